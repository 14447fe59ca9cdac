// ctx.h — the library's context object and error helpers (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gl.h"
#include "hostsync.h"
#include "kernels.h"

namespace whir_gpu {

std::string& last_error();  // per thread

// Serialises the contexts' grow-only buffer (re)allocations across threads:
// serving-loop workers each own a context, and a worker that grows a buffer
// (hipFree + hipMalloc, or the pinned-host forms) does so while other workers
// keep launching.  Growth is rare after warm-up, so one process-wide lock
// costs nothing measurable and keeps those runtime calls from ever running
// concurrently with each other.
std::mutex& alloc_mutex();

inline int fail(const char* what, hipError_t e) {
    last_error() = std::string(what) + ": " + hipGetErrorString(e);
    return -1;
}

inline int bad_arg(const char* what) {
    last_error() = what;
    return 3;
}

inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline uint32_t log2u(uint64_t x) {
    uint32_t k = 0;
    while ((1ull << k) < x) ++k;
    return k;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// The host waits for a stream by polling a word of pinned, device-mapped host
// memory that a one-lane kernel (signal_launch) sets after the stream's
// earlier work: about 6 us from a short kernel's end to the host seeing it,
// against about 11 us through hipStreamSynchronize.  h / d: host and device
// addresses of the word; seq: the last number enqueued.
struct StreamSignal {
    uint64_t* h = nullptr;
    uint64_t* d = nullptr;
    uint64_t seq = 0;
    WaitHub* hub = nullptr;  // when set, waits sleep on the hub instead of spinning
};

int signal_init(StreamSignal& s);  // allocates the word once
void signal_free(StreamSignal& s);
// Enqueues the signal on st and waits for it: polls, then yields, and after
// 2 ms falls back to hipStreamSynchronize (which also reports stream errors).
int signal_wait(StreamSignal& s, hipStream_t st);

// One proof-of-work search and its current window.
constexpr uint32_t POW_MAX_LG_LANES = 22;
struct PowJob {
    uint8_t ch[32];
    uint64_t threshold = 0, base = 0, window = 0;
    uint32_t blocks = 0, per_thread = 0;
    uint32_t hash = HASH_BLAKE3;  // engine of the nonce hash; one per launch
};
void pow_job_init(PowJob& j, const uint8_t ch[32], uint64_t threshold, uint32_t extra_bits = 1);
// One launch over jobs[0..n) (n <= POW_BATCH_MAX, their blocks <= cap entries
// of the mapped block-minimum buffer hmin / hmin_d).  found[k], nonce[k] get
// job k's answer; a job without a hit moves to its next window.  slots holds
// two sets of POW_BATCH_MAX device early-exit slots (~0 at start); `set`
// alternates between them.  Waits for the stream (through sig when given).
int pow_run(hipStream_t st, unsigned long long* slots, int& set, const uint64_t* hmin, uint64_t* hmin_d, size_t cap,
            PowJob* const* jobs, int n, bool* found, uint64_t* nonce, StreamSignal* sig = nullptr);

}  // namespace whir_gpu

#define HIPCHK(expr)                                              \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return ::whir_gpu::fail(#expr, e_); \
    } while (0)

struct whir_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    whir_gpu::DevBuf scratch, nttbuf, partials, hin, hout, hnodes, hidx, hcomp, tbuf;
    // prove loop: PoW words, term descriptors and tables, opened rows, vector, covector
    whir_gpu::DevBuf powbuf, tdesc, ttab;
    int pow_slot = 0;  // early-exit slot set of the next PoW launch
    // hash engine of the commits, trees and PoW searches run on this context
    // (HASH_BLAKE3 or HASH_SHA2; whir_ctx_set_hash)
    uint32_t hash = whir_gpu::HASH_BLAKE3;
    // when set, LDE passes use this caller-owned scratch instead of `scratch`
    void* ext_scratch = nullptr;
    size_t ext_scratch_bytes = 0;
    std::map<std::pair<uint32_t, bool>, whir_gpu::Twiddles> tw;
    uint64_t* t1024[2] = {nullptr, nullptr};
    // stage timing (HIP events on `stream`)
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double acc_ms[4] = {0, 0, 0, 0};
    uint64_t cnt[4] = {0, 0, 0, 0};

    int mark(int k) {
        if (!timing) return 0;
        if (!ev[k]) HIPCHK(hipEventCreate(&ev[k]));
        HIPCHK(hipEventRecord(ev[k], stream));
        return 0;
    }
    // adds the time between marks a and b to stage st (waits for mark b)
    int collect(int a, int b, int st) {
        if (!timing) return 0;
        HIPCHK(hipEventSynchronize(ev[b]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[a], ev[b]));
        acc_ms[st] += ms;
        cnt[st] += 1;
        return 0;
    }

    int ensure(whir_gpu::DevBuf& b, size_t bytes) {
        if (b.bytes >= bytes && b.p) return 0;
        std::lock_guard<std::mutex> lk(whir_gpu::alloc_mutex());
        if (b.p) {
            HIPCHK(hipStreamSynchronize(stream));
            HIPCHK(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        HIPCHK(hipMalloc(&b.p, want));
        b.bytes = want;
        return 0;
    }

    // partial-sum buffer of the two-stage reductions
    int ensure_partials() { return ensure(partials, 6 * whir_gpu::SUMCHECK_MAX_BLOCKS * 8); }

    // Block partials written by the kernels straight into pinned, device-mapped
    // host memory (fine-grained: the stores cross PCIe as the blocks finish),
    // so a reduction whose result the host needs costs no second-stage kernel
    // and no copy: the host sums the g block totals after the stream sync.
    // Waits for the stream, then sums the nv Ext3 totals of g blocks left in
    // hpart by a launcher called with device_sum = false (lazy 128-bit sums:
    // g <= 2048 canonical limbs stay below 2^75).
    int host_sums(unsigned g, int nv, gl::e3* out) {
        if (int rc = wait()) return rc;
        for (int v = 0; v < nv; ++v) {
            unsigned __int128 acc[3] = {0, 0, 0};
            for (unsigned b = 0; b < g; ++b) {
                const uint64_t* q = hpart + 3 * ((uint64_t)nv * b + v);
                acc[0] += q[0];
                acc[1] += q[1];
                acc[2] += q[2];
            }
            uint64_t c[3];
            for (int k = 0; k < 3; ++k) c[k] = gl::reduce128((uint64_t)acc[k], (uint64_t)(acc[k] >> 64));
            out[v] = {c[0], c[1], c[2]};
        }
        return 0;
    }
    // u64 entries: the sumcheck partials (6 per block) and one PoW minimum per
    // block of the widest search window (2^22 lanes / 256)
    static constexpr size_t HPART_U64 = 16384;
    static_assert((size_t{1} << (whir_gpu::POW_MAX_LG_LANES - 8)) <= HPART_U64, "one PoW minimum per block");

    // A serving loop may take over the nonce searches (to batch them across
    // its requests): the hook returns the nonce for (challenge, threshold).
    int (*pow_hook)(void* arg, const uint8_t* ch, uint64_t threshold, uint64_t* nonce) = nullptr;
    void* pow_hook_arg = nullptr;
    // ... and the encodes of its requests' initial commits (BLAKE3, 16-column
    // rows, 2^10 <= m <= 2^18): the hook encodes `seg` (leaves included) with
    // the encodes of other requests, after the work enqueued so far on
    // `stream`, and returns once `stream` waits for the result.
    int (*encode_hook)(void* arg, const whir_gpu::LdeMultiSeg& seg, const uint64_t* t1024, hipStream_t st) = nullptr;
    void* encode_hook_arg = nullptr;
    // waits for all work enqueued on `stream` (through the mapped signal word)
    whir_gpu::StreamSignal sig;
    // Failure injection for tests of the callers' error paths: when nonzero,
    // the fail_at_wait-th wait() from now returns an error without waiting,
    // so the work enqueued since the previous wait is still in flight when
    // the caller unwinds.  Host-side only; no kernel is affected.
    uint32_t fail_at_wait = 0;
    int wait() {
        if (fail_at_wait && --fail_at_wait == 0) {
            whir_gpu::last_error() = "injected failure at a device wait";
            return -1;
        }
        if (int rc = whir_gpu::signal_init(sig)) return rc;
        return whir_gpu::signal_wait(sig, stream);
    }
    uint64_t* hpart = nullptr;   // host address
    uint64_t* hpart_d = nullptr;  // device address of the same memory

    // Pinned staging of the term descriptors: the copy out of it runs
    // asynchronously; before the host rewrites it, it waits for the event
    // recorded after the previous copy.
    void* hdesc = nullptr;
    size_t hdesc_bytes = 0;
    hipEvent_t hdesc_ev = nullptr;
    int stage_desc(const std::vector<uint64_t>& v, void* d_dst) {
        const size_t bytes = v.size() * 8;
        if (!bytes) return 0;
        if (hdesc_ev) HIPCHK(hipEventSynchronize(hdesc_ev));
        if (hdesc_bytes < bytes) {
            std::lock_guard<std::mutex> lk(whir_gpu::alloc_mutex());
            if (hdesc) HIPCHK(hipHostFree(hdesc));
            hdesc = nullptr;
            hdesc_bytes = 0;
            size_t want = std::max<size_t>(bytes, 64 << 10);
            HIPCHK(hipHostMalloc(&hdesc, want, hipHostMallocDefault));
            hdesc_bytes = want;
        }
        std::memcpy(hdesc, v.data(), bytes);
        HIPCHK(hipMemcpyAsync(d_dst, hdesc, bytes, hipMemcpyHostToDevice, stream));
        if (!hdesc_ev) HIPCHK(hipEventCreateWithFlags(&hdesc_ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(hdesc_ev, stream));
        return 0;
    }

    // Two pinned chunks for copies to and from pageable host memory
    // (whir_gpu::copy_h2d / copy_d2h); each chunk's last DMA is recorded in
    // its event before the host reuses the chunk.
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};

    // Grow-only mapped staging for openings: the gather kernels read the
    // indices from it and write the rows and hashes into it.
    void* hio = nullptr;
    void* hio_d = nullptr;
    size_t hio_bytes = 0;
    int ensure_hio(size_t bytes) {
        if (hio && hio_bytes >= bytes) return 0;
        std::lock_guard<std::mutex> lk(whir_gpu::alloc_mutex());
        if (hio) {
            HIPCHK(hipStreamSynchronize(stream));
            HIPCHK(hipHostFree(hio));
            hio = hio_d = nullptr;
            hio_bytes = 0;
        }
        size_t want = std::max<size_t>(bytes, 64 << 10);
        HIPCHK(hipHostMalloc(&hio, want, hipHostMallocMapped | hipHostMallocCoherent));
        hipError_t e = hipHostGetDevicePointer(&hio_d, hio, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(hio);
            hio = nullptr;
            return whir_gpu::fail("hipHostGetDevicePointer", e);
        }
        hio_bytes = want;
        return 0;
    }
    int ensure_hpart() {
        if (hpart) return 0;
        std::lock_guard<std::mutex> lk(whir_gpu::alloc_mutex());
        void* h = nullptr;
        HIPCHK(hipHostMalloc(&h, HPART_U64 * 8, hipHostMallocMapped | hipHostMallocCoherent));
        void* d = nullptr;
        hipError_t e = hipHostGetDevicePointer(&d, h, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(h);
            return whir_gpu::fail("hipHostGetDevicePointer", e);
        }
        hpart = (uint64_t*)h;
        hpart_d = (uint64_t*)d;
        return 0;
    }

    int twiddles(uint32_t log_n, bool inverse, whir_gpu::Twiddles& out) {
        auto key = std::make_pair(log_n, inverse);
        auto it = tw.find(key);
        if (it != tw.end()) {
            out = it->second;
            return 0;
        }
        int inv = inverse ? 1 : 0;
        std::lock_guard<std::mutex> lk(whir_gpu::alloc_mutex());
        if (!t1024[inv]) {
            uint64_t w = gl::root_of_unity(10);
            if (inverse) w = gl::inv(w);
            std::vector<uint64_t> h(1024);
            uint64_t acc = 1;
            for (int i = 0; i < 1024; ++i) {
                h[i] = acc;
                acc = gl::mul(acc, w);
            }
            HIPCHK(hipMalloc(&t1024[inv], 1024 * 8));
            HIPCHK(hipMemcpy(t1024[inv], h.data(), 1024 * 8, hipMemcpyHostToDevice));
        }
        uint64_t n = 1ull << log_n;
        uint64_t w = gl::root_of_unity(log_n);
        if (inverse) w = gl::inv(w);
        std::vector<uint64_t> lo(2048), hi(std::max<uint64_t>(1, n / 2048));
        uint64_t acc = 1;
        for (int i = 0; i < 2048; ++i) {
            lo[i] = acc;
            acc = gl::mul(acc, w);
        }
        uint64_t w2048 = gl::pow(w, 2048);
        acc = 1;
        for (size_t i = 0; i < hi.size(); ++i) {
            hi[i] = acc;
            acc = gl::mul(acc, w2048);
        }
        whir_gpu::Twiddles t;
        t.log_n = log_n;
        HIPCHK(hipMalloc(&t.lo, lo.size() * 8));
        HIPCHK(hipMalloc(&t.hi, hi.size() * 8));
        HIPCHK(hipMemcpy(t.lo, lo.data(), lo.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t.hi, hi.data(), hi.size() * 8, hipMemcpyHostToDevice));
        t.t1024 = t1024[inv];
        tw[key] = t;
        out = t;
        return 0;
    }
};

namespace whir_gpu {

// Host <-> device copies on ctx's stream for the host-pointer entry points.
// Pageable host memory goes through the context's pinned staging chunks (DMA
// of one chunk overlapping the host copy of the other); pinned memory and
// small copies take one DMA.  copy_h2d returns once the source may be
// reused (the last DMAs may still run, in stream order); copy_d2h returns
// once the destination is written.
int copy_h2d(whir_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
int copy_d2h(whir_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
// Staged D2H of `bytes` (a multiple of `unit`) where consume(off, p, len)
// moves bytes [off, off + len) of the source out of the pinned chunk p; it is
// called for pieces of whole units, several at once on the copy threads.
using D2hConsume = std::function<void(size_t off, const uint8_t* p, size_t len)>;
int copy_d2h_each(whir_ctx* ctx, const void* src_dev, size_t bytes, size_t unit, const D2hConsume& consume);

// Autotuned LDE splits (log2 M2) per shape; 0 = not tuned.
uint32_t tuned_split(const LdeShape& s);
void set_tuned_split(const LdeShape& s, uint32_t lg_m2);
// Flat records of 9 words (8 key words, then log2 M2); returns the word count.
size_t tuned_entries(uint64_t* out, size_t cap);

// One LDE on ctx's stream (tuned split unless s.lg_m2 is set).
int lde_device(whir_ctx* ctx, const LdeShape& s, const uint64_t* in, uint64_t* out, uint8_t* leaves = nullptr,
               bool* fused = nullptr);

// RS encode + leaves + Merkle tree of num_polys vectors already on the device;
// the root is copied to out_root (host) when non-null.  With tree false only
// the encode and the leaves (in d_nodes' first cw nodes) run.
int commit_device(whir_ctx* ctx, const uint64_t* d_coeffs, size_t num_polys, size_t poly_size, size_t cw,
                  size_t depth, uint64_t* d_matrix, uint8_t* d_nodes, uint8_t* out_root, int ext, bool tree = true);

}  // namespace whir_gpu
