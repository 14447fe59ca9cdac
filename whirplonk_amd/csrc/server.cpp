// server.cpp — a serving loop for commit+prove requests (C ABI: whir_server_*).
//
// Requests of different sizes share one parameter set.  A request is admitted
// when a slot is idle and its device footprint (the prover's workspace plus
// its coefficients) can be carved from a fixed HBM arena of the configured
// budget; otherwise it waits, and smaller requests behind it may pass it
// until it has been passed too often.  Each slot is a worker thread with its
// own context and HIP stream, so admitted requests run concurrently on the
// device: their kernels, copies and host transcript work overlap, and the
// codewords, trees and vectors of each request stay resident in its arena
// block across all WHIR rounds.  Same-stage work is batched across requests
// where it has no data dependence on a request's device state: every
// proof-of-work search of the running requests goes through one batcher that
// launches the searches pending at that moment as one kernel, and the encodes
// of their initial commits through another that runs the encodes pending at
// that moment as one level-synchronous launch pair.  Results are
// byte-identical to the plain single-request path (the same prover code runs
// on the same inputs; a search's nonce does not depend on its batch).
#include "../../include/whir_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ctx.h"

using namespace whir_gpu;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over one device region, with coalescing on free.
class Arena {
public:
    void init(uint8_t* base, size_t bytes) {
        base_ = base;
        free_.clear();
        free_[0] = bytes;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~size_t{255};
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                size_t rest = it->second - bytes, at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~size_t{255};
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint8_t* at(size_t off) const { return base_ + off; }

private:
    uint8_t* base_ = nullptr;
    std::map<size_t, size_t> free_;
};

struct Request {
    uint64_t id = 0;
    uint64_t size = 0;
    size_t num_forms = 0;
    std::vector<uint64_t> coeffs, points, evals;
    const uint64_t* borrowed = nullptr;  // caller-owned coefficients (no copy)
    bool have_evals = false;
    size_t footprint = 0, ws_bytes = 0, coeff_bytes = 0, arena_off = 0;
    size_t skipped = 0;
    uint32_t fail_at_wait = 0;  // injected failure (whir_server_inject_failure), 0 = none
    bool done = false;
    int status = 0;
    std::string error;
    std::vector<uint8_t> narg, hints;
    Clock::time_point t_submit, t_admit, t_done;
};

// Proof-of-work searches of all running requests, batched: a worker that
// reaches a search queues it; if one of the LANES launch paths is free it
// becomes the leader and launches everything queued (up to POW_BATCH_MAX
// jobs) as one kernel on that path's stream, while searches that arrive
// meanwhile queue for the next launch.  No request waits for a batch to fill: a lone search launches
// at once, and under load the batches form by themselves.  A search is a pure
// function of (challenge, threshold), so its nonce is the one it gets alone.
#ifndef SERVE_POW_EXTRA_BITS
#define SERVE_POW_EXTRA_BITS 1  // search windows of 1-2x the expected work (pow_job_init)
#endif

class PowBatcher {
public:
    static constexpr int LANES = 2;  // launches in flight at once

    int init(int device, uint32_t hash, WaitHub* hub) {
        HIPCHK(hipSetDevice(device));
        hash_ = hash;
        cap_ = (size_t)POW_BATCH_MAX << (POW_MAX_LG_LANES - 8);  // one minimum per block
        for (Lane& l : lanes_) {
            HIPCHK(hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking));
            HIPCHK(hipMalloc((void**)&l.slots, 2 * POW_BATCH_MAX * 8));
            fill_u64_launch((uint64_t*)l.slots, 2 * POW_BATCH_MAX, ~0ull, l.st);
            void* h = nullptr;
            HIPCHK(hipHostMalloc(&h, cap_ * 8, hipHostMallocMapped | hipHostMallocCoherent));
            l.hmin = (uint64_t*)h;
            HIPCHK(hipHostGetDevicePointer((void**)&l.hmin_d, h, 0));
            if (int rc = signal_init(l.sig)) return rc;
            l.sig.hub = hub;
            HIPCHK(hipStreamSynchronize(l.st));
        }
        return 0;
    }
    ~PowBatcher() {
        for (Lane& l : lanes_) {
            if (l.st) (void)hipStreamSynchronize(l.st);
            if (l.slots) (void)hipFree(l.slots);
            if (l.hmin) (void)hipHostFree(l.hmin);
            signal_free(l.sig);
            if (l.st) (void)hipStreamDestroy(l.st);
        }
    }
    static int hook(void* self, const uint8_t* ch, uint64_t threshold, uint64_t* nonce) {
        return static_cast<PowBatcher*>(self)->search(ch, threshold, nonce);
    }
    void stats(uint64_t* launches, uint64_t* jobs) {
        std::lock_guard<std::mutex> lk(mu_);
        *launches = launches_;
        *jobs = jobs_;
    }

private:
    struct Waiter {
        PowJob job;
        bool done = false;
        int rc = 0;
        std::string error;  // the leader's message, re-raised in this thread
        uint64_t nonce = 0;
    };
    // one launch path: its stream, early-exit slot sets and block minima
    uint32_t hash_ = HASH_BLAKE3;  // the engine of every search (the server's hash_id)
    struct Lane {
        hipStream_t st = nullptr;
        unsigned long long* slots = nullptr;
        int set = 0;
        uint64_t* hmin = nullptr;
        uint64_t* hmin_d = nullptr;
        StreamSignal sig;  // the leader waits for its launch through this word
        bool busy = false;
    };

    int search(const uint8_t* ch, uint64_t threshold, uint64_t* nonce) {
        Waiter w;
        pow_job_init(w.job, ch, threshold, SERVE_POW_EXTRA_BITS);
        w.job.hash = hash_;
        std::unique_lock<std::mutex> lk(mu_);
        queue_.push_back(&w);
        while (!w.done) {
            Lane* lane = nullptr;
            for (Lane& l : lanes_)
                if (!l.busy) lane = &l;
            if (!lane || queue_.empty()) {
                cv_.wait(lk);
                continue;
            }
            // lead one launch over the queue
            lane->busy = true;
            std::vector<Waiter*> batch;
            size_t blocks = 0;
            while (!queue_.empty() && batch.size() < POW_BATCH_MAX && blocks + queue_.front()->job.blocks <= cap_) {
                blocks += queue_.front()->job.blocks;
                batch.push_back(queue_.front());
                queue_.pop_front();
            }
            lk.unlock();
            std::vector<PowJob*> jobs;
            for (Waiter* x : batch) jobs.push_back(&x->job);
            bool found[POW_BATCH_MAX];
            uint64_t got[POW_BATCH_MAX];
            int rc = pow_run(lane->st, lane->slots, lane->set, lane->hmin, lane->hmin_d, cap_, jobs.data(),
                             (int)jobs.size(), found, got, &lane->sig);
            std::string err = rc ? last_error() : std::string();
            lk.lock();
            lane->busy = false;
            ++launches_;
            jobs_ += batch.size();
            for (size_t k = 0; k < batch.size(); ++k) {
                Waiter* x = batch[k];
                if (rc) {
                    x->rc = rc;
                    x->error = err;
                    x->done = true;
                } else if (found[k]) {
                    x->nonce = got[k];
                    x->done = true;
                } else {
                    queue_.push_front(x);  // its next window goes first
                }
            }
            cv_.notify_all();
        }
        if (w.rc) {
            last_error() = w.error;  // per-thread error text of this worker
            return w.rc;
        }
        *nonce = w.nonce;
        return 0;
    }

    Lane lanes_[LANES];
    size_t cap_ = 0;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Waiter*> queue_;
    uint64_t launches_ = 0, jobs_ = 0;
};

// The encodes of the running requests' initial commits, batched: a worker
// that reaches its commit waits (on the host) for its stream's earlier work
// (the coefficient upload), queues the encode (one LdeMultiSeg: its
// coefficients, codeword, leaf level and scratch in its arena block) and, if
// the launch path is free, leads: it runs every queued encode (up to
// MAX_SEGS) as one level-synchronous launch pair on the batcher's stream and
// waits for it, while encodes that arrive meanwhile queue for the next
// launch.  Streams never wait on one another on the device (a stream waiting
// on another's event would hold its hardware queue, which other slots'
// streams share).  As with the searches, no request waits for a batch to
// fill, and an encode's output does not depend on its batch.
class EncodeBatcher {
public:
    static constexpr uint32_t MAX_SEGS = 16;

    int init(int device, WaitHub* hub) {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        HIPCHK(hipMalloc((void**)&table_, MAX_SEGS * sizeof(LdeMultiSeg)));
        HIPCHK(hipHostMalloc((void**)&staging_, MAX_SEGS * sizeof(LdeMultiSeg), hipHostMallocDefault));
        hub_ = hub;
        if (int rc = signal_init(sig_)) return rc;
        sig_.hub = hub;
        return 0;
    }
    ~EncodeBatcher() {
        if (st_) (void)hipStreamSynchronize(st_);
        if (table_) (void)hipFree(table_);
        if (staging_) (void)hipHostFree(staging_);
        signal_free(sig_);
        if (st_) (void)hipStreamDestroy(st_);
    }
    static int hook(void* self, const LdeMultiSeg& seg, const uint64_t* t1024, hipStream_t st) {
        return static_cast<EncodeBatcher*>(self)->encode(seg, t1024, st);
    }
    void stats(uint64_t* launches, uint64_t* jobs) {
        std::lock_guard<std::mutex> lk(mu_);
        *launches = launches_;
        *jobs = jobs_;
    }

private:
    struct Waiter {
        LdeMultiSeg seg;
        const uint64_t* t1024;
        bool done = false;
        int rc = 0;
        std::string error;
    };

    // one completion word per worker thread, for its own stream
    int thread_signal(StreamSignal** out) {
        struct Holder {
            StreamSignal s;
            ~Holder() { signal_free(s); }
        };
        thread_local Holder h;
        if (int rc = signal_init(h.s)) return rc;
        h.s.hub = hub_;
        *out = &h.s;
        return 0;
    }

    // runs on the leader's thread with the path held: launch and wait
    int lead(std::vector<Waiter*>& batch) {
        std::stable_sort(batch.begin(), batch.end(),
                         [](const Waiter* a, const Waiter* b) { return a->seg.lg_m > b->seg.lg_m; });
        std::vector<LdeMultiSeg> segs;
        for (Waiter* w : batch) segs.push_back(w->seg);
        LdeMultiPlan plan;
        if (lde_multi_prepare(segs, plan)) return bad_arg("encode batch rejected by the level-synchronous plan");
        std::memcpy(staging_, segs.data(), segs.size() * sizeof(LdeMultiSeg));  // the last launch is done
        HIPCHK(hipMemcpyAsync(table_, staging_, segs.size() * sizeof(LdeMultiSeg), hipMemcpyHostToDevice, st_));
        lde_multi_launch(table_, (uint32_t)segs.size(), plan, batch[0]->t1024, st_);
        HIPCHK(hipGetLastError());
        return signal_wait(sig_, st_);
    }

    int encode(const LdeMultiSeg& seg, const uint64_t* t1024, hipStream_t st) {
        StreamSignal* own = nullptr;
        if (int rc = thread_signal(&own)) return rc;
        if (int rc = signal_wait(*own, st)) return rc;  // the coefficients are on the device
        Waiter w;
        w.seg = seg;
        w.t1024 = t1024;
        std::unique_lock<std::mutex> lk(mu_);
        queue_.push_back(&w);
        while (!w.done) {
            if (busy_ || queue_.empty()) {
                cv_.wait(lk);
                continue;
            }
            busy_ = true;
            std::vector<Waiter*> batch;
            while (!queue_.empty() && batch.size() < MAX_SEGS) {
                batch.push_back(queue_.front());
                queue_.pop_front();
            }
            lk.unlock();
            int rc = lead(batch);
            std::string err = rc ? last_error() : std::string();
            if (rc && hipStreamSynchronize(st_) != hipSuccess && err.empty()) err = "encode batch stream failed";
            lk.lock();
            busy_ = false;
            ++launches_;
            jobs_ += batch.size();
            for (Waiter* x : batch) {
                x->rc = rc;
                x->error = err;
                x->done = true;
            }
            cv_.notify_all();
        }
        if (w.rc) {
            last_error() = w.error;
            return w.rc;
        }
        return 0;
    }

    hipStream_t st_ = nullptr;
    LdeMultiSeg* table_ = nullptr;    // device copy of the launch's segments
    LdeMultiSeg* staging_ = nullptr;  // pinned source of that copy
    StreamSignal sig_;
    WaitHub* hub_ = nullptr;
    bool busy_ = false;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Waiter*> queue_;
    uint64_t launches_ = 0, jobs_ = 0;
};

struct Slot {
    whir_ctx* ctx = nullptr;
    uint8_t* pinned = nullptr;
    size_t pinned_bytes = 0;
    std::thread th;
};

}  // namespace

struct whir_server {
    int device = 0;
    whir_params params{};
    uint8_t protocol_id[64], session_id[32];
    size_t budget = 0, in_use = 0, peak = 0;
    uint8_t* region = nullptr;
    Arena arena;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::map<uint64_t, uint32_t> inject;  // request id -> device wait that fails
    // the slots' waits sleep on this hub (its poller outlives the batchers and
    // the workers, which are declared or joined before it goes)
    WaitHub hub;
    PowBatcher pow;
    EncodeBatcher enc;
    uint64_t next_id = 1;
    size_t completed = 0, running = 0;
    bool stop = false;

    size_t footprint(uint64_t size, size_t* ws, size_t* cb) const {
        *ws = whir_prover_workspace_bytes(&params, size);
        *cb = ((size_t)params.batch_size * size * 8 + 255) & ~size_t{255};
        return *ws ? *ws + *cb : 0;
    }

    // the first admissible request (and its arena block), or null
    Request* pick() {
        const size_t max_skips = 4 * slots.size();
        for (size_t i = 0; i < queue.size(); ++i) {
            Request* r = queue[i];
            size_t off;
            if (arena.alloc(r->footprint, &off)) {
                r->arena_off = off;
                for (size_t j = 0; j < i; ++j) queue[j]->skipped++;
                queue.erase(queue.begin() + (long)i);
                return r;
            }
            if (r->skipped >= max_skips) return nullptr;  // keep its place: no more passing
        }
        return nullptr;
    }

    void run(Slot& s, Request& r);
    void worker(Slot* s);
};

void whir_server::run(Slot& s, Request& r) {
    // Every exit waits for the slot's stream first: a request that fails part
    // way may still have copies and kernels queued there that read or write
    // its arena block and read the slot's pinned staging buffer (or the
    // caller's borrowed buffer).  Only then may the worker release the block,
    // the caller free a borrowed buffer, and the next request reuse the slot.
    struct Drain {
        whir_ctx* ctx;
        ~Drain() {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->fail_at_wait = 0;
        }
    } drain{s.ctx};
    s.ctx->fail_at_wait = r.fail_at_wait;
    auto fail_with = [&r](int rc) {
        r.status = rc;
        r.error = last_error();
    };
    uint8_t* blk = arena.at(r.arena_off);
    uint64_t* d_coeffs = (uint64_t*)(blk + r.ws_bytes);
    // coefficients: a borrowed buffer is copied by DMA straight from the
    // caller's memory (pinned memory is fastest); an owned copy goes through
    // this slot's pinned staging buffer
    const size_t cbytes = (size_t)params.batch_size * r.size * 8;
    const void* src = r.borrowed;
    if (!src) {
        if (s.pinned_bytes < cbytes) {
            if (s.pinned) (void)hipHostFree(s.pinned);
            s.pinned = nullptr;
            s.pinned_bytes = 0;
            hipError_t e = hipHostMalloc((void**)&s.pinned, cbytes, hipHostMallocDefault);
            if (e != hipSuccess) return fail_with(fail("hipHostMalloc", e));
            s.pinned_bytes = cbytes;
        }
        std::memcpy(s.pinned, r.coeffs.data(), cbytes);
        src = s.pinned;
    }
    hipError_t e = hipMemcpyAsync(d_coeffs, src, cbytes, hipMemcpyHostToDevice, s.ctx->stream);
    if (e != hipSuccess) return fail_with(fail("hipMemcpyAsync", e));
    const uint32_t nv = log2u(r.size);
    const size_t nvec = params.batch_size;
    int rc;
    if (!r.have_evals && r.num_forms) {
        // every constraint of a vector in one pass over it
        r.evals.assign(3 * r.num_forms * nvec, 0);
        std::vector<uint64_t> ev(3 * r.num_forms);
        for (size_t v = 0; v < nvec; ++v) {
            if ((rc = whir_dev_mle_evaluate_many(s.ctx, d_coeffs + v * r.size, 1, nv, r.points.data(), r.num_forms,
                                                 ev.data())))
                return fail_with(rc);
            for (size_t f = 0; f < r.num_forms; ++f) std::memcpy(r.evals.data() + 3 * (f * nvec + v), &ev[3 * f], 24);
        }
    }
    whir_prover* P = nullptr;
    if ((rc = whir_prover_create_in(s.ctx, &params, r.size, protocol_id, session_id, nullptr, 0, blk, r.ws_bytes,
                                    &P)))
        return fail_with(rc);
    std::unique_ptr<whir_prover, int (*)(whir_prover*)> guard(P, whir_prover_destroy);
    if ((rc = whir_prover_commit(P, d_coeffs, nvec))) return fail_with(rc);
    if ((rc = whir_prover_prove(P, r.num_forms, r.num_forms ? r.points.data() : nullptr,
                                r.num_forms ? r.evals.data() : nullptr, nullptr, nullptr)))
        return fail_with(rc);
    size_t na = 0, nh = 0;
    whir_prover_proof(P, nullptr, 0, &na, nullptr, 0, &nh);
    r.narg.resize(na);
    r.hints.resize(nh);
    if ((rc = whir_prover_proof(P, r.narg.data(), na, &na, r.hints.data(), nh, &nh))) return fail_with(rc);
    r.status = 0;
}

void whir_server::worker(Slot* s) {
    (void)hipSetDevice(device);
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || (r = pick()) != nullptr; });
            if (!r) return;  // stopping with nothing admissible
            auto inj = inject.find(r->id);
            if (inj != inject.end()) {
                r->fail_at_wait = inj->second;
                inject.erase(inj);
            }
            r->t_admit = Clock::now();
            in_use += r->footprint;
            peak = std::max(peak, in_use);
            ++running;
        }
        run(*s, *r);
        r->coeffs = std::vector<uint64_t>();  // inputs are no longer needed
        {
            std::lock_guard<std::mutex> lk(mu);
            arena.release(r->arena_off, r->footprint);
            in_use -= r->footprint;
            --running;
            r->t_done = Clock::now();
            r->done = true;
            ++completed;
        }
        cv_done.notify_all();
        cv_work.notify_all();  // freed memory may admit a waiting request
    }
}

extern "C" {

int whir_host_malloc(size_t bytes, void** out) {
    if (!out) return bad_arg("null output");
    HIPCHK(hipHostMalloc(out, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    return 0;
}

int whir_host_free(void* p) {
    if (p) HIPCHK(hipHostFree(p));
    return 0;
}

int whir_server_create(int device, const whir_params* params, uint32_t num_slots, size_t hbm_budget_bytes,
                       whir_server** out) {
    if (!params || !out || num_slots == 0 || hbm_budget_bytes == 0) return bad_arg("bad server arguments");
    if (params->hash_id != WHIR_HASH_BLAKE3 && params->hash_id != WHIR_HASH_SHA2) return bad_arg("unknown hash engine");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<whir_server> S(new whir_server());
    S->device = device;
    S->params = *params;
    int rc = whir_domain_separator_cli(params, nullptr, S->protocol_id, S->session_id);
    if (rc) return rc;
    S->budget = (hbm_budget_bytes + 255) & ~size_t{255};
    HIPCHK(hipMalloc((void**)&S->region, S->budget));
    S->arena.init(S->region, S->budget);
    if ((rc = S->pow.init(device, params->hash_id == WHIR_HASH_SHA2 ? HASH_SHA2 : HASH_BLAKE3, &S->hub)) ||
        (rc = S->enc.init(device, &S->hub))) {
        (void)hipFree(S->region);
        return rc;
    }
    for (uint32_t i = 0; i < num_slots; ++i) {
        auto sl = std::make_unique<Slot>();
        if ((rc = whir_ctx_create(device, &sl->ctx))) {
            for (auto& x : S->slots) whir_ctx_destroy(x->ctx);
            (void)hipFree(S->region);
            return rc;
        }
        sl->ctx->pow_hook = &PowBatcher::hook;
        sl->ctx->pow_hook_arg = &S->pow;
        sl->ctx->encode_hook = &EncodeBatcher::hook;
        sl->ctx->encode_hook_arg = &S->enc;
        sl->ctx->sig.hub = &S->hub;
        S->slots.push_back(std::move(sl));
    }
    for (auto& sl : S->slots) sl->th = std::thread(&whir_server::worker, S.get(), sl.get());
    *out = S.release();
    return 0;
}

int whir_server_destroy(whir_server* S) {
    if (!S) return 0;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        // let everything queued finish first
        S->cv_done.wait(lk, [&] { return S->queue.empty() && S->running == 0; });
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& sl : S->slots) {
        if (sl->th.joinable()) sl->th.join();
        whir_ctx_destroy(sl->ctx);
        if (sl->pinned) (void)hipHostFree(sl->pinned);
    }
    (void)hipFree(S->region);
    delete S;
    return 0;
}

size_t whir_server_request_bytes(whir_server* S, uint64_t size) {
    size_t ws, cb;
    return S ? S->footprint(size, &ws, &cb) : 0;
}

int whir_server_submit(whir_server* S, uint64_t size, const uint64_t* coeffs, int borrow, size_t num_forms,
                       const uint64_t* points, const uint64_t* evaluations, uint64_t* out_id) {
    if (!S || !coeffs || !out_id || (num_forms && !points)) return bad_arg("null argument");
    auto r = std::make_unique<Request>();
    r->size = size;
    r->footprint = S->footprint(size, &r->ws_bytes, &r->coeff_bytes);
    if (!r->footprint) return bad_arg("invalid request size");
    if (r->footprint > S->budget) return bad_arg("request larger than the HBM budget");
    const size_t nvec = S->params.batch_size;
    const uint32_t nv = log2u(size);
    if (borrow)
        r->borrowed = coeffs;
    else
        r->coeffs.assign(coeffs, coeffs + nvec * size);
    r->num_forms = num_forms;
    if (num_forms) r->points.assign(points, points + 3 * nv * num_forms);
    if (evaluations) {
        r->evals.assign(evaluations, evaluations + 3 * num_forms * nvec);
        r->have_evals = true;
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        *out_id = r->id;
        S->queue.push_back(r.get());
        S->reqs[r->id] = std::move(r);
    }
    S->cv_work.notify_all();
    return 0;
}

int whir_server_wait(whir_server* S, uint64_t id, uint8_t* narg, size_t narg_cap, size_t* narg_len, uint8_t* hints,
                     size_t hints_cap, size_t* hints_len, double* out_ms) {
    if (!S) return bad_arg("null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return bad_arg("unknown request id");
    Request* r = it->second.get();
    S->cv_done.wait(lk, [&] { return r->done; });
    if (narg_len) *narg_len = r->narg.size();
    if (hints_len) *hints_len = r->hints.size();
    if (out_ms) {
        out_ms[0] = ms_between(r->t_submit, r->t_admit);
        out_ms[1] = ms_between(r->t_admit, r->t_done);
        out_ms[2] = ms_between(r->t_submit, r->t_done);
    }
    if (r->status) {
        last_error() = r->error;
        int st = r->status;
        S->reqs.erase(it);
        return st;
    }
    if ((narg && r->narg.size() > narg_cap) || (hints && r->hints.size() > hints_cap))
        return bad_arg("proof buffer too small");  // the result is kept for another call
    if (narg) std::memcpy(narg, r->narg.data(), r->narg.size());
    if (hints) std::memcpy(hints, r->hints.data(), r->hints.size());
    if (narg || hints) S->reqs.erase(it);
    return 0;
}

int whir_server_inject_failure(whir_server* S, uint64_t id, uint32_t at_wait) {
    if (!S) return bad_arg("null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (at_wait)
        S->inject[id] = at_wait;
    else
        S->inject.erase(id);
    return 0;
}

int whir_server_stats(whir_server* S, uint64_t* out) {
    if (!S || !out) return bad_arg("null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out[0] = S->completed;
    out[1] = S->running;
    out[2] = S->queue.size();
    out[3] = S->in_use;
    out[4] = S->peak;
    out[5] = S->budget;
    out[6] = S->slots.size();
    S->pow.stats(&out[7], &out[8]);
    S->enc.stats(&out[9], &out[10]);
    return 0;
}

}  // extern "C"
