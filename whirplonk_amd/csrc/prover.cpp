// prover.cpp — the WHIR commit and prove loop (C ABI: whir_prover_*, plus the
// prove-path building blocks whir_dev_pow_search / _eq_accumulate /
// _geometric_accumulate / _mle_evaluate).
//
// Follows protocols/whir/whir_prover.hpp:42-400 (prove), irs_commit.hpp
// (commit :328-553, open :636-713), sumcheck_protocol.hpp:82-145 (rounds) and
// proof_of_work.hpp:237-250 (PoW).  The transcript lives on the host
// (transcript.h); every vector, codeword and Merkle tree stays on the device,
// and only messages, challenges, opened rows and path hashes cross PCIe.
#include "../../include/whir_gpu.h"

#include <chrono>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "config.h"
#include "ctx.h"
#include "profile.h"
#include "transcript.h"

using namespace whir_gpu;
using gl::e3;

namespace {

const e3 E3_ZERO = {0, 0, 0};
const e3 E3_ONE = {1, 0, 0};

// A committed codeword kept on the device (irs_commit::Witness).
struct DevWitness {
    uint64_t* matrix = nullptr;  // rows x row_u64 (base, or Ext3 AoS)
    uint8_t* nodes = nullptr;    // flat Merkle tree
    size_t rows = 0, num_cols = 0;
    bool ext = false;
    std::vector<e3> ood_points, ood_values;  // values point-major, one per vector
    size_t row_u64() const { return num_cols * (ext ? 3 : 1); }
};

// Opened rows (irs_commit::Evaluations): points and rows, row-major.  The
// rows and path hashes arrive in the context's mapped buffer; until
// finish_open has read them (and hinted them) the opening is pending.
struct Opened {
    std::vector<uint64_t> points;  // base field (gen^index)
    std::vector<uint64_t> rows;    // count x row_u64
    bool pending = false;
    const void* buf = nullptr;  // ctx->hio when the gather was enqueued
    uint64_t seq = 0;           // the context's last completion number then
    size_t ru = 0, cols = 0, nh = 0, hoff = 0;
    bool ext = false;
};

uint32_t lg(uint64_t n) { return log2u(n); }

size_t align256(size_t x) { return (x + 255) & ~size_t{255}; }

// Device memory of one proof, carved from one region: the committed
// codewords and trees (the initial one and two alternating round slots), the
// folded vector and the covector, and the LDE scratch.  Offsets in bytes.
struct Layout {
    size_t init_matrix = 0, init_nodes = 0, rw_matrix[2] = {0, 0}, rw_nodes[2] = {0, 0}, vec = 0, cov = 0,
           scratch = 0, scratch_bytes = 0, total = 0;
};

size_t lde_scratch_bytes(size_t vector_size, size_t depth, size_t num_vectors, size_t cw, bool ext) {
    LdeShape s;
    s.m = vector_size / depth;
    s.C = (uint64_t)num_vectors * depth * (ext ? 3 : 1);
    s.n = cw;
    s.row_major = true;
    if (ext) {
        s.in_el = 3;
        s.split3 = true;
    }
    return lde_scratch_elems_any_split(s) * 8;
}

Layout plan_layout(const whir_host::WhirConfig& cfg, size_t size) {
    Layout L;
    const auto& ic = cfg.initial;
    size_t mat[2] = {0, 0}, nod[2] = {0, 0};
    size_t scratch = lde_scratch_bytes(ic.vector_size, ic.depth, ic.num_vectors, ic.codeword_length, false);
    for (size_t k = 0; k < cfg.rounds.size(); ++k) {
        const auto& r = cfg.rounds[k].irs;
        mat[k & 1] = std::max(mat[k & 1], r.codeword_length * r.num_cols() * 24);
        nod[k & 1] = std::max(nod[k & 1], whir_merkle_num_nodes(r.codeword_length) * 32);
        scratch = std::max(scratch, lde_scratch_bytes(r.vector_size, r.depth, 1, r.codeword_length, true));
    }
    size_t off = 0;
    auto put = [&off](size_t bytes) {
        size_t o = off;
        off += align256(std::max<size_t>(bytes, 8));
        return o;
    };
    L.init_matrix = put(ic.codeword_length * ic.num_cols() * 8);
    L.init_nodes = put(whir_merkle_num_nodes(ic.codeword_length) * 32);
    for (int k = 0; k < 2; ++k) {
        L.rw_matrix[k] = put(mat[k]);
        L.rw_nodes[k] = put(nod[k]);
    }
    L.vec = put((ic.num_vectors > 1 ? size : size / 2 + 1) * 24);
    L.cov = put(size * 24);
    L.scratch_bytes = align256(scratch);
    L.scratch = put(scratch);
    L.total = off;
    return L;
}

// EngineId of the BLAKE3 hash engine (hash/blake3_engine.hpp:50-55): an
// identifier fixed by the protocol, hashed into the domain separator.
const uint8_t kBlake3EngineId[32] = {0x03, 0xe0, 0x17, 0x49, 0xeb, 0xcc, 0x04, 0x77, 0x92, 0x42, 0x54,
                                     0xeb, 0x48, 0x20, 0x66, 0xb8, 0x64, 0xa8, 0xdd, 0x4d, 0x77, 0x25,
                                     0x24, 0x64, 0xca, 0x6f, 0x5b, 0x6f, 0x5c, 0xc0, 0x5b, 0x4c};
// EngineId of the SHA-256 engine (hash/sha2_engine.hpp:40-44), likewise fixed.
const uint8_t kSha2EngineId[32] = {0x01, 0x8e, 0xaa, 0x24, 0x7c, 0xb8, 0x95, 0x7a, 0xb1, 0xe9, 0xfd,
                                   0xac, 0x88, 0x54, 0x50, 0x40, 0x3c, 0x5e, 0x7b, 0xda, 0x1a, 0xca,
                                   0xaa, 0x50, 0x4e, 0x4c, 0xc8, 0xf2, 0xf7, 0x6b, 0xb0, 0x76};

// The prover's hash engine on its context for the duration of a call.
struct HashScope {
    whir_ctx* ctx;
    uint32_t prev;
    HashScope(whir_ctx* c, uint32_t h) : ctx(c), prev(c->hash) { c->hash = h; }
    ~HashScope() { ctx->hash = prev; }
};

}  // namespace

struct whir_prover {
    whir_ctx* ctx = nullptr;
    whir_host::Params params;
    whir_host::WhirConfig cfg;
    whir_host::Transcript ts;
    size_t size = 0;
    const uint64_t* d_vectors = nullptr;
    size_t num_vectors = 0;
    DevWitness initial;
    DevWitness rw[2];
    Layout L;
    uint8_t* region = nullptr;  // L.total bytes of device memory
    bool own_region = false;
    bool committed = false, proved = false;
    std::vector<e3> eval_point, forms_rlc;
    size_t pow_hashes = 0;
    double commit_ms = 0;         // host wall time of whir_prover_commit (WHIR_PROFILE's commit_total)
    uint32_t hash = HASH_BLAKE3;  // engine of the trees and PoW (params.hash_id)
    // host wall time per phase (include the waits on the device)
    double phase_ms[WHIR_PROVER_PHASES] = {};
    int phase = WHIR_PROVER_PHASES - 1;
    std::chrono::steady_clock::time_point t_last = std::chrono::steady_clock::now();

    int enter(int k) {
        auto now = std::chrono::steady_clock::now();
        phase_ms[phase] += std::chrono::duration<double, std::milli>(now - t_last).count();
        t_last = now;
        int prev = phase;
        phase = k;
        return prev;
    }

    ~whir_prover() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        if (own_region && region) (void)hipFree(region);
    }
    hipStream_t st() const { return ctx->stream; }
};

namespace {

// Attributes the wall time of a scope to one phase (nested scopes pause the outer one).
struct Phase {
    whir_prover* P;
    int prev;
    Phase(whir_prover* p, int k) : P(p), prev(p->enter(k)) {}
    ~Phase() { P->enter(prev); }
};
enum { PH_COMMIT = 0, PH_OOD, PH_WEIGHTS, PH_SUMCHECK, PH_POW, PH_ROUND_COMMIT, PH_OPEN, PH_OTHER };

int sync(whir_prover* P) { return P->ctx->wait(); }

// ------------------------------------------------------------------- PoW
// PowConfig::prove: nothing at zero difficulty; else squeeze a 32-byte
// challenge, find the smallest nonce, send it as 8 LE bytes.
int pow_device(whir_ctx* ctx, const uint8_t ch[32], uint64_t threshold, uint64_t* nonce, size_t* scanned) {
    int rc;
    if (ctx->pow_hook) {  // a serving loop batches the searches of its requests
        if ((rc = ctx->pow_hook(ctx->pow_hook_arg, ch, threshold, nonce))) return rc;
        if (scanned) *scanned += *nonce + 1;
        return 0;
    }
    if ((rc = ctx->ensure_hpart()) || (rc = signal_init(ctx->sig))) return rc;
    if (!ctx->powbuf.p) {
        if ((rc = ctx->ensure(ctx->powbuf, 2 * POW_BATCH_MAX * 8))) return rc;
        fill_u64_launch((uint64_t*)ctx->powbuf.p, 2 * POW_BATCH_MAX, ~0ull, ctx->stream);
        ctx->pow_slot = 0;
    }
    PowJob job;
    pow_job_init(job, ch, threshold);
    job.hash = ctx->hash;
    PowJob* jp = &job;
    for (;;) {
        bool found = false;
        uint64_t n = 0;
        if ((rc = pow_run(ctx->stream, (unsigned long long*)ctx->powbuf.p, ctx->pow_slot, ctx->hpart, ctx->hpart_d,
                          whir_ctx::HPART_U64, &jp, 1, &found, &n, &ctx->sig)))
            return rc;
        if (found) {
            if (scanned) *scanned += n + 1;  // the serial scan's work
            *nonce = n;
            return 0;
        }
    }
}

int pow_prove(whir_prover* P, uint64_t threshold) {
    if (threshold == UINT64_MAX) return 0;
    Phase ph(P, PH_POW);
    uint8_t ch[32];
    P->ts.squeeze(ch, 32);
    uint64_t nonce = 0;
    int rc = pow_device(P->ctx, ch, threshold, &nonce, &P->pow_hashes);
    if (rc) return rc;
    P->ts.message_u64(nonce);
    return 0;
}

// ------------------------------------------------------------- weight terms
struct Terms {
    std::vector<uint64_t> desc_e, desc_b, coords;
    uint32_t ne = 0, nb = 0;

    void add(std::vector<uint64_t>& d, uint64_t kind, e3 s, e3 z, e3 Z, uint64_t off) {
        for (uint64_t v : {kind, s.c0, s.c1, s.c2, z.c0, z.c1, z.c2, Z.c0, Z.c1, Z.c2, off}) d.push_back(v);
    }
    // eq(x, i) * s, x of nv Ext3 coordinates (first = top index bit)
    void eq(const uint64_t* x, uint32_t nv, e3 s) {
        uint64_t off = coords.size() / 3;
        coords.insert(coords.end(), x, x + 3 * nv);
        add(desc_e, 0, s, E3_ZERO, E3_ZERO, off);
        ++ne;
    }
    // z^i * s; base-field z goes to the cheaper group
    void power(e3 z, e3 s, uint32_t L) {
        e3 Z = gl::e3_pow(z, 1ull << L);
        if (z.c1 == 0 && z.c2 == 0) {
            add(desc_b, 1, s, z, Z, 0);
            ++nb;
        } else {
            add(desc_e, 1, s, z, Z, 0);
            ++ne;
        }
    }
};

uint32_t split_bits(uint32_t nv) { return nv < 10 ? nv : 10; }

// acc[i] = (accumulate ? acc[i] : 0) + sum of the terms, i < 2^nv
int terms_apply(whir_ctx* ctx, const Terms& T, uint64_t* d_acc, uint32_t nv, bool accumulate, whir_prover* P = nullptr) {
    std::unique_ptr<Phase> ph;
    if (P) ph.reset(new Phase(P, PH_WEIGHTS));
    const uint32_t L = split_bits(nv);
    const uint64_t H = 1ull << (nv - L), LO = 1ull << L;
    int rc;
    size_t desc_bytes = (T.desc_e.size() + T.desc_b.size() + T.coords.size() + 4) * 8;
    size_t he = (size_t)T.ne * H * 24, le = (size_t)T.ne * LO * 24, hb = (size_t)T.nb * H * 24,
           lb = (size_t)T.nb * LO * 8, hlb = term_limb_bytes(nv, L, T.ne + T.nb);
    if ((rc = ctx->ensure(ctx->tdesc, desc_bytes))) return rc;
    if ((rc = ctx->ensure(ctx->ttab, he + le + hb + lb + hlb + 256))) return rc;
    uint64_t* dd = (uint64_t*)ctx->tdesc.p;
    uint64_t* de = dd;
    uint64_t* db = de + T.desc_e.size();
    uint64_t* dc = db + T.desc_b.size();
    std::vector<uint64_t> all;
    all.reserve(T.desc_e.size() + T.desc_b.size() + T.coords.size());
    all.insert(all.end(), T.desc_e.begin(), T.desc_e.end());
    all.insert(all.end(), T.desc_b.begin(), T.desc_b.end());
    all.insert(all.end(), T.coords.begin(), T.coords.end());
    if ((rc = ctx->stage_desc(all, dd))) return rc;  // asynchronous: no stream sync here
    uint64_t* tab = (uint64_t*)ctx->ttab.p;
    uint64_t* hi_e = tab;
    uint64_t* lo_e = hi_e + he / 8;
    uint64_t* hi_b = lo_e + le / 8;
    uint64_t* lo_b = hi_b + hb / 8;
    uint32_t* hl = hlb ? (uint32_t*)(lo_b + lb / 8) : nullptr;
    term_tables_launch(de, T.ne, db, T.nb, dc, nv, L, hi_e, lo_e, hi_b, lo_b, hl, ctx->stream);
    term_accumulate_launch(d_acc, 1ull << nv, L, hi_e, lo_e, T.ne, hi_b, lo_b, T.nb, hl, accumulate, ctx->stream);
    HIPCHK(hipGetLastError());
    return 0;
}

// --------------------------------------------------------------- sumcheck
// (c0, c2) of the g block partials a launcher left in the mapped host buffer
int read_pair(whir_ctx* ctx, unsigned g, e3& c0, e3& c2) {
    e3 c[2];
    int rc = ctx->host_sums(g, 2, c);
    c0 = c[0];
    c2 = c[1];
    return rc;
}

// sumcheck::Config::prove over (a, b) of length n: a_out receives the folded
// a (may equal a when a is Ext3); b is folded in place.  Returns the
// challenges in `rs`.
// `settle`, when given, runs after the first round's sums have arrived and
// before `sum` is read (or, with no rounds, at once): it may finish host work
// that waits on earlier launches of the stream, such as an opening whose
// rows update `sum`.
int sumcheck_prove(whir_prover* P, const uint64_t* a, bool a_base, uint64_t* a_out, uint64_t* b, size_t n,
                   size_t rounds, uint64_t round_pow, e3& sum, std::vector<e3>& rs,
                   const std::function<int()>& settle = {}) {
    whir_ctx* ctx = P->ctx;
    rs.clear();
    if (rounds == 0) return settle ? settle() : 0;
    Phase ph(P, PH_SUMCHECK);
    int rc;
    if ((rc = ctx->ensure_hpart())) return rc;
    uint64_t* part = ctx->hpart_d;
    e3 c0, c2;
    unsigned g = sumcheck_compute_launch(a, a_base, b, n, part, ctx->stream, false);
    HIPCHK(hipGetLastError());
    if ((rc = read_pair(ctx, g, c0, c2))) return rc;
    if (settle && (rc = settle())) return rc;
    for (size_t r = 0; r < rounds; ++r) {
        e3 c1 = gl::e3_sub(gl::e3_sub(sum, gl::e3_add(c0, c0)), c2);
        P->ts.message_e3(c0);
        P->ts.message_e3(c2);
        if ((rc = pow_prove(P, round_pow))) return rc;
        e3 rnd = P->ts.challenge_e3();
        rs.push_back(rnd);
        uint64_t rr[3] = {rnd.c0, rnd.c1, rnd.c2};
        e3 nc0 = E3_ZERO, nc2 = E3_ZERO;
        if (r + 1 < rounds) {
            g = fold_compute_launch(a, a_base, b, n, rr, a_out, b, part, ctx->stream, false);
            HIPCHK(hipGetLastError());
            if ((rc = read_pair(ctx, g, nc0, nc2))) return rc;
        } else {
            sumcheck_fold2_launch(a, a_base, a_out, b, b, n, rr, ctx->stream);
            HIPCHK(hipGetLastError());
        }
        sum = gl::e3_add(gl::e3_mul(gl::e3_add(gl::e3_mul(c2, rnd), c1), rnd), c0);
        c0 = nc0;
        c2 = nc2;
        n /= 2;
        a = a_out;
        a_base = false;
    }
    return 0;
}

// --------------------------------------------------------------- commit
// irs_commit::Config::commit: encode + leaves + tree, root message, OOD points
// squeezed, OOD values (point-major, vector-minor) sent.
int commit_witness(whir_prover* P, const whir_host::IrsConfig& c, const uint64_t* d_vecs, bool ext, DevWitness& w,
                   size_t matrix_off, size_t nodes_off) {
    whir_ctx* ctx = P->ctx;
    int rc;
    w.rows = c.codeword_length;
    w.num_cols = c.num_cols();
    w.ext = ext;
    w.matrix = (uint64_t*)(P->region + matrix_off);
    w.nodes = P->region + nodes_off;
    uint8_t root[32];
    {
        Phase ph(P, ext ? PH_ROUND_COMMIT : PH_COMMIT);
        ctx->ext_scratch = P->region + P->L.scratch;
        ctx->ext_scratch_bytes = P->L.scratch_bytes;
        // WHIR_PROFILE: the stage times of this commit from the context's HIP events
        const bool prof_on = prof::enabled();
        const bool was_timing = ctx->timing;
        double before[3] = {ctx->acc_ms[0], ctx->acc_ms[1], ctx->acc_ms[2]};
        if (prof_on) ctx->timing = true;
        rc = commit_device(ctx, d_vecs, c.num_vectors, c.vector_size, c.codeword_length, c.depth, w.matrix, w.nodes,
                           root, ext ? 1 : 0);
        ctx->timing = was_timing;
        ctx->ext_scratch = nullptr;
        ctx->ext_scratch_bytes = 0;
        if (rc) return rc;
        if (prof_on) {
            // witness_encoding covers the encode and the row leaves (fused into
            // the encode's last pass for 16-column base rows)
            prof::record("prover", c.num_vectors * c.vector_size, "witness_encoding",
                         (ctx->acc_ms[0] - before[0]) + (ctx->acc_ms[1] - before[1]));
            prof::record("prover", c.codeword_length, "merkle_build_total", ctx->acc_ms[2] - before[2]);
        }
    }
    P->ts.message(root, 32);
    const auto t_ood = std::chrono::steady_clock::now();
    struct OodRow {
        size_t n;
        std::chrono::steady_clock::time_point t0;
        ~OodRow() {
            prof::record("prover", n, "ood_evaluation",
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
    } ood_row{c.vector_size, t_ood};
    Phase ph(P, PH_OOD);
    w.ood_points.clear();
    w.ood_values.clear();
    for (size_t k = 0; k < c.out_domain; ++k) w.ood_points.push_back(P->ts.challenge_e3());
    for (const e3& z : w.ood_points) {
        for (size_t v = 0; v < c.num_vectors; ++v) {
            uint64_t zz[3] = {z.c0, z.c1, z.c2}, out[3];
            const uint64_t* f = d_vecs + v * c.vector_size * (ext ? 3 : 1);
            if ((rc = whir_dev_horner(ctx, f, ext ? 1 : 0, c.vector_size, zz, out))) return rc;
            e3 val = {out[0], out[1], out[2]};
            P->ts.message_e3(val);
            w.ood_values.push_back(val);
        }
    }
    return 0;
}

// irs_commit::Config::open for one witness: squeeze the indices, hint the
// opened rows (one vector of elements), then the Merkle path hashes.  The
// gather is only enqueued here; finish_open reads its output and appends the
// hints (hints are not absorbed, so appending them later changes no
// challenge), usually after the next wait the prove loop needs anyway.
int open_witness(whir_prover* P, const whir_host::IrsConfig& c, const DevWitness& w, Opened& o) {
    whir_ctx* ctx = P->ctx;
    int rc;
    Phase ph(P, PH_OPEN);
    std::vector<size_t> idx = P->ts.indices(c.codeword_length, c.in_domain, false);
    const size_t k = idx.size(), ru = w.row_u64();
    const uint64_t gen = gl::root_of_unity(lg(c.codeword_length));
    o.points.resize(k);
    for (size_t i = 0; i < k; ++i) o.points[i] = gl::pow(gen, idx[i]);
    std::vector<uint64_t> idx64(idx.begin(), idx.end());
    size_t nh = whir_merkle_hint_node_indices(w.rows, idx64.data(), k, nullptr);
    std::vector<uint64_t> hidx(nh);
    whir_merkle_hint_node_indices(w.rows, idx64.data(), k, hidx.data());
    // Indices in, rows and hashes out, all through the mapped staging buffer:
    // the gather kernels read the indices over PCIe and write their output
    // straight to host memory, so the opening costs no copies.
    // Layout (u64): rows [0, k*ru), hashes from the next even offset (their
    // 16-byte stores stay aligned), then the row and node indices.
    const size_t hoff = (k * ru + 1) & ~size_t{1}, ioff = hoff + 4 * nh;
    if ((rc = ctx->ensure_hio((ioff + k + nh) * 8))) return rc;
    uint64_t* back = (uint64_t*)ctx->hio;
    uint64_t* back_d = (uint64_t*)ctx->hio_d;
    std::memcpy(back + ioff, idx64.data(), k * 8);
    std::memcpy(back + ioff + k, hidx.data(), nh * 8);
    gather_open_launch(w.matrix, ru, back_d + ioff, k, back_d, w.nodes, back_d + ioff + k, nh,
                       (uint8_t*)(back_d + hoff), ctx->stream);
    HIPCHK(hipGetLastError());
    o.pending = true;
    o.buf = back;
    o.seq = ctx->sig.seq;
    o.ru = ru;
    o.cols = w.num_cols;
    o.nh = nh;
    o.hoff = hoff;
    o.ext = w.ext;
    return 0;
}

// Reads a pending opening's rows and hashes and appends its hints; waits
// first unless a completion signal enqueued after its gather has been seen.
// Nothing may grow the context's mapped buffer between the two calls.
int finish_open(whir_prover* P, Opened& o) {
    if (!o.pending) return 0;
    whir_ctx* ctx = P->ctx;
    int rc;
    if (ctx->sig.seq == o.seq && (rc = ctx->wait())) return rc;
    if (ctx->hio != o.buf) return bad_arg("the opening's buffer moved before its rows were read");
    const uint64_t* back = (const uint64_t*)o.buf;
    const size_t k = o.points.size();
    o.rows.assign(back, back + k * o.ru);
    P->ts.hint_vector(o.rows.data(), k * o.cols, o.ext ? 24 : 8);
    for (size_t h = 0; h < o.nh; ++h) P->ts.hint(back + o.hoff + 4 * h, 32);
    o.pending = false;
    return 0;
}

e3 row_elem(const Opened& o, size_t r, size_t c, size_t cols, bool ext) {
    if (!ext) return {o.rows[r * cols + c], 0, 0};
    const uint64_t* p = o.rows.data() + 3 * (r * cols + c);
    return {p[0], p[1], p[2]};
}

// MultilinearPoint::eq_weights (multilinear_point.hpp:56-62): entry i is the
// product over coordinates, the last coordinate pairing with bit 0 of i.
std::vector<e3> eq_weights(const std::vector<e3>& x) {
    const size_t n = size_t{1} << x.size();
    std::vector<e3> out(n);
    for (size_t i = 0; i < n; ++i) {
        e3 acc = E3_ONE;
        size_t bits = i;
        for (size_t k = x.size(); k-- > 0;) {
            acc = gl::e3_mul(acc, (bits & 1) ? x[k] : gl::e3_sub(E3_ONE, x[k]));
            bits >>= 1;
        }
        out[i] = acc;
    }
    return out;
}

int prove_impl(whir_prover* P, size_t num_forms, const uint64_t* points, const uint64_t* evaluations) {
    whir_ctx* ctx = P->ctx;
    const auto& cfg = P->cfg;
    const size_t n = P->size, nvec = P->num_vectors;
    const uint32_t nv = lg(n);
    int rc;

    // OOD completion: one witness covers every vector, so nothing is sent and
    // the committed OOD values are reused (whir_prover.hpp:96-131).
    std::vector<e3> vrlc = P->ts.geometric(nvec);
    std::vector<e3> crlc = P->ts.geometric(num_forms + P->initial.ood_points.size());
    const bool has_constraints = !crlc.empty();
    P->forms_rlc.assign(crlc.begin(), crlc.begin() + std::min(num_forms, crlc.size()));

    // vector and covector buffers
    uint64_t* vec = (uint64_t*)(P->region + P->L.vec);
    uint64_t* cov = (uint64_t*)(P->region + P->L.cov);
    const uint64_t* a = P->d_vectors;
    bool a_base = true;
    if (nvec > 1) {
        std::vector<uint64_t> cf(3 * nvec);
        for (size_t k = 0; k < nvec; ++k) {
            cf[3 * k] = vrlc[k].c0;
            cf[3 * k + 1] = vrlc[k].c1;
            cf[3 * k + 2] = vrlc[k].c2;
        }
        if ((rc = ctx->ensure(ctx->tdesc, cf.size() * 8))) return rc;
        HIPCHK(hipMemcpyAsync(ctx->tdesc.p, cf.data(), cf.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        rlc_combine_launch(P->d_vectors, n, (uint32_t)nvec, (const uint64_t*)ctx->tdesc.p, vec, ctx->stream);
        HIPCHK(hipGetLastError());
        if ((rc = sync(P))) return rc;
        a = vec;
        a_base = false;
    }

    // covector = sum_i crlc_i eq(point_i, .) + sum_o oods_rlc_o z_o^(.)
    if (has_constraints) {
        Terms T;
        const uint32_t L = split_bits(nv);
        for (size_t f = 0; f < num_forms; ++f) T.eq(points + 3 * nv * f, nv, crlc[f]);
        for (size_t o = 0; o < P->initial.ood_points.size(); ++o)
            T.power(P->initial.ood_points[o], crlc[num_forms + o], L);
        if ((rc = terms_apply(ctx, T, cov, nv, false, P))) return rc;
    }

    // the_sum (whir_prover.hpp:171-197)
    e3 sum = E3_ZERO;
    for (size_t i = 0; i < num_forms; ++i) {
        e3 row = E3_ZERO;
        for (size_t j = 0; j < nvec; ++j) {
            const uint64_t* e = evaluations + 3 * (i * nvec + j);
            row = gl::e3_add(row, gl::e3_mul(vrlc[j], e3{e[0], e[1], e[2]}));
        }
        sum = gl::e3_add(sum, gl::e3_mul(crlc[i], row));
    }
    for (size_t o = 0; o < P->initial.ood_points.size(); ++o) {
        e3 row = E3_ZERO;
        for (size_t j = 0; j < nvec; ++j)
            row = gl::e3_add(row, gl::e3_mul(vrlc[j], P->initial.ood_values[o * nvec + j]));
        sum = gl::e3_add(sum, gl::e3_mul(crlc[num_forms + o], row));
    }

    // initial sumcheck, or the fold without one (whir_prover.hpp:200-225)
    std::vector<e3> fr;
    size_t cur = n;
    if (has_constraints) {
        if ((rc = sumcheck_prove(P, a, a_base, vec, cov, n, cfg.initial_sumcheck.num_rounds,
                                 cfg.initial_sumcheck.round_pow, sum, fr)))
            return rc;
    } else {
        for (size_t r = 0; r < cfg.initial_sumcheck.num_rounds; ++r) fr.push_back(P->ts.challenge_e3());
        if ((rc = pow_prove(P, cfg.initial_skip_pow))) return rc;
        for (const e3& f : fr) {
            uint64_t rr[3] = {f.c0, f.c1, f.c2};
            sumcheck_fold_launch(a, a_base, cur, rr, vec, ctx->stream);
            a = vec;
            a_base = false;
            cur /= 2;
        }
        HIPCHK(hipGetLastError());
        fill_zero_launch((uint8_t*)cov, cfg.initial_sumcheck.final_size() * 24, ctx->stream);
    }
    if (cfg.initial_sumcheck.num_rounds == 0 && a != vec) {
        // no fold happened: the prove loop continues on an Ext3 copy
        return bad_arg("zero initial folding rounds are not supported");
    }
    cur = cfg.initial_sumcheck.final_size();
    P->eval_point = fr;
    std::vector<e3> last_fr = fr;

    const DevWitness* prev = nullptr;
    const whir_host::IrsConfig* prev_cfg = nullptr;
    for (size_t k = 0; k < cfg.rounds.size(); ++k) {
        const auto& rc_cfg = cfg.rounds[k];
        DevWitness& w = P->rw[k & 1];
        if ((rc = commit_witness(P, rc_cfg.irs, vec, true, w, P->L.rw_matrix[k & 1], P->L.rw_nodes[k & 1]))) return rc;
        if ((rc = pow_prove(P, rc_cfg.pow))) return rc;
        Opened o;
        bool o_ext;
        size_t o_cols;
        if (!prev) {
            if ((rc = open_witness(P, cfg.initial, P->initial, o))) return rc;
            o_ext = false;
            o_cols = P->initial.num_cols;
        } else {
            if ((rc = open_witness(P, *prev_cfg, *prev, o))) return rc;
            o_ext = true;
            o_cols = prev->num_cols;
        }
        // STIR constraints (whir_prover.hpp:286-345)
        std::vector<e3> eqw = last_fr.empty() ? std::vector<e3>{E3_ONE} : eq_weights(last_fr);
        std::vector<e3> tp;
        for (const e3& v : vrlc)
            for (const e3& e : eqw) tp.push_back(gl::e3_mul(v, e));
        if (tp.size() != o_cols) return bad_arg("opened row width does not match the folding weights");
        const size_t nstir = w.ood_points.size() + o.points.size();
        std::vector<e3> srlc = P->ts.geometric(nstir);
        {
            Terms T;
            const uint32_t lv = lg(cur), L = split_bits(lv);
            for (size_t j = 0; j < w.ood_points.size(); ++j) T.power(w.ood_points[j], srlc[j], L);
            for (size_t j = 0; j < o.points.size(); ++j)
                T.power(e3{o.points[j], 0, 0}, srlc[w.ood_points.size() + j], L);
            if (nstir && (rc = terms_apply(ctx, T, cov, lv, true, P))) return rc;
        }
        // the opened rows' values at the query points join the claimed sum
        // once they have arrived: after the round sumcheck's first sums (one
        // wait for both), before its first message
        auto settle = [&]() -> int {
            if (int e = finish_open(P, o)) return e;
            std::vector<e3> stir_eval(w.ood_values);
            for (size_t r = 0; r < o.points.size(); ++r) {
                e3 acc = E3_ZERO;
                for (size_t c = 0; c < o_cols; ++c)
                    acc = gl::e3_add(acc, gl::e3_mul(row_elem(o, r, c, o_cols, o_ext), tp[c]));
                stir_eval.push_back(acc);
            }
            for (size_t j = 0; j < nstir; ++j) sum = gl::e3_add(sum, gl::e3_mul(srlc[j], stir_eval[j]));
            return 0;
        };
        // round sumcheck
        if ((rc = sumcheck_prove(P, vec, false, vec, cov, cur, rc_cfg.sumcheck.num_rounds,
                                 rc_cfg.sumcheck.round_pow, sum, fr, settle)))
            return rc;
        P->eval_point.insert(P->eval_point.end(), fr.begin(), fr.end());
        last_fr = fr;
        cur = rc_cfg.sumcheck.final_size();
        vrlc = {E3_ONE};
        prev = &w;
        prev_cfg = &rc_cfg.irs;
    }

    // final vector, PoW, opening, sumcheck (whir_prover.hpp:355-398)
    if (cur != cfg.final_sumcheck.initial_size) return bad_arg("final vector size mismatch");
    {
        // into the pinned staging buffer (an asynchronous copy), then the
        // completion word instead of a pageable copy and a stream sync
        if ((rc = ctx->ensure_hio(24 * cur))) return rc;
        HIPCHK(hipMemcpyAsync(ctx->hio, vec, 24 * cur, hipMemcpyDeviceToHost, ctx->stream));
        if ((rc = ctx->wait())) return rc;
        const uint64_t* fin = (const uint64_t*)ctx->hio;
        for (size_t i = 0; i < cur; ++i) P->ts.message(fin + 3 * i, 24);
    }
    if ((rc = pow_prove(P, cfg.final_pow))) return rc;
    Opened fo;
    if (!prev) {
        if ((rc = open_witness(P, cfg.initial, P->initial, fo))) return rc;
    } else {
        if ((rc = open_witness(P, *prev_cfg, *prev, fo))) return rc;
    }
    // the final opening's hints, after the final sumcheck's first sums
    if ((rc = sumcheck_prove(P, vec, false, vec, cov, cur, cfg.final_sumcheck.num_rounds,
                             cfg.final_sumcheck.round_pow, sum, fr, [&] { return finish_open(P, fo); })))
        return rc;
    P->eval_point.insert(P->eval_point.end(), fr.begin(), fr.end());
    // with no final rounds, reading the final opening was the last wait and
    // nothing has been enqueued since; otherwise the last fold is in flight
    return cfg.final_sumcheck.num_rounds == 0 ? 0 : sync(P);
}

whir_host::Params to_params(const whir_params* p) {
    whir_host::Params q;
    q.unique_decoding = p->unique_decoding != 0;
    q.starting_log_inv_rate = p->starting_log_inv_rate;
    q.initial_folding_factor = p->initial_folding_factor;
    q.folding_factor = p->folding_factor;
    q.security_level = p->security_level;
    q.pow_bits = p->pow_bits;
    q.batch_size = p->batch_size;
    return q;
}

int check_params(const whir_params* p, uint64_t size) {
    if (!p) return bad_arg("null parameters");
    if (p->hash_id != WHIR_HASH_BLAKE3 && p->hash_id != WHIR_HASH_SHA2) return bad_arg("unknown hash engine");
    if (!is_pow2(size) || size < 2) return bad_arg("size must be a power of two >= 2");
    if (p->batch_size == 0) return bad_arg("batch_size must be >= 1");
    if (p->folding_factor == 0 || p->initial_folding_factor == 0) return bad_arg("folding factors must be >= 1");
    if (p->pow_bits > p->security_level) return bad_arg("pow_bits above security_level");
    if ((1ull << p->initial_folding_factor) > size) return bad_arg("initial folding factor larger than log2(size)");
    return 0;
}

void put_e3s(const std::vector<e3>& v, uint64_t* out) {
    for (size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = v[i].c0;
        out[3 * i + 1] = v[i].c1;
        out[3 * i + 2] = v[i].c2;
    }
}

}  // namespace

extern "C" {

void whir_params_default(whir_params* p) {
    // main.cpp Args defaults (PCS, list decoding, BLAKE3)
    p->unique_decoding = 0;
    p->starting_log_inv_rate = 1;
    p->initial_folding_factor = 4;
    p->folding_factor = 4;
    p->security_level = 128;
    p->pow_bits = 20;
    p->batch_size = 1;
    p->hash_id = WHIR_HASH_BLAKE3;
}

size_t whir_config_flat(uint64_t size, const whir_params* p, uint64_t* out, size_t cap) {
    if (check_params(p, size)) return 0;
    auto v = whir_host::WhirConfig::derive(size, to_params(p)).flat();
    if (out) std::memcpy(out, v.data(), std::min(cap, v.size()) * 8);
    return v.size();
}

int whir_domain_separator_cli(const whir_params* p, const char* session, uint8_t* protocol_id,
                              uint8_t* session_id) {
    if (!p) return bad_arg("null parameters");
    // CBOR map of ProtocolParameters in the CLI's key order (main.cpp:176-215)
    std::vector<uint8_t> m;
    whir_host::cbor_head(m, 5, 8);
    whir_host::cbor_text(m, "unique_decoding");
    m.push_back(p->unique_decoding ? 0xF5 : 0xF4);
    whir_host::cbor_text(m, "starting_log_inv_rate");
    whir_host::cbor_uint(m, p->starting_log_inv_rate);
    whir_host::cbor_text(m, "initial_folding_factor");
    whir_host::cbor_uint(m, p->initial_folding_factor);
    whir_host::cbor_text(m, "folding_factor");
    whir_host::cbor_uint(m, p->folding_factor);
    whir_host::cbor_text(m, "security_level");
    whir_host::cbor_uint(m, p->security_level);
    whir_host::cbor_text(m, "pow_bits");
    whir_host::cbor_uint(m, p->pow_bits);
    whir_host::cbor_text(m, "batch_size");
    whir_host::cbor_uint(m, p->batch_size);
    whir_host::cbor_text(m, "hash_id");
    // EngineId bytes of the hash engine, as an array of byte-valued integers
    const uint8_t* id = p->hash_id == WHIR_HASH_SHA2 ? kSha2EngineId : kBlake3EngineId;
    whir_host::cbor_head(m, 4, 32);
    for (int i = 0; i < 32; ++i) whir_host::cbor_uint(m, id[i]);
    whir_host::sha3_512(m.data(), m.size(), protocol_id);
    std::vector<uint8_t> s;
    whir_host::cbor_text(s, session ? session : "Example at src/bin/main.rs:133");
    whir_host::sha3_256(s.data(), s.size(), session_id);
    return 0;
}

static int prover_create(whir_ctx* ctx, const whir_params* params, uint64_t size, const uint8_t* protocol_id,
                         const uint8_t* session_id, const uint8_t* instance, size_t instance_len, void* d_workspace,
                         size_t workspace_bytes, whir_prover** out) {
    if (!ctx || !out || !protocol_id || !session_id) return bad_arg("null argument");
    int rc = check_params(params, size);
    if (rc) return rc;
    whir_prover* P = new whir_prover();
    P->ctx = ctx;
    P->params = to_params(params);
    P->hash = params->hash_id == WHIR_HASH_SHA2 ? HASH_SHA2 : HASH_BLAKE3;
    P->size = size;
    P->cfg = whir_host::WhirConfig::derive(size, P->params);
    P->L = plan_layout(P->cfg, size);
    if (d_workspace) {
        if (workspace_bytes < P->L.total || ((uintptr_t)d_workspace & 255)) {
            delete P;
            return bad_arg("workspace too small or not 256-byte aligned");
        }
        P->region = (uint8_t*)d_workspace;
    } else {
        hipError_t e = hipMalloc((void**)&P->region, P->L.total);
        if (e != hipSuccess) {
            delete P;
            return fail("hipMalloc(prover region)", e);
        }
        P->own_region = true;
    }
    P->ts.start(protocol_id, session_id, instance, instance_len);
    *out = P;
    return 0;
}

int whir_prover_create(whir_ctx* ctx, const whir_params* params, uint64_t size, const uint8_t* protocol_id,
                       const uint8_t* session_id, const uint8_t* instance, size_t instance_len, whir_prover** out) {
    return prover_create(ctx, params, size, protocol_id, session_id, instance, instance_len, nullptr, 0, out);
}

int whir_prover_create_in(whir_ctx* ctx, const whir_params* params, uint64_t size, const uint8_t* protocol_id,
                          const uint8_t* session_id, const uint8_t* instance, size_t instance_len, void* d_workspace,
                          size_t workspace_bytes, whir_prover** out) {
    if (!d_workspace) return bad_arg("null workspace");
    return prover_create(ctx, params, size, protocol_id, session_id, instance, instance_len, d_workspace,
                         workspace_bytes, out);
}

size_t whir_prover_workspace_bytes(const whir_params* params, uint64_t size) {
    if (check_params(params, size)) return 0;
    auto cfg = whir_host::WhirConfig::derive(size, to_params(params));
    return plan_layout(cfg, size).total;
}

int whir_prover_destroy(whir_prover* P) {
    if (!P) return 0;
    std::lock_guard<std::recursive_mutex> g(P->ctx->mu);
    delete P;
    return 0;
}

int whir_prover_commit(whir_prover* P, const uint64_t* d_vectors, size_t num_vectors) {
    if (!P || !d_vectors) return bad_arg("null argument");
    if (P->committed) return bad_arg("already committed");
    if (num_vectors != P->cfg.initial.num_vectors) return bad_arg("num_vectors must equal batch_size");
    std::lock_guard<std::recursive_mutex> g(P->ctx->mu);
    HashScope hs(P->ctx, P->hash);
    prof::PhaseGuard phase("commit");
    const auto t0 = std::chrono::steady_clock::now();
    P->d_vectors = d_vectors;
    P->num_vectors = num_vectors;
    int rc = commit_witness(P, P->cfg.initial, d_vectors, false, P->initial, P->L.init_matrix, P->L.init_nodes);
    if (rc) return rc;
    P->committed = true;
    P->commit_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    prof::record("prover", P->size, "commit_total", P->commit_ms);
    return 0;
}

int whir_prover_prove(whir_prover* P, size_t num_forms, const uint64_t* points, const uint64_t* evaluations,
                      uint64_t* out_eval_point, uint64_t* out_forms_rlc) {
    if (!P) return bad_arg("null prover");
    if (!P->committed) return bad_arg("commit before prove");
    if (P->proved) return bad_arg("already proved");
    if (num_forms && (!points || !evaluations)) return bad_arg("null constraint arrays");
    std::lock_guard<std::recursive_mutex> g(P->ctx->mu);
    HashScope hs(P->ctx, P->hash);
    double prove_ms = 0;
    {
        prof::PhaseGuard phase("prove");
        const auto t0 = std::chrono::steady_clock::now();
        int rc = prove_impl(P, num_forms, points, evaluations);
        if (rc) return rc;
        prove_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        prof::record("prover", P->size, "prove_total", prove_ms);
    }
    prof::record("prover", P->size, "total_prover", P->commit_ms + prove_ms);
    P->proved = true;
    if (out_eval_point) put_e3s(P->eval_point, out_eval_point);
    if (out_forms_rlc) put_e3s(P->forms_rlc, out_forms_rlc);
    return 0;
}

int whir_prover_proof(whir_prover* P, uint8_t* narg, size_t narg_cap, size_t* narg_len, uint8_t* hints,
                      size_t hints_cap, size_t* hints_len) {
    if (!P) return bad_arg("null prover");
    const auto& a = P->ts.narg();
    const auto& h = P->ts.hints();
    if (narg_len) *narg_len = a.size();
    if (hints_len) *hints_len = h.size();
    if ((narg && a.size() > narg_cap) || (hints && h.size() > hints_cap)) return bad_arg("proof buffer too small");
    if (narg && !a.empty()) std::memcpy(narg, a.data(), a.size());
    if (hints && !h.empty()) std::memcpy(hints, h.data(), h.size());
    return 0;
}

size_t whir_prover_pow_hashes(const whir_prover* P) { return P ? P->pow_hashes : 0; }

int whir_prover_phase_times(whir_prover* P, double* out_ms) {
    if (!P || !out_ms) return bad_arg("null argument");
    P->enter(P->phase);  // close the running interval
    for (int k = 0; k < WHIR_PROVER_PHASES; ++k) out_ms[k] = P->phase_ms[k];
    return 0;
}

// ------------------------------------------------------------ transcript

struct whir_transcript {
    whir_host::Transcript ts;
};

int whir_transcript_create(const uint8_t* protocol_id, const uint8_t* session_id, const uint8_t* instance,
                           size_t instance_len, whir_transcript** out) {
    if (!out) return bad_arg("null output");
    if ((protocol_id == nullptr) != (session_id == nullptr)) return bad_arg("give both ids or neither");
    whir_transcript* t = new whir_transcript();
    if (protocol_id) t->ts.start(protocol_id, session_id, instance, instance_len);
    *out = t;
    return 0;
}

int whir_transcript_destroy(whir_transcript* t) {
    delete t;
    return 0;
}

int whir_transcript_absorb(whir_transcript* t, const uint8_t* data, size_t len) {
    if (!t || (len && !data)) return bad_arg("null argument");
    t->ts.absorb(data, len);
    return 0;
}

int whir_transcript_message(whir_transcript* t, const uint8_t* data, size_t len) {
    if (!t || (len && !data)) return bad_arg("null argument");
    t->ts.message(data, len);
    return 0;
}

int whir_transcript_hint(whir_transcript* t, const uint8_t* data, size_t len) {
    if (!t || (len && !data)) return bad_arg("null argument");
    t->ts.hint(data, len);
    return 0;
}

int whir_transcript_hint_vector(whir_transcript* t, const uint8_t* data, size_t count, size_t elem_bytes) {
    if (!t || (count && !data)) return bad_arg("null argument");
    t->ts.hint_vector(data, count, elem_bytes);
    return 0;
}

int whir_transcript_squeeze(whir_transcript* t, uint8_t* out, size_t len) {
    if (!t || (len && !out)) return bad_arg("null argument");
    t->ts.squeeze(out, len);
    return 0;
}

int whir_transcript_challenge_gl(whir_transcript* t, uint64_t* out) {
    if (!t || !out) return bad_arg("null argument");
    *out = t->ts.challenge_gl();
    return 0;
}

size_t whir_transcript_indices(whir_transcript* t, size_t num_leaves, size_t count, int dedup, uint64_t* out) {
    if (!t || !is_pow2(num_leaves)) return 0;
    std::vector<size_t> v = t->ts.indices(num_leaves, count, dedup != 0);
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return v.size();
}

int whir_transcript_proof(whir_transcript* t, uint8_t* narg, size_t narg_cap, size_t* narg_len, uint8_t* hints,
                          size_t hints_cap, size_t* hints_len) {
    if (!t) return bad_arg("null transcript");
    const auto& a = t->ts.narg();
    const auto& h = t->ts.hints();
    if (narg_len) *narg_len = a.size();
    if (hints_len) *hints_len = h.size();
    if ((narg && a.size() > narg_cap) || (hints && h.size() > hints_cap)) return bad_arg("proof buffer too small");
    if (narg && !a.empty()) std::memcpy(narg, a.data(), a.size());
    if (hints && !h.empty()) std::memcpy(hints, h.data(), h.size());
    return 0;
}

// ------------------------------------------------------------ building blocks

int whir_dev_pow_search(whir_ctx* ctx, const uint8_t* challenge, uint64_t threshold, uint64_t* out_nonce) {
    if (!ctx || !challenge || !out_nonce) return bad_arg("null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (threshold == UINT64_MAX) {
        *out_nonce = 0;  // find_nonce returns 0 at zero difficulty
        return 0;
    }
    return pow_device(ctx, challenge, threshold, out_nonce, nullptr);
}

int whir_dev_eq_accumulate(whir_ctx* ctx, uint64_t* d_acc, size_t num_variables, const uint64_t* point,
                           const uint64_t* scalar, int accumulate) {
    if (!ctx || !d_acc || (num_variables && !point) || !scalar) return bad_arg("null argument");
    if (num_variables > 30) return bad_arg("num_variables above 30");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    Terms T;
    T.eq(point, (uint32_t)num_variables, e3{scalar[0], scalar[1], scalar[2]});
    return terms_apply(ctx, T, d_acc, (uint32_t)num_variables, accumulate != 0);
}

int whir_dev_geometric_accumulate(whir_ctx* ctx, uint64_t* d_acc, size_t n, const uint64_t* points,
                                  const uint64_t* scalars, size_t count) {
    if (!ctx || !d_acc || (count && (!points || !scalars))) return bad_arg("null argument");
    if (!is_pow2(n) || n > (1ull << 30)) return bad_arg("accumulator length must be a power of two <= 2^30");
    if (!count) return 0;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    Terms T;
    const uint32_t nv = lg(n), L = split_bits(nv);
    for (size_t j = 0; j < count; ++j)
        T.power(e3{points[3 * j], points[3 * j + 1], points[3 * j + 2]},
                e3{scalars[3 * j], scalars[3 * j + 1], scalars[3 * j + 2]}, L);
    return terms_apply(ctx, T, d_acc, nv, true);
}

int whir_dev_mle_evaluate_many(whir_ctx* ctx, const uint64_t* d_f, int f_is_base, size_t num_variables,
                               const uint64_t* points, size_t count, uint64_t* out) {
    if (!ctx || !d_f || (count && !out) || (num_variables && count && !points)) return bad_arg("null argument");
    if (num_variables > 30) return bad_arg("num_variables above 30");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    static_assert(3 * 2 * SUMCHECK_MAX_BLOCKS <= whir_ctx::HPART_U64 &&
                      3 * MLE_MAX_POINTS * (SUMCHECK_MAX_BLOCKS / 2) <= whir_ctx::HPART_U64,
                  "the evaluation's block totals fit the mapped partials");
    const uint32_t nv = (uint32_t)num_variables, L = split_bits(nv);
    const uint64_t H = 1ull << (nv - L), LO = 1ull << L;
    int rc;
    if ((rc = ctx->ensure_hpart())) return rc;
    // up to MLE_MAX_POINTS points share one term-table launch, one pass over f
    // and one wait
    for (size_t p0 = 0; p0 < count; p0 += MLE_MAX_POINTS) {
        const uint32_t K = (uint32_t)std::min<size_t>(MLE_MAX_POINTS, count - p0);
        Terms T;
        for (uint32_t k = 0; k < K; ++k) T.eq(points + 3 * nv * (p0 + k), nv, E3_ONE);
        if ((rc = ctx->ensure(ctx->tdesc, (T.desc_e.size() + T.coords.size() + 4) * 8))) return rc;
        if ((rc = ctx->ensure(ctx->ttab, K * (H + LO) * 24 + 256))) return rc;
        std::vector<uint64_t> all(T.desc_e);
        all.insert(all.end(), T.coords.begin(), T.coords.end());
        uint64_t* dd = (uint64_t*)ctx->tdesc.p;
        if ((rc = ctx->stage_desc(all, dd))) return rc;
        uint64_t* hi = (uint64_t*)ctx->ttab.p;
        uint64_t* lo = hi + 3 * K * H;
        term_tables_launch(dd, K, nullptr, 0, dd + T.desc_e.size(), nv, L, hi, lo, nullptr, nullptr, nullptr,
                           ctx->stream);
        const unsigned blocks = mle_eval_launch(d_f, f_is_base != 0, 1ull << nv, L, hi, lo, K, ctx->hpart_d, ctx->stream);
        if (!blocks) return bad_arg("bad MLE point count");
        HIPCHK(hipGetLastError());
        e3 v[MLE_MAX_POINTS];
        if ((rc = ctx->host_sums(blocks, (int)K, v))) return rc;
        for (uint32_t k = 0; k < K; ++k) std::memcpy(out + 3 * (p0 + k), &v[k], 24);
    }
    return 0;
}

int whir_dev_mle_evaluate(whir_ctx* ctx, const uint64_t* d_f, int f_is_base, size_t num_variables,
                          const uint64_t* point, uint64_t* out) {
    if (!ctx || !d_f || (num_variables && !point) || !out) return bad_arg("null argument");
    return whir_dev_mle_evaluate_many(ctx, d_f, f_is_base, num_variables, point, 1, out);
}

}  // extern "C"
