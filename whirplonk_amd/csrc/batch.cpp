// batch.cpp — batched commits of many independent polynomials (C ABI:
// whir_batch_*) and the LDE autotuner (whir_autotune_lde, whir_tuning_*).
//
// A batch is planned once for a list of (possibly different) sizes: one
// device region holds every polynomial's coefficients, codeword, Merkle tree
// and the roots; the commits (irs_commit::Config::commit's encode + leaves +
// tree, irs_commit.hpp:344-460, per polynomial) are spread over branch
// streams, captured once into a HIP graph, and each run is a single graph
// launch whose last node gathers all roots into one array.
#include "../../include/whir_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "ctx.h"

using namespace whir_gpu;

namespace {

size_t a256(size_t x) { return (x + 255) & ~size_t{255}; }

LdeShape commit_shape(uint64_t size, uint32_t depth, uint32_t rate) {
    LdeShape s;
    s.m = size / depth;
    s.C = depth;
    s.n = s.m << rate;
    s.row_major = true;
    return s;
}

// Times every valid split of a shape under both pass-B forms (one warm-up,
// then the median of `reps` runs with HIP events on the context's stream) and
// records the best pair, packed as log2 M2 | form << 8.
int tune_shape(whir_ctx* ctx, LdeShape s, int reps, double* best_ms, uint32_t* best_lg) {
    uint32_t cand[11];
    int nc = lde_splits(s, cand, 11);
    if (nc <= 0) return bad_arg("LDE shape not supported");
    const size_t in_words = s.m * s.C * (s.split3 ? 1 : s.in_el), out_words = s.n * s.C * s.out_el;
    uint64_t *din = nullptr, *dout = nullptr;
    uint8_t* leaves = nullptr;
    hipEvent_t e0, e1;
    HIPCHK(hipMalloc(&din, in_words * 8));
    HIPCHK(hipMalloc(&dout, out_words * 8));
    HIPCHK(hipMalloc(&leaves, s.n * 32));
    HIPCHK(hipMemsetAsync(din, 0x11, in_words * 8, ctx->stream));  // any canonical values will do
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e30;
    uint32_t arg = cand[0], arg_form = 2;
    int rc = 0;
    for (int c = 0; c < 2 * nc && !rc; ++c) {
        s.lg_m2 = cand[c / 2];
        s.rows_form = 2 - (uint32_t)(c % 2);
        std::vector<float> t;
        for (int r = 0; r <= reps && !rc; ++r) {
            (void)hipEventRecord(e0, ctx->stream);
            bool fused = false;
            rc = lde_device(ctx, s, din, dout, s.C == 16 && s.out_el == 1 && !s.split3 ? leaves : nullptr, &fused);
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r) t.push_back(ms);
        }
        if (rc) break;
        std::sort(t.begin(), t.end());
        double med = t[t.size() / 2];
        if (med < best) {
            best = med;
            arg = cand[c / 2];
            arg_form = s.rows_form;
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(din);
    (void)hipFree(dout);
    (void)hipFree(leaves);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    s.lg_m2 = 0;
    s.rows_form = 0;
    set_tuned_split(s, arg | (arg_form << 8));
    if (best_ms) *best_ms = best;
    if (best_lg) *best_lg = arg;
    return 0;
}

}  // namespace

struct whir_batch {
    whir_ctx* ctx = nullptr;
    std::vector<uint64_t> sizes;
    uint32_t depth = 16, rate = 2;
    uint32_t hash = HASH_BLAKE3;  // the context's engine when the batch was made
    std::vector<size_t> coeff_off, matrix_off, nodes_off, rows;
    size_t roots_off = 0, idx_off = 0, total = 0;
    std::vector<size_t> scratch_off, scratch_bytes;
    std::vector<std::vector<size_t>> branch_polys;
    // BLAKE3 batches: the trees are built after the encodes, level by level
    // across all polynomials (a few launches for the whole batch)
    bool multi_tree = false;
    size_t segs_off = 0;
    std::vector<MerkleMultiLaunch> tree_launches;
    // level-synchronous encodes (BLAKE3, folding factor 4, 2^10 <= m <= 2^18):
    // the polynomials in `level_polys` are encoded, leaves included, by one
    // first-pass and one second-pass launch together (lde_multi_launch); the
    // others keep their per-polynomial commits on the branch streams
    std::vector<size_t> level_polys;
    std::vector<std::vector<size_t>> level_branch_polys;  // the rest, per branch
    LdeMultiPlan lplan;
    size_t lsegs_off = 0;
    uint32_t nlseg = 0;
    const uint64_t* t1024 = nullptr;
    bool use_level = false;  // the form the graph was captured with
    double form_ms[2] = {-1.0, -1.0};  // measured replay: per-polynomial form, level-synchronous form
    uint8_t* region = nullptr;
    std::vector<hipStream_t> branches;
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~whir_batch() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (auto st : branches) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        for (auto e : events)
            if (e) (void)hipEventDestroy(e);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (region) (void)hipFree(region);
    }

    // enqueue every per-polynomial commit: branch b runs its polynomials in
    // order on `on[b]` (with the level-synchronous form, only those the
    // multi-polynomial launches do not cover)
    int enqueue(const std::vector<hipStream_t>& on, hipStream_t main) {
        hipStream_t keep = ctx->stream;
        const uint32_t keep_hash = ctx->hash;
        ctx->hash = hash;
        int rc = 0;
        const auto& per_branch = use_level ? level_branch_polys : branch_polys;
        for (size_t b = 0; b < per_branch.size() && !rc; ++b) {
            ctx->stream = on[b];
            ctx->ext_scratch = scratch_bytes[b] ? region + scratch_off[b] : nullptr;
            ctx->ext_scratch_bytes = scratch_bytes[b];
            for (size_t k : per_branch[b]) {
                rc = commit_device(ctx, (const uint64_t*)(region + coeff_off[k]), 1, sizes[k], rows[k], depth,
                                   (uint64_t*)(region + matrix_off[k]), region + nodes_off[k], nullptr, 0,
                                   !multi_tree);
                if (rc) break;
            }
        }
        ctx->stream = keep;
        ctx->hash = keep_hash;
        ctx->ext_scratch = nullptr;
        ctx->ext_scratch_bytes = 0;
        if (rc) return rc;
        (void)main;
        return 0;
    }

    // fork from `main`, run the branches, join, gather the roots on `main`
    int record(hipStream_t main) {
        const size_t nb = branch_polys.size();
        HIPCHK(hipEventRecord(events[0], main));
        for (size_t b = 0; b < nb; ++b) HIPCHK(hipStreamWaitEvent(branches[b], events[0], 0));
        int rc = enqueue(branches, main);
        if (rc) return rc;
        if (use_level) lde_multi_launch((const LdeMultiSeg*)(region + lsegs_off), nlseg, lplan, t1024, main);
        for (size_t b = 0; b < nb; ++b) {
            HIPCHK(hipEventRecord(events[1 + b], branches[b]));
            HIPCHK(hipStreamWaitEvent(main, events[1 + b], 0));
        }
        if (multi_tree) merkle_multi_launch(region, (const MerkleSeg*)(region + segs_off), tree_launches, main);
        gather_hashes_launch(region, (const uint64_t*)(region + idx_off), sizes.size(), region + roots_off, main);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // record the current form into `graph` / `exec`
    int capture() {
        const bool timing = ctx->timing;
        ctx->timing = false;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = record(ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        ctx->timing = timing;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail("hipStreamEndCapture", e);
        graph = g;
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return 0;
    }

    // median of 3 replays (after one warm-up) of the captured graph, HIP events
    int time_replays(double* ms_out) {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        std::vector<float> t;
        hipError_t e = hipSuccess;
        for (int r = 0; r < 4 && e == hipSuccess; ++r) {
            (void)hipEventRecord(e0, ctx->stream);
            e = hipGraphLaunch(exec, ctx->stream);
            (void)hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r) t.push_back(ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (e != hipSuccess) return fail("batch replay", e);
        std::sort(t.begin(), t.end());
        *ms_out = t[t.size() / 2];
        return 0;
    }
};

extern "C" {

int whir_autotune_lde(whir_ctx* ctx, size_t message_length, size_t codeword_length, size_t columns, int ext3,
                      double* out_best_ms, uint32_t* out_lg_m2) {
    if (!ctx) return bad_arg("null context");
    LdeShape s;
    s.m = message_length;
    s.n = codeword_length;
    s.C = columns * (ext3 ? 3 : 1);
    s.row_major = true;
    if (ext3) {
        s.in_el = 3;
        s.split3 = true;
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return tune_shape(ctx, s, 3, out_best_ms, out_lg_m2);
}

size_t whir_tuning_export(uint64_t* out, size_t cap) { return tuned_entries(out, cap); }

int whir_tuning_import(const uint64_t* rec, size_t words) {
    if (words % 9) return bad_arg("tuning records are 9 words each");
    for (size_t k = 0; k < words; k += 9) {
        LdeShape s;
        s.m = rec[k];
        s.n = rec[k + 1];
        s.C = rec[k + 2];
        s.in_el = (uint32_t)rec[k + 3];
        s.out_el = (uint32_t)rec[k + 4];
        s.split3 = rec[k + 5] != 0;
        s.row_major = rec[k + 6] != 0;
        s.inverse = rec[k + 7] != 0;
        uint32_t cand[11];
        int nc = lde_splits(s, cand, 11);
        const uint64_t lg = rec[k + 8] & 0xFF, form = rec[k + 8] >> 8;
        bool ok = false;
        for (int c = 0; c < nc; ++c) ok |= cand[c] == lg;
        if (!ok) return bad_arg("tuning record names an invalid split");
        if (form > 2) return bad_arg("tuning record names an invalid pass-B form");
        set_tuned_split(s, (uint32_t)(lg | (form << 8)));
    }
    return 0;
}

int whir_batch_create(whir_ctx* ctx, size_t count, const uint64_t* sizes, uint32_t log_inv_rate,
                      uint32_t folding_factor, uint32_t num_branches, int autotune, whir_batch** out) {
    return whir_batch_create_form(ctx, count, sizes, log_inv_rate, folding_factor, num_branches, autotune, -1, out);
}

int whir_batch_create_form(whir_ctx* ctx, size_t count, const uint64_t* sizes, uint32_t log_inv_rate,
                           uint32_t folding_factor, uint32_t num_branches, int autotune, int form, whir_batch** out) {
    if (!ctx || !out || !count || !sizes) return bad_arg("null argument");
    if (form < -1 || form > 1) return bad_arg("batch form must be -1, 0 or 1");
    if (num_branches == 0) num_branches = 1;
    const uint32_t depth = 1u << folding_factor;
    std::unique_ptr<whir_batch> B(new whir_batch());
    B->ctx = ctx;
    B->hash = ctx->hash;
    B->sizes.assign(sizes, sizes + count);
    B->depth = depth;
    B->rate = log_inv_rate;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    int rc;
    // shapes: validate, tune, and make the twiddle tables now (none may be
    // built while the graph is being captured)
    for (size_t k = 0; k < count; ++k) {
        uint64_t sz = sizes[k];
        if (!is_pow2(sz) || sz < depth) return bad_arg("sizes must be powers of two >= 2^folding_factor");
        LdeShape s = commit_shape(sz, depth, log_inv_rate);
        if (s.n > (1ull << 26) || s.m > (1ull << 20)) return bad_arg("codeword above 2^26 rows");
        if (autotune && !tuned_split(s)) {
            if ((rc = tune_shape(ctx, s, 3, nullptr, nullptr))) return rc;
        }
        Twiddles tw;
        if ((rc = ctx->twiddles(log2u(s.n), false, tw))) return rc;
        B->rows.push_back(s.n);
    }
    // branches: largest first onto the least loaded branch (work ~ rows)
    const size_t nb = std::min<size_t>(num_branches, count);
    std::vector<size_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return B->rows[a] > B->rows[b]; });
    B->branch_polys.assign(nb, {});
    std::vector<uint64_t> load(nb, 0);
    for (size_t k : order) {
        size_t b = std::min_element(load.begin(), load.end()) - load.begin();
        B->branch_polys[b].push_back(k);
        load[b] += B->rows[k];
    }
    for (auto& v : B->branch_polys) std::sort(v.begin(), v.end());
    // layout
    size_t off = 0;
    auto put = [&off](size_t bytes) {
        size_t o = off;
        off += a256(std::max<size_t>(bytes, 8));
        return o;
    };
    for (size_t k = 0; k < count; ++k) {
        B->coeff_off.push_back(put(sizes[k] * 8));
        B->matrix_off.push_back(put(B->rows[k] * depth * 8));
        B->nodes_off.push_back(put(whir_merkle_num_nodes(B->rows[k]) * 32));
    }
    B->roots_off = put(count * 32);
    B->idx_off = put(count * 8);
    std::vector<MerkleSeg> segs;
    if (B->hash == HASH_BLAKE3) {
        std::vector<uint64_t> noff(B->nodes_off.begin(), B->nodes_off.end());
        B->multi_tree = merkle_multi_plan(noff.data(), B->rows.data(), count, segs, B->tree_launches) == 0;
    }
    if (B->multi_tree) B->segs_off = put(segs.size() * sizeof(MerkleSeg));
    // level-synchronous encodes: BLAKE3 trees built level by level, 16-column
    // rows (folding factor 4), at least two cosets, 2^10 <= m <= 2^18;
    // segments in decreasing m (the largest tiles first)
    std::vector<LdeMultiSeg> lsegs;
    if (B->multi_tree && depth == 16 && log_inv_rate >= 1) {
        for (size_t k : order) {
            const uint64_t m = sizes[k] / depth;
            if (m >= 1024 && m <= (1ull << 18)) B->level_polys.push_back(k);
        }
    }
    for (size_t k : B->level_polys) {
        LdeShape s = commit_shape(sizes[k], depth, log_inv_rate);
        Twiddles tw;
        if ((rc = ctx->twiddles(log2u(s.n), false, tw))) return rc;
        LdeMultiSeg g{};
        g.in = (const uint64_t*)(uintptr_t)B->coeff_off[k];  // offsets for now, rebased below
        g.out = (uint64_t*)(uintptr_t)B->matrix_off[k];
        g.leaves = (uint8_t*)(uintptr_t)B->nodes_off[k];
        const size_t zw = lde_multi_scratch_elems(s.m, s.n);
        g.z = zw ? (uint64_t*)(uintptr_t)put(zw * 8) : nullptr;
        g.tw_lo = tw.lo;
        g.tw_hi = tw.hi;
        g.lg_m = (uint32_t)log2u(s.m);
        g.lg_n = (uint32_t)log2u(s.n);
        B->t1024 = tw.t1024;
        lsegs.push_back(g);
    }
    if (!lsegs.empty()) {
        B->nlseg = (uint32_t)lsegs.size();
        B->lsegs_off = put(lsegs.size() * sizeof(LdeMultiSeg));
    }
    // the per-polynomial commits left to the branches in that form
    B->level_branch_polys.assign(nb, {});
    {
        std::vector<bool> lev(count, false);
        for (size_t k : B->level_polys) lev[k] = true;
        for (size_t b = 0; b < nb; ++b)
            for (size_t k : B->branch_polys[b])
                if (!lev[k]) B->level_branch_polys[b].push_back(k);
    }
    for (size_t b = 0; b < nb; ++b) {
        size_t need = 0;
        for (size_t k : B->branch_polys[b]) {
            LdeShape s = commit_shape(sizes[k], depth, log_inv_rate);
            need = std::max(need, lde_scratch_elems_any_split(s) * 8);
        }
        B->scratch_bytes.push_back(a256(need));
        B->scratch_off.push_back(put(need));
    }
    B->total = off;
    HIPCHK(hipMalloc((void**)&B->region, B->total));
    std::vector<uint64_t> idx(count);
    for (size_t k = 0; k < count; ++k)
        idx[k] = (B->nodes_off[k] + 32 * (whir_merkle_num_nodes(B->rows[k]) - 1)) / 32;
    HIPCHK(hipMemcpyAsync(B->region + B->idx_off, idx.data(), count * 8, hipMemcpyHostToDevice, ctx->stream));
    if (B->multi_tree && !segs.empty())
        HIPCHK(hipMemcpyAsync(B->region + B->segs_off, segs.data(), segs.size() * sizeof(MerkleSeg),
                              hipMemcpyHostToDevice, ctx->stream));
    if (!lsegs.empty()) {
        for (auto& g : lsegs) {  // offsets -> device addresses
            g.in = (const uint64_t*)(B->region + (uintptr_t)g.in);
            g.out = (uint64_t*)(B->region + (uintptr_t)g.out);
            g.leaves = B->region + (uintptr_t)g.leaves;
            if (g.z) g.z = (uint64_t*)(B->region + (uintptr_t)g.z);
        }
        if (lde_multi_prepare(lsegs, B->lplan)) return bad_arg("batch shapes rejected by the level-synchronous plan");
        HIPCHK(hipMemcpyAsync(B->region + B->lsegs_off, lsegs.data(), lsegs.size() * sizeof(LdeMultiSeg),
                              hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t b = 0; b < nb; ++b) {
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        B->branches.push_back(st);
    }
    B->events.assign(nb + 1, nullptr);
    for (auto& e : B->events) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // capture the level-synchronous form where it applies; with autotune,
    // also the per-polynomial form, and keep the one whose replays are faster
    B->use_level = !B->level_polys.empty() && form != 0;
    if ((rc = B->capture())) return rc;
    if (autotune && B->use_level && form == -1) {
        if ((rc = B->time_replays(&B->form_ms[1]))) return rc;
        hipGraph_t g1 = B->graph;
        hipGraphExec_t x1 = B->exec;
        B->graph = nullptr;
        B->exec = nullptr;
        B->use_level = false;
        if ((rc = B->capture()) || (rc = B->time_replays(&B->form_ms[0]))) {
            (void)hipGraphExecDestroy(x1);
            (void)hipGraphDestroy(g1);
            return rc;
        }
        if (B->form_ms[1] <= B->form_ms[0]) {
            (void)hipGraphExecDestroy(B->exec);
            (void)hipGraphDestroy(B->graph);
            B->graph = g1;
            B->exec = x1;
            B->use_level = true;
        } else {
            (void)hipGraphExecDestroy(x1);
            (void)hipGraphDestroy(g1);
        }
    }
    *out = B.release();
    return 0;
}

int whir_batch_destroy(whir_batch* B) {
    delete B;
    return 0;
}

uint64_t* whir_batch_coeffs(whir_batch* B, size_t k) {
    return (B && k < B->sizes.size()) ? (uint64_t*)(B->region + B->coeff_off[k]) : nullptr;
}
uint64_t* whir_batch_matrix(whir_batch* B, size_t k) {
    return (B && k < B->sizes.size()) ? (uint64_t*)(B->region + B->matrix_off[k]) : nullptr;
}
uint8_t* whir_batch_nodes(whir_batch* B, size_t k) {
    return (B && k < B->sizes.size()) ? B->region + B->nodes_off[k] : nullptr;
}
size_t whir_batch_rows(whir_batch* B, size_t k) { return (B && k < B->sizes.size()) ? B->rows[k] : 0; }
size_t whir_batch_device_bytes(whir_batch* B) { return B ? B->total : 0; }

int whir_batch_form(whir_batch* B, double* per_poly_ms, double* level_ms) {
    if (!B) return -1;
    if (per_poly_ms) *per_poly_ms = B->form_ms[0];
    if (level_ms) *level_ms = B->form_ms[1];
    return B->use_level ? 1 : 0;
}

int whir_batch_run(whir_batch* B) {
    if (!B) return bad_arg("null batch");
    std::lock_guard<std::recursive_mutex> g(B->ctx->mu);
    HIPCHK(hipGraphLaunch(B->exec, B->ctx->stream));
    return 0;
}

int whir_batch_run_eager(whir_batch* B) {
    if (!B) return bad_arg("null batch");
    std::lock_guard<std::recursive_mutex> g(B->ctx->mu);
    return B->record(B->ctx->stream);
}

int whir_batch_roots(whir_batch* B, uint8_t* out_roots) {
    if (!B || !out_roots) return bad_arg("null argument");
    std::lock_guard<std::recursive_mutex> g(B->ctx->mu);
    HIPCHK(hipMemcpyAsync(out_roots, B->region + B->roots_off, B->sizes.size() * 32, hipMemcpyDeviceToHost,
                          B->ctx->stream));
    HIPCHK(hipStreamSynchronize(B->ctx->stream));
    return 0;
}

}  // extern "C"
