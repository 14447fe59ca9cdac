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

    // enqueue every commit: branch b runs its polynomials in order on `on[b]`
    int enqueue(const std::vector<hipStream_t>& on, hipStream_t main) {
        hipStream_t keep = ctx->stream;
        const uint32_t keep_hash = ctx->hash;
        ctx->hash = hash;
        int rc = 0;
        for (size_t b = 0; b < branch_polys.size() && !rc; ++b) {
            ctx->stream = on[b];
            ctx->ext_scratch = scratch_bytes[b] ? region + scratch_off[b] : nullptr;
            ctx->ext_scratch_bytes = scratch_bytes[b];
            for (size_t k : branch_polys[b]) {
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
        for (size_t b = 0; b < nb; ++b) {
            HIPCHK(hipEventRecord(events[1 + b], branches[b]));
            HIPCHK(hipStreamWaitEvent(main, events[1 + b], 0));
        }
        if (multi_tree) merkle_multi_launch(region, (const MerkleSeg*)(region + segs_off), tree_launches, main);
        gather_hashes_launch(region, (const uint64_t*)(region + idx_off), sizes.size(), region + roots_off, main);
        HIPCHK(hipGetLastError());
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
    if (!ctx || !out || !count || !sizes) return bad_arg("null argument");
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
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t b = 0; b < nb; ++b) {
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        B->branches.push_back(st);
    }
    B->events.assign(nb + 1, nullptr);
    for (auto& e : B->events) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // capture
    const bool timing = ctx->timing;
    ctx->timing = false;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = B->record(ctx->stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    ctx->timing = timing;
    if (rc) return rc;
    if (e != hipSuccess) return fail("hipStreamEndCapture", e);
    B->graph = graph;
    HIPCHK(hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0));
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
