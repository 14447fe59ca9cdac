// sumcheck.hip — sumcheck round polynomial, fold, and Horner evaluation over
// the Goldilocks cubic extension.
//
// compute (algebra/sumcheck.hpp:43-98), half-split over n:
//   c0 = sum_{i<n/2} a_i b_i,   c2 = sum_{i<n/2} (a_{i+n/2} - a_i)(b_{i+n/2} - b_i)
// fold (sumcheck.hpp:103-116):  v'_i = v_i + (v_{i+n/2} - v_i) * r
// horner (utilities.hpp:195-246, mixed_univariate_evaluate): sum_i f_i z^i
// Field sums are exact, so any reduction order gives the reference's bits.
// Reductions: wave64 shuffles, then LDS across the workgroup's waves, then a
// one-workgroup pass over the per-workgroup partials.
#include "../../include/whir_gpu.h"

#include <mutex>

#include "ctx.h"
#include "gl.h"
#include "kernels.h"
#include "reduce_dev.h"

namespace whir_gpu {
namespace {

using gl::e3;

__device__ __forceinline__ e3 load_e3(const uint64_t* p, uint64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void store_e3(uint64_t* p, uint64_t i, e3 v) {
    p[3 * i] = v.c0;
    p[3 * i + 1] = v.c1;
    p[3 * i + 2] = v.c2;
}

using red::block_sum;
using red::store_block_totals;

__global__ __launch_bounds__(256) void sumcheck_compute_kernel(const uint64_t* __restrict__ a, int a_base,
                                                               const uint64_t* __restrict__ b, uint64_t half,
                                                               uint64_t* partials) {
    // exact sums of products per lane, reduced once
    gl::e3acc s0 = gl::e3acc_zero(), s2 = gl::e3acc_zero();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += stride) {
        e3 b0 = load_e3(b, i), b1 = load_e3(b, i + half);
        e3 db = gl::e3_sub(b1, b0);
        if (a_base) {
            uint64_t a0 = a[i], a1 = a[i + half];
            gl::e3_mac_base(s0, b0, a0);
            gl::e3_mac_base(s2, db, gl::sub(a1, a0));
        } else {
            e3 a0 = load_e3(a, i), a1 = load_e3(a, i + half);
            gl::e3_mac(s0, a0, b0);
            gl::e3_mac(s2, gl::e3_sub(a1, a0), db);
        }
    }
    e3 acc[2] = {gl::e3acc_reduce(s0), gl::e3acc_reduce(s2)};
    block_sum<2>(acc);
    store_block_totals<2>(acc, partials);
}

// One sumcheck round in a single sweep: (c0, c2) partials of (a, b) and both
// folds with the round's challenge r (sumcheck_protocol.hpp:108-134, where the
// reference computes, then folds a, then folds b: three passes over the data).
// Here r is known up front (given by the caller), so each (i, i + n/2) pair is
// read once.
// Lane i reads indices i and i + n/2 and writes index i only, so the folded
// vectors may overwrite the lower halves of Ext3 inputs (no __restrict__).
__global__ __launch_bounds__(256) void sumcheck_round_kernel(const uint64_t* a, int a_base, const uint64_t* b,
                                                             uint64_t half, e3 r, uint64_t* a_out, uint64_t* b_out,
                                                             uint64_t* partials) {
    gl::e3acc s0 = gl::e3acc_zero(), s2 = gl::e3acc_zero();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += stride) {
        e3 b0 = load_e3(b, i), b1 = load_e3(b, i + half);
        e3 db = gl::e3_sub(b1, b0);
        e3 fa;
        if (a_base) {
            uint64_t a0 = a[i], a1 = a[i + half];
            uint64_t da = gl::sub(a1, a0);
            gl::e3_mac_base(s0, b0, a0);
            gl::e3_mac_base(s2, db, da);
            fa = gl::e3_mul_base(r, da);
            fa.c0 = gl::add(fa.c0, a0);
        } else {
            e3 a0 = load_e3(a, i), a1 = load_e3(a, i + half);
            e3 da = gl::e3_sub(a1, a0);
            gl::e3_mac(s0, a0, b0);
            gl::e3_mac(s2, da, db);
            fa = gl::e3_add(a0, gl::e3_mul(r, da));
        }
        store_e3(a_out, i, fa);
        store_e3(b_out, i, gl::e3_add(b0, gl::e3_mul(r, db)));
    }
    e3 acc[2] = {gl::e3acc_reduce(s0), gl::e3acc_reduce(s2)};
    block_sum<2>(acc);
    store_block_totals<2>(acc, partials);
}

// blockIdx.y = 1 folds the second (Ext3) vector in2 into out2 in the same launch
__global__ __launch_bounds__(256) void sumcheck_fold_kernel(const uint64_t* __restrict__ in, int in_base,
                                                            uint64_t half, e3 r, uint64_t* __restrict__ out,
                                                            const uint64_t* __restrict__ in2,
                                                            uint64_t* __restrict__ out2) {
    if (blockIdx.y == 1) {
        in = in2;
        out = out2;
        in_base = 0;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += stride) {
        e3 v;
        if (in_base) {
            uint64_t lo = in[i], hi = in[i + half];
            v = gl::e3_mul_base(r, gl::sub(hi, lo));
            v.c0 = gl::add(v.c0, lo);
        } else {
            e3 lo = load_e3(in, i), hi = load_e3(in, i + half);
            v = gl::e3_add(lo, gl::e3_mul(r, gl::e3_sub(hi, lo)));
        }
        store_e3(out, i, v);
    }
}

// sum_i f_i z^i with i = t + T q, lane t < T = 2^logT (= gridDim.x * 256):
//   lane t: acc_t = sum_q f[t + T q] P[q],  P[q] = z^(T q),
// from a per-workgroup LDS table of the powers (256 of them per chunk of q;
// the chunk's sum is scaled by z^(256 T c)), so each coefficient costs one
// multiply-add without reduction and a lane's loads do not wait on one
// another.  acc_t is then scaled by z^t = z^tau z^(256 b) (t = 256 b + tau):
// z^tau = A[tau % 16] B[tau / 16] from two 16-entry tables, and z^(256 b)
// once per workgroup on its total.  The tables and z^(256 b) are products of
// the powers z^(2^l) (l < HORNER_POWS, by value) over the bits of the
// exponent: each table entry by its own lane, z^(256 b) as a product tree
// over one wave's lanes (one factor per bit of b).
template <bool EXT>
__global__ __launch_bounds__(256) void horner_kernel(const uint64_t* __restrict__ f, uint64_t n, HornerPowers zp,
                                                     uint32_t logT, uint64_t* __restrict__ partials) {
    __shared__ e3 P[256];
    __shared__ e3 A[16], B[16];
    __shared__ e3 zb;
    const uint32_t tid = threadIdx.x;
    const uint64_t T = 1ull << logT;
    const uint64_t Q = (n + T - 1) >> logT;
    const e3 one = {1, 0, 0};
    auto zpow2 = [&](uint32_t l) -> e3 { return {zp.w[3 * l], zp.w[3 * l + 1], zp.w[3 * l + 2]}; };
    if (tid < 64) {
        // wave 0: z^(256 b) = product over the set bits j of b of z^(2^(8 + j))
        e3 v = one;
#pragma unroll
        for (uint32_t j = 0; j < HORNER_MAX_LOGT - 8; ++j)
            if (tid == j && j + 8 < logT && ((blockIdx.x >> j) & 1)) v = zpow2(8 + j);
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
            const e3 o = {red::shfl_xor64(v.c0, m), red::shfl_xor64(v.c1, m), red::shfl_xor64(v.c2, m)};
            v = gl::e3_mul(v, o);
        }
        if (tid == 0) zb = v;
    } else if (tid < 96) {
        // A[k] = z^k, B[k] = z^(16 k), k < 16
        const uint32_t k = tid & 15, l0 = tid < 80 ? 0 : 4;
        e3 a = one;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
            if ((k >> j) & 1) a = gl::e3_mul(a, zpow2(l0 + j));
        if (tid < 80)
            A[k] = a;
        else
            B[k] = a;
    }
    // P[q] = z^(T q), q < min(Q, 256): lane q, product over the bits of q
    const uint32_t qmax = (uint32_t)(Q < 256 ? Q : 256);
    if (tid < qmax) {
        e3 p = one;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j)
            if ((tid >> j) & 1) p = gl::e3_mul(p, zpow2(logT + j));
        P[tid] = p;
    }
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * 256 + tid;
    e3 acc = {0, 0, 0};
    e3 zc = {1, 0, 0};  // z^(256 T c) for chunk c
    constexpr uint32_t U = EXT ? 8 : 16;  // loads in flight per lane
    for (uint64_t c0 = 0; c0 < Q; c0 += 256) {
        const uint32_t qn = (uint32_t)(Q - c0 < 256 ? Q - c0 : 256);
        gl::e3acc s = gl::e3acc_zero();
        // full groups of U coefficients (loads first), then the remainder
        uint32_t q = 0;
        for (; q + U <= qn; q += U) {
            e3 v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t i = t + T * (c0 + q + u);
                if constexpr (EXT)
                    v[u] = i < n ? load_e3(f, i) : e3{0, 0, 0};
                else
                    v[u] = {i < n ? f[i] : 0, 0, 0};
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                if constexpr (EXT)
                    gl::e3_mac(s, P[q + u], v[u]);
                else
                    gl::e3_mac_base(s, P[q + u], v[u].c0);
            }
        }
        for (; q < qn; ++q) {
            const uint64_t i = t + T * (c0 + q);
            if (i < n) {
                if constexpr (EXT)
                    gl::e3_mac(s, P[q], load_e3(f, i));
                else
                    gl::e3_mac_base(s, P[q], f[i]);
            }
        }
        e3 part = gl::e3acc_reduce(s);
        if (c0) part = gl::e3_mul(part, zc);
        acc = gl::e3_add(acc, part);
        if (c0 + 256 < Q) zc = gl::e3_mul(zc, zpow2(logT + 8));
    }
    acc = gl::e3_mul(acc, gl::e3_mul(A[tid & 15], B[tid >> 4]));
    e3 a1[1] = {acc};
    block_sum<1>(a1);
    if (tid == 0) store_e3(partials, blockIdx.x, gl::e3_mul(a1[0], zb));
}

unsigned grid_for(uint64_t work, unsigned cap) {
    uint64_t b = (work + 255) / 256;
    if (b < 1) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

}  // namespace

unsigned sumcheck_compute_launch(const uint64_t* a, bool a_base, const uint64_t* b, uint64_t n, uint64_t* partials,
                                 hipStream_t st, bool device_sum) {
    unsigned g = grid_for(n / 2, SUMCHECK_MAX_BLOCKS);
    hipLaunchKernelGGL(sumcheck_compute_kernel, dim3(g), dim3(256), 0, st, a, a_base ? 1 : 0, b, n / 2, partials);
    if (device_sum) hipLaunchKernelGGL(red::sum_blocks_kernel<2>, dim3(1), dim3(256), 0, st, partials, g, partials);
    return g;
}

unsigned sumcheck_round_launch(const uint64_t* a, bool a_base, const uint64_t* b, uint64_t n, const uint64_t r[3],
                               uint64_t* a_out, uint64_t* b_out, uint64_t* partials, hipStream_t st,
                               bool device_sum) {
    unsigned g = grid_for(n / 2, SUMCHECK_MAX_BLOCKS);
    e3 rr = {r[0], r[1], r[2]};
    hipLaunchKernelGGL(sumcheck_round_kernel, dim3(g), dim3(256), 0, st, a, a_base ? 1 : 0, b, n / 2, rr, a_out, b_out,
                       partials);
    if (device_sum) hipLaunchKernelGGL(red::sum_blocks_kernel<2>, dim3(1), dim3(256), 0, st, partials, g, partials);
    return g;
}

void sumcheck_fold_launch(const uint64_t* in, bool in_base, uint64_t n, const uint64_t r[3], uint64_t* out,
                          hipStream_t st) {
    unsigned g = grid_for(n / 2, 8192);
    e3 rr = {r[0], r[1], r[2]};
    hipLaunchKernelGGL(sumcheck_fold_kernel, dim3(g), dim3(256), 0, st, in, in_base ? 1 : 0, n / 2, rr, out,
                       (const uint64_t*)nullptr, (uint64_t*)nullptr);
}

void sumcheck_fold2_launch(const uint64_t* a, bool a_base, uint64_t* a_out, const uint64_t* b, uint64_t* b_out,
                           uint64_t n, const uint64_t r[3], hipStream_t st) {
    unsigned g = grid_for(n / 2, 8192);
    e3 rr = {r[0], r[1], r[2]};
    hipLaunchKernelGGL(sumcheck_fold_kernel, dim3(g, 2), dim3(256), 0, st, a, a_base ? 1 : 0, n / 2, rr, a_out, b,
                       b_out);
}

unsigned horner_launch(const uint64_t* f, bool f_ext3, uint64_t n, const HornerPowers& zp, uint64_t* partials,
                       hipStream_t st) {
    // T = blocks*256 threads, a power of two, about 16+ terms per thread
    // (logT + 8 < HORNER_POWS: the kernel reads z^(2^l) up to l = logT + 8)
    uint32_t logT = 8;
    while (logT < HORNER_MAX_LOGT && (1ull << (logT + 4)) < n) ++logT;
    unsigned blocks = (unsigned)((1ull << logT) / 256);
    if (f_ext3)
        hipLaunchKernelGGL(horner_kernel<true>, dim3(blocks), dim3(256), 0, st, f, n, zp, logT, partials);
    else
        hipLaunchKernelGGL(horner_kernel<false>, dim3(blocks), dim3(256), 0, st, f, n, zp, logT, partials);
    return blocks;
}

}  // namespace whir_gpu

// ------------------------------------------- the allocation-free sumcheck round
// whir_dev_sumcheck_round_noalloc: the per-workgroup (c0, c2) partials live
// in one file-scope device array per GPU instead of a buffer the library
// allocates.  Host code never uses the array's name as an address (it is a
// host-side stand-in there): its device address is looked up once per device
// with hipGetSymbolAddress at context creation and kept in g_scratch_ptr.
namespace whir_gpu {
namespace {
constexpr size_t NOALLOC_SCRATCH_U64 = 6 * (size_t)SUMCHECK_MAX_BLOCKS;
static_assert(NOALLOC_SCRATCH_U64 * 8 == WHIR_NOALLOC_SCRATCH_BYTES, "the header's stated scratch size");
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_U64];
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return bad_arg("device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // `device` is current (whir_ctx_create)
    if (e != hipSuccess || !p) return fail("hipGetSymbolAddress(scratch)", e != hipSuccess ? e : hipErrorUnknown);
    g_scratch_ptr[device] = (uint64_t*)p;
    return 0;
}

}  // namespace whir_gpu

extern "C" int whir_dev_sumcheck_round_noalloc(whir_ctx* ctx, const uint64_t* d_a, int a_is_base, const uint64_t* d_b,
                                               size_t n, const uint64_t* r, uint64_t* d_a_out, uint64_t* d_b_out,
                                               uint64_t* d_c0c2) {
    using namespace whir_gpu;
    if (!ctx || !d_a || !d_b || !r || !d_a_out || !d_b_out || !d_c0c2) return bad_arg("null argument");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (n % 2) return bad_arg("sumcheck length must be even");
    if (n > WHIR_NOALLOC_SUMCHECK_MAX_N) return bad_arg("sumcheck length above 2^32 (the no-allocation maximum)");
    if (a_is_base && (const void*)d_a == (const void*)d_a_out) return bad_arg("base-field fold cannot run in place");
    uint64_t* sc = nullptr;
    {
        std::lock_guard<std::mutex> s(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) sc = g_scratch_ptr[ctx->device];
    }
    if (!sc) return bad_arg("no-allocation scratch not set up for this context's device");
    const unsigned blocks = grid_for(n / 2, SUMCHECK_MAX_BLOCKS);
    if ((size_t)6 * blocks > NOALLOC_SCRATCH_U64) return bad_arg("sumcheck partials exceed the no-allocation scratch");
    // every workgroup overwrites its own partials, so nothing needs clearing
    const e3 rr = {r[0], r[1], r[2]};
    hipLaunchKernelGGL(sumcheck_round_kernel, dim3(blocks), dim3(256), 0, ctx->stream, d_a, a_is_base ? 1 : 0, d_b,
                       (uint64_t)(n / 2), rr, d_a_out, d_b_out, sc);
    hipLaunchKernelGGL(red::sum_blocks_kernel<2>, dim3(1), dim3(256), 0, ctx->stream, (const uint64_t*)sc, blocks,
                       d_c0c2);
    HIPCHK(hipGetLastError());
    return 0;
}
