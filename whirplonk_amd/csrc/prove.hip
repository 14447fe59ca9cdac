// prove.hip — kernels of the WHIR prove loop beyond commit and plain sumcheck.
//
//  * pow_search: the smallest nonce n with LE64(BLAKE3(ch || LE64(n) || 0^24))
//    <= threshold (protocols/proof_of_work.hpp:95-193).  Each launch covers one
//    window of nonces and keeps the minimum hit (a lane stops once the best
//    hit is below its next nonce), so windows taken in increasing order return
//    the global minimum, as the reference's scan does.
//  * term tables + term_accumulate: acc[i] (+)= sum_t s_t * w_t(i) for terms
//      eq terms   w(i) = prod_k (bit_k(i) ? x_k : 1 - x_k), x_0 the top bit
//                 (eval_eq, algebra/multilinear.hpp:120-133)
//      power terms w(i) = z^i (UnivariateEvaluation::accumulate /
//                 geometric_accumulate, algebra/utilities.hpp:260-307)
//    Every term factors over i = h * 2^L + l as hi_t[h] * lo_t[l]; a small
//    kernel writes the two tables and the sweep does one multiply-add per
//    term and entry (an Ext3 x base multiply when z is in the base field).
//  * mle_evaluate: sum_i f_i eq(x, i) from the same tables
//    (mixed_multilinear_extend, algebra/multilinear.hpp:41-103).
//  * fold_compute: one sumcheck fold of (a, b) by r fused with the next round
//    polynomial of the folded pair (sumcheck.hpp:43-116).
//  * gather_rows: opened codeword rows (irs_commit.hpp:656-668).
//  * rlc_combine: vector = f_0 + sum_k c_k f_k lifted to Ext3
//    (lift / mixed_scalar_mul_add, utilities.hpp:152-181).
#include "blake3_dev.h"
#include "sha256_dev.h"
#include "gl.h"
#include "kernels.h"
#include "reduce_dev.h"

#include <algorithm>
#include <type_traits>

namespace whir_gpu {
namespace {

using gl::e3;

__device__ __forceinline__ e3 ld3(const uint64_t* p, uint64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(uint64_t* p, uint64_t i, e3 v) {
    p[3 * i] = v.c0;
    p[3 * i + 1] = v.c1;
    p[3 * i + 2] = v.c2;
}

using red::block_sum;
using red::store_block_totals;

// ------------------------------------------------------------------ PoW
// Up to POW_BATCH_MAX independent searches in one launch (a single search is
// a batch of one).  Job k owns blocks [block0_k, block0_{k+1}); its lanes
// walk nonces base_k + lane, + lanes_k, ... in increasing order.  best[k]
// (device) is job k's running minimum hit, which lets lanes and whole blocks
// stop early; each block also writes the minimum of its own hits (or ~0) to
// block_min[blockIdx.x], and the host takes the minimum over a job's blocks.
template <uint32_t HASH>
__global__ __launch_bounds__(256) void pow_batch_kernel(PowBatchArgs a, unsigned long long* best,
                                                        unsigned long long* reset, uint64_t* block_min) {
    __shared__ unsigned long long smin;
    __shared__ int skip;
    // the other set of early-exit slots is cleared for the next launch
    if (blockIdx.x == 0 && threadIdx.x < POW_BATCH_MAX) reset[threadIdx.x] = ~0ull;
    // this block's job (block-uniform; constant indices keep the arguments in
    // scalar registers)
    uint32_t jb = 0;
#pragma unroll
    for (uint32_t k = 1; k < POW_BATCH_MAX; ++k)
        if (k < a.njobs && blockIdx.x >= a.job[k].block0) jb = k;
    PowJobDesc J = a.job[0];
    uint32_t next0 = a.total_blocks;
#pragma unroll
    for (uint32_t k = 0; k < POW_BATCH_MAX; ++k) {
        if (k == jb) J = a.job[k];
        if (k == jb + 1 && k < a.njobs) next0 = a.job[k].block0;
    }
    unsigned long long* bj = best + jb;
    const uint64_t lanes = (uint64_t)(next0 - J.block0) * blockDim.x;
    const uint64_t first = J.base + (uint64_t)(blockIdx.x - J.block0) * blockDim.x;
    if (threadIdx.x == 0) {
        smin = ~0ull;
        // a block dispatched after a hit below its first nonce has nothing to
        // add: with more blocks than the GPU holds at once, the later ones
        // mostly return here
        skip = __hip_atomic_load(bj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < first;
    }
    __syncthreads();
    if (skip) {
        if (threadIdx.x == 0) block_min[blockIdx.x] = ~0ull;
        return;
    }
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) m[i] = J.ch[i];
#pragma unroll
    for (int i = 10; i < 16; ++i) m[i] = 0;
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = b3::iv(i);
    uint64_t nonce = first + threadIdx.x;
    for (uint32_t k = 0; k < J.per_thread; ++k, nonce += lanes) {
        // once a hit below the lane's next nonce is known, the lane cannot improve it
        if (k && nonce > __hip_atomic_load(bj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        m[8] = (uint32_t)nonce;
        m[9] = (uint32_t)(nonce >> 32);
        uint32_t out[8];  // the digest's bytes, little-endian words
        if constexpr (HASH == HASH_SHA2)
            s256::hash64(m, out);
        else
            b3::compress(cv, m, 0, 64, b3::CHUNK_START | b3::CHUNK_END | b3::ROOT, out);
        uint64_t h = ((uint64_t)out[1] << 32) | out[0];
        if (h <= J.threshold) {
            atomicMin(bj, (unsigned long long)nonce);
            atomicMin(&smin, (unsigned long long)nonce);
            break;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) block_min[blockIdx.x] = smin;
}

// ------------------------------------------------------------ term tables
// Term descriptor (TERM_WORDS u64): kind, s (3), z or unused (3), Z = z^(2^L)
// (3), offset of the eq coordinates in `coords` (in Ext3 elements).
enum : uint64_t { TERM_EQ = 0, TERM_POW = 1 };

// Limb form of a hi entry x (Ext3) for term_sweep_kernel: TERM_LIMB_WORDS
// u32, the 22/22/20-bit limbs of x.c0, x.c1, x.c2, 2 x.c1, 2 x.c2 (the
// doubled components are the factors X^3 = 2 brings into an Ext3 product),
// padded to one 64-byte scalar load.
constexpr uint32_t TERM_LIMB_WORDS = 16;

// hi tables: H entries per term (Ext3); lo tables: 2^L entries per term, Ext3
// for eq terms and Ext3 powers, u64 for base-field powers.  Both term groups
// in one launch: blockIdx.y < ne is term y of the Ext3 group (desc_e, hi_e,
// lo_e), the rest the base-field power group (desc_b, hi_b, lo_b).
__global__ __launch_bounds__(256) void term_tables_kernel(const uint64_t* __restrict__ desc_e,
                                                          const uint64_t* __restrict__ desc_b,
                                                          const uint64_t* __restrict__ coords, uint32_t nv,
                                                          uint32_t L, uint32_t ne, uint64_t* __restrict__ hi_e,
                                                          uint64_t* __restrict__ lo_e, uint64_t* __restrict__ hi_b,
                                                          uint64_t* __restrict__ lo_b, uint32_t* __restrict__ hl) {
    const int lo_base = blockIdx.y >= ne;
    const uint32_t t = lo_base ? blockIdx.y - ne : blockIdx.y;
    const uint64_t* d = (lo_base ? desc_b : desc_e) + (uint64_t)TERM_WORDS * t;
    uint64_t* hi = lo_base ? hi_b : hi_e;
    uint64_t* lo = lo_base ? lo_b : lo_e;
    const uint64_t H = 1ull << (nv - L), LO = 1ull << L;
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const e3 s = {d[1], d[2], d[3]};
    const e3 one = {1, 0, 0};
    // limb form of hi entry (term y of the launch, row j) for the sweep
    auto limbs = [&](e3 x) {
        if (!hl) return;
        uint32_t* o = hl + TERM_LIMB_WORDS * ((uint64_t)blockIdx.y * H + j);
        const uint64_t v[5] = {x.c0, x.c1, x.c2, gl::dbl(x.c1), gl::dbl(x.c2)};
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            o[3 * k] = (uint32_t)v[k] & 0x3FFFFF;
            o[3 * k + 1] = (uint32_t)(v[k] >> 22) & 0x3FFFFF;
            o[3 * k + 2] = (uint32_t)(v[k] >> 44);
        }
        o[15] = 0;
    };
    if (d[0] == TERM_EQ) {
        const uint64_t* x = coords + 3 * d[10];
        if (j < H) {
            // top nv - L coordinates; coordinate k pairs with bit (nv - L - 1 - k) of h
            e3 acc = s;
            for (uint32_t k = 0; k < nv - L; ++k) {
                e3 xk = ld3(x, k);
                bool bit = (j >> (nv - L - 1 - k)) & 1;
                acc = gl::e3_mul(acc, bit ? xk : gl::e3_sub(one, xk));
            }
            st3(hi, (uint64_t)t * H + j, acc);
            limbs(acc);
        }
        if (j < LO) {
            e3 acc = one;
            for (uint32_t k = 0; k < L; ++k) {
                e3 xk = ld3(x, nv - L + k);
                bool bit = (j >> (L - 1 - k)) & 1;
                acc = gl::e3_mul(acc, bit ? xk : gl::e3_sub(one, xk));
            }
            st3(lo, (uint64_t)t * LO + j, acc);
        }
    } else {
        const e3 z = {d[4], d[5], d[6]}, Z = {d[7], d[8], d[9]};
        if (j < H) {
            const e3 x = gl::e3_mul(s, gl::e3_pow(Z, j));
            st3(hi, (uint64_t)t * H + j, x);
            limbs(x);
        }
        if (j < LO) {
            if (lo_base) {
                lo[(uint64_t)t * LO + j] = gl::pow(z.c0, j);
            } else {
                st3(lo, (uint64_t)t * LO + j, gl::e3_pow(z, j));
            }
        }
    }
}

// acc[i] (+)= sum over Ext3-lo terms of hi*lo + sum over base-lo terms of hi*lo
// (2^L < 256: one lane per entry).
__global__ __launch_bounds__(256) void term_accumulate_kernel(uint64_t* __restrict__ acc, uint64_t n, uint32_t L,
                                                              const uint64_t* __restrict__ hi_e,
                                                              const uint64_t* __restrict__ lo_e, uint32_t te,
                                                              const uint64_t* __restrict__ hi_b,
                                                              const uint64_t* __restrict__ lo_b, uint32_t tb,
                                                              int accumulate) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t H = n >> L, LO = 1ull << L;
    const uint64_t h = i >> L, l = i & (LO - 1);
    // exact sum of all term products, reduced once
    gl::e3acc s = gl::e3acc_zero();
    for (uint32_t t = 0; t < te; ++t) gl::e3_mac(s, ld3(hi_e, t * H + h), ld3(lo_e, t * LO + l));
    for (uint32_t t = 0; t < tb; ++t) gl::e3_mac_base(s, ld3(hi_b, t * H + h), lo_b[t * LO + l]);
    e3 sum = gl::e3acc_reduce(s);
    if (accumulate) sum = gl::e3_add(sum, ld3(acc, i));
    st3(acc, i, sum);
}

// x * 2^(22 i + 32 j) summed over the six limb-product accumulators A[2 i + j]
// (any values below 2^64), reduced mod p
__device__ __forceinline__ uint64_t limb_fold(const uint64_t* A) {
    uint64_t v = gl::canon(A[0]);
    v = gl::add(v, gl::mul_pow2<32>(A[1]));
    v = gl::add(v, gl::mul_pow2<22>(A[2]));
    v = gl::add(v, gl::mul_pow2<54>(A[3]));
    v = gl::add(v, gl::mul_pow2<44>(A[4]));
    v = gl::add(v, gl::mul_pow2<76>(A[5]));
    return v;
}

// The same sweep for 2^L >= 256.  Workgroup w covers l = 256 (w % (2^L / 256))
// + tid of row h = w / (2^L / 256), so every hi entry it reads is the same for
// all its lanes (workgroup-uniform addresses: scalar loads).  Terms go in
// groups of TERM_GROUP whose lo loads are all issued before their
// multiply-adds, so a lane waits once per group instead of once per term.
// No per-term 128-bit products: each (uniform) hi component x is split into
// limbs x_0 + 2^22 x_1 + 2^44 x_2 (22, 22, 20 bits) and the lane's lo value y
// into 32-bit halves y_0, y_1; the six products x_i y_j < 2^54 add into 64-bit
// accumulators A_ij with one mad each and no carries, and
// sum_ij A_ij 2^(22 i + 32 j) is reduced mod p once, or once every
// TERM_LIMB_PASS products per accumulator: a base-field term (one per STIR
// query, usually most of the terms) adds one, an Ext3 term three (Ext3 x Ext3
// with X^3 = 2, the doubled x components computed once per term).
constexpr uint32_t TERM_GROUP = 4;
constexpr uint32_t TERM_LIMB_PASS = 1024;  // 1024 * (2^22 - 1) * (2^32 - 1) < 2^64

struct LimbAcc {
    uint64_t a[3][6];  // component, slot 2 i + j
};

// A += x * y for x given by its limbs xl[0..3), y any u64
__device__ __forceinline__ void limb_mac(uint64_t* A, const uint32_t* xl, uint64_t y) {
    const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
    A[0] += (uint64_t)xl[0] * y0;
    A[1] += (uint64_t)xl[0] * y1;
    A[2] += (uint64_t)xl[1] * y0;
    A[3] += (uint64_t)xl[1] * y1;
    A[4] += (uint64_t)xl[2] * y0;
    A[5] += (uint64_t)xl[2] * y1;
}

// acc += x * y, x Ext3 in limb form (TERM_LIMB_WORDS, uniform), y Ext3
__device__ __forceinline__ void limb_mac_e3(LimbAcc& A, const uint32_t* x, e3 y) {
    limb_mac(A.a[0], x + 0, y.c0);   // x0 y0
    limb_mac(A.a[0], x + 12, y.c1);  // 2 x2 y1
    limb_mac(A.a[0], x + 9, y.c2);   // 2 x1 y2
    limb_mac(A.a[1], x + 3, y.c0);   // x1 y0
    limb_mac(A.a[1], x + 0, y.c1);   // x0 y1
    limb_mac(A.a[1], x + 12, y.c2);  // 2 x2 y2
    limb_mac(A.a[2], x + 6, y.c0);   // x2 y0
    limb_mac(A.a[2], x + 3, y.c1);   // x1 y1
    limb_mac(A.a[2], x + 0, y.c2);   // x0 y2
}

// acc += x * y, x Ext3 in limb form (uniform), y base
__device__ __forceinline__ void limb_mac_base(LimbAcc& A, const uint32_t* x, uint64_t y) {
    limb_mac(A.a[0], x + 0, y);
    limb_mac(A.a[1], x + 3, y);
    limb_mac(A.a[2], x + 6, y);
}

// sum_ij A_ij 2^(22 i + 32 j) mod p of one component: the six shifted
// accumulators summed exactly in 192 bits (acc128), reduced once
__device__ __forceinline__ uint64_t limb_reduce(const uint64_t* a) {
    gl::acc128 s = {a[0], 0, 0};
    gl::acc_add(s, a[1] << 32, a[1] >> 32);
    gl::acc_add(s, a[2] << 22, a[2] >> 42);
    gl::acc_add(s, a[3] << 54, a[3] >> 10);
    gl::acc_add(s, a[4] << 44, a[4] >> 20);
    gl::acc_add(s, 0, a[5] << 12);
    s.top += a[5] >> 52;
    return gl::acc_reduce(s);
}

// the three components' sums mod p; clears A
__device__ __forceinline__ e3 limb_flush(LimbAcc& A) {
    const e3 r = {limb_reduce(A.a[0]), limb_reduce(A.a[1]), limb_reduce(A.a[2])};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 6; ++k) A.a[c][k] = 0;
    return r;
}

__global__ __launch_bounds__(256) void term_sweep_kernel(uint64_t* __restrict__ acc, uint64_t H, uint32_t L,
                                                         const uint32_t* __restrict__ hl,
                                                         const uint64_t* __restrict__ lo_e, uint32_t te,
                                                         const uint64_t* __restrict__ lo_b, uint32_t tb,
                                                         int accumulate) {
    const uint64_t LO = 1ull << L;
    const uint32_t lblocks = (uint32_t)(LO >> 8);
    const uint64_t l = (uint64_t)(blockIdx.x % lblocks) * 256 + threadIdx.x;
    const uint64_t h = blockIdx.x / lblocks;
    LimbAcc A;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 6; ++k) A.a[c][k] = 0;
    e3 sum = {0, 0, 0};
    // products added to each accumulator since the last flush (uniform)
    uint32_t used = 0;
    auto reserve = [&](uint32_t k) {
        if (used + k > TERM_LIMB_PASS) {
            sum = gl::e3_add(sum, limb_flush(A));
            used = 0;
        }
        used += k;
    };
    // Ext3 terms (few: eq constraints and out-of-domain samples)
    uint32_t t = 0;
    for (; t + TERM_GROUP <= te; t += TERM_GROUP) {
        e3 lv[TERM_GROUP];
#pragma unroll
        for (uint32_t u = 0; u < TERM_GROUP; ++u) lv[u] = ld3(lo_e, (uint64_t)(t + u) * LO + l);
        reserve(3 * TERM_GROUP);
#pragma unroll
        for (uint32_t u = 0; u < TERM_GROUP; ++u) limb_mac_e3(A, hl + TERM_LIMB_WORDS * ((uint64_t)(t + u) * H + h), lv[u]);
    }
    for (; t < te; ++t) {
        reserve(3);
        limb_mac_e3(A, hl + TERM_LIMB_WORDS * ((uint64_t)t * H + h), ld3(lo_e, (uint64_t)t * LO + l));
    }
    // base-field terms (their limb rows follow the Ext3 terms'): the next
    // group's lo loads are in flight while this group's products run
    const uint32_t* hb = hl + TERM_LIMB_WORDS * ((uint64_t)te * H);
    const uint32_t full = tb / TERM_GROUP * TERM_GROUP;
    uint64_t cur[TERM_GROUP];
    if (full) {
#pragma unroll
        for (uint32_t u = 0; u < TERM_GROUP; ++u) cur[u] = lo_b[(uint64_t)u * LO + l];
    }
    for (t = 0; t < full; t += TERM_GROUP) {
        const uint32_t tn = t + TERM_GROUP < full ? t + TERM_GROUP : t;
        uint64_t nxt[TERM_GROUP];
#pragma unroll
        for (uint32_t u = 0; u < TERM_GROUP; ++u) nxt[u] = lo_b[(uint64_t)(tn + u) * LO + l];
        reserve(TERM_GROUP);
#pragma unroll
        for (uint32_t u = 0; u < TERM_GROUP; ++u) limb_mac_base(A, hb + TERM_LIMB_WORDS * ((uint64_t)(t + u) * H + h), cur[u]);
#pragma unroll
        for (uint32_t u = 0; u < TERM_GROUP; ++u) cur[u] = nxt[u];
    }
    for (t = full; t < tb; ++t) {
        reserve(1);
        limb_mac_base(A, hb + TERM_LIMB_WORDS * ((uint64_t)t * H + h), lo_b[(uint64_t)t * LO + l]);
    }
    if (used) sum = gl::e3_add(sum, limb_flush(A));
    const uint64_t i = h * LO + l;
    if (accumulate) sum = gl::e3_add(sum, ld3(acc, i));
    st3(acc, i, sum);
}

// For each of K points: the sum over a lane's indices i = b + threadIdx.x
// (b = blockIdx.x * 256 + k * stride) of f_i * hi_p[i >> L], accumulated
// without reduction; every load of f serves all K points, and the loads of U
// iterations are issued before their products.  ROW: 2^L >= 256, so the
// workgroup's 256 indices share one row and its hi entries are uniform loads.
template <bool F_BASE, bool ROW, int K>
__device__ __forceinline__ void mle_sweep(const uint64_t* __restrict__ f, uint64_t n, uint32_t L,
                                          const uint64_t* __restrict__ hi, uint64_t H, gl::e3acc* s) {
    using V = typename std::conditional<F_BASE, uint64_t, e3>::type;
    constexpr int U = std::max(1, (F_BASE ? 8 : 4) / K);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, t = threadIdx.x;
    auto ld = [&](uint64_t i) -> V {
        if constexpr (F_BASE)
            return f[i];
        else
            return ld3(f, i);
    };
    auto mac = [&](gl::e3acc& acc, const e3& w, const V& v) {
        if constexpr (F_BASE)
            gl::e3_mac_base(acc, w, v);
        else
            gl::e3_mac(acc, w, v);
    };
    // all points' products for the U iterations from index base b
    auto step = [&](uint64_t b, int u) {
        V v[U];
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (k < u) v[k] = ld(b + k * stride + t);
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (k < u) {
                const uint64_t row = (ROW ? b + k * stride : b + k * stride + t) >> L;
#pragma unroll
                for (int p = 0; p < K; ++p) mac(s[p], ld3(hi, (uint64_t)p * H + row), v[k]);
            }
    };
#pragma unroll
    for (int p = 0; p < K; ++p) s[p] = gl::e3acc_zero();
    uint64_t b = (uint64_t)blockIdx.x * blockDim.x;
    for (; b + (U - 1) * stride + t < n; b += U * stride) step(b, U);
    for (; b + t < n; b += stride) step(b, 1);
}

// For each of K points p: partial sums of f_i * hi_p[i >> L] * lo_p[i & mask]
// (one eq term per point; point p's tables at hi + 3 H p and lo + 3 2^L p),
// as sum_l lo_p[l] * (sum_h f[h 2^L + l] hi_p[h]): the stride (a multiple of
// 2^L: mle_eval_launch keeps the grid a multiple of 2^L / 256 when 2^L > 256)
// leaves a lane's column l = i & mask fixed across its iterations, so the
// lane accumulates f_i hi_p[h] without reduction and multiplies its total by
// lo_p[l] once.  Block totals: partials[3 (K b + p) ..].
template <int K>
__global__ __launch_bounds__(256) void mle_eval_kernel(const uint64_t* __restrict__ f, int f_base, uint64_t n,
                                                       uint32_t L, const uint64_t* __restrict__ hi,
                                                       const uint64_t* __restrict__ lo, uint64_t* partials) {
    const bool row = L >= 8;
    const uint64_t H = n >> L, LO = 1ull << L;
    gl::e3acc s[K];
    if (f_base) {
        if (row)
            mle_sweep<true, true, K>(f, n, L, hi, H, s);
        else
            mle_sweep<true, false, K>(f, n, L, hi, H, s);
    } else {
        if (row)
            mle_sweep<false, true, K>(f, n, L, hi, H, s);
        else
            mle_sweep<false, false, K>(f, n, L, hi, H, s);
    }
    const uint64_t l = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & (LO - 1);
    e3 acc[K];
#pragma unroll
    for (int p = 0; p < K; ++p) acc[p] = gl::e3_mul(gl::e3acc_reduce(s[p]), ld3(lo, (uint64_t)p * LO + l));
    block_sum<K>(acc);
    store_block_totals<K>(acc, partials);
}

// ------------------------------------------------------------ fold+compute
// (a, b) of length 4q folded by r into (a', b') of length 2q, plus
//   c0 = sum_{i<q} a'_i b'_i,  c2 = sum_{i<q} (a'_{i+q} - a'_i)(b'_{i+q} - b'_i).
// Lane i reads indices i, i+q, i+2q, i+3q and writes i, i+q only, so the
// outputs may overwrite Ext3 inputs in place.
__global__ __launch_bounds__(256) void fold_compute_kernel(const uint64_t* a, int a_base, const uint64_t* b,
                                                           uint64_t q, e3 r, uint64_t* a_out, uint64_t* b_out,
                                                           uint64_t* partials) {
    gl::e3acc s0 = gl::e3acc_zero(), s2 = gl::e3acc_zero();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q; i += stride) {
        e3 fa0, fa1;
        if (a_base) {
            uint64_t x0 = a[i], x1 = a[i + q], x2 = a[i + 2 * q], x3 = a[i + 3 * q];
            fa0 = gl::e3_mul_base(r, gl::sub(x2, x0));
            fa0.c0 = gl::add(fa0.c0, x0);
            fa1 = gl::e3_mul_base(r, gl::sub(x3, x1));
            fa1.c0 = gl::add(fa1.c0, x1);
        } else {
            e3 x0 = ld3(a, i), x1 = ld3(a, i + q), x2 = ld3(a, i + 2 * q), x3 = ld3(a, i + 3 * q);
            fa0 = gl::e3_add(x0, gl::e3_mul(r, gl::e3_sub(x2, x0)));
            fa1 = gl::e3_add(x1, gl::e3_mul(r, gl::e3_sub(x3, x1)));
        }
        e3 y0 = ld3(b, i), y1 = ld3(b, i + q), y2 = ld3(b, i + 2 * q), y3 = ld3(b, i + 3 * q);
        e3 fb0 = gl::e3_add(y0, gl::e3_mul(r, gl::e3_sub(y2, y0)));
        e3 fb1 = gl::e3_add(y1, gl::e3_mul(r, gl::e3_sub(y3, y1)));
        st3(a_out, i, fa0);
        st3(a_out, i + q, fa1);
        st3(b_out, i, fb0);
        st3(b_out, i + q, fb1);
        gl::e3_mac(s0, fa0, fb0);
        gl::e3_mac(s2, gl::e3_sub(fa1, fa0), gl::e3_sub(fb1, fb0));
    }
    e3 acc[2] = {gl::e3acc_reduce(s0), gl::e3acc_reduce(s2)};
    block_sum<2>(acc);
    store_block_totals<2>(acc, partials);
}

// ------------------------------------------------------------ row gather
__global__ __launch_bounds__(256) void gather_rows_kernel(const uint64_t* __restrict__ matrix, uint64_t row_u64,
                                                          const uint64_t* __restrict__ idx, uint64_t count,
                                                          uint64_t* __restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count * row_u64) return;
    const uint64_t r = k / row_u64, c = k - r * row_u64;
    out[k] = matrix[idx[r] * row_u64 + c];
}

// An opening in one launch: y = 0 gathers the rows (as gather_rows_kernel),
// y = 1 the 32-byte node hashes (16 bytes per lane).
__global__ __launch_bounds__(256) void gather_open_kernel(const uint64_t* __restrict__ matrix, uint64_t row_u64,
                                                          const uint64_t* __restrict__ ridx, uint64_t count,
                                                          uint64_t* __restrict__ rows, const uint8_t* __restrict__ nodes,
                                                          const uint64_t* __restrict__ hidx, uint64_t nh,
                                                          uint8_t* __restrict__ hashes) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.y == 0) {
        if (k >= count * row_u64) return;
        const uint64_t r = k / row_u64, c = k - r * row_u64;
        rows[k] = matrix[ridx[r] * row_u64 + c];
    } else {
        if (k >= 2 * nh) return;
        const uint4* s = reinterpret_cast<const uint4*>(nodes + 32 * hidx[k >> 1]);
        reinterpret_cast<uint4*>(hashes)[k] = s[k & 1];
    }
}

// ------------------------------------------------------------ vector RLC
// out[i] = f_0[i] + sum_{k>=1} c_k f_k[i], f_k base-field vectors of length n
__global__ __launch_bounds__(256) void rlc_combine_kernel(const uint64_t* __restrict__ f, uint64_t n, uint32_t nvec,
                                                          const uint64_t* __restrict__ coef,
                                                          uint64_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    e3 acc = {f[i], 0, 0};
    for (uint32_t k = 1; k < nvec; ++k) acc = gl::e3_add(acc, gl::e3_mul_base(ld3(coef, k), f[(uint64_t)k * n + i]));
    st3(out, i, acc);
}

// Completion signal: one lane makes every earlier write of the stream visible
// system-wide and then stores the sequence number into mapped host memory.
__global__ void signal_kernel(uint64_t* flag, uint64_t seq) {
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned blocks_of(uint64_t work) { return (unsigned)((work + 255) / 256); }

}  // namespace

void pow_batch_launch(const PowBatchArgs& a, unsigned long long* d_best, unsigned long long* d_reset,
                      uint64_t* block_min, hipStream_t st) {
    if (a.hash == HASH_SHA2)
        hipLaunchKernelGGL(pow_batch_kernel<HASH_SHA2>, dim3(a.total_blocks), dim3(256), 0, st, a, d_best, d_reset,
                           block_min);
    else
        hipLaunchKernelGGL(pow_batch_kernel<HASH_BLAKE3>, dim3(a.total_blocks), dim3(256), 0, st, a, d_best, d_reset,
                           block_min);
}

void signal_launch(uint64_t* flag_d, uint64_t seq, hipStream_t st) {
    hipLaunchKernelGGL(signal_kernel, dim3(1), dim3(64), 0, st, flag_d, seq);
}

void fill_u64_launch(uint64_t* p, uint64_t n, uint64_t v, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks_of(n)), dim3(256), 0, st, p, n, v);
}

void term_tables_launch(const uint64_t* d_desc_e, uint32_t ne, const uint64_t* d_desc_b, uint32_t nb,
                        const uint64_t* d_coords, uint32_t nv, uint32_t L, uint64_t* hi_e, uint64_t* lo_e,
                        uint64_t* hi_b, uint64_t* lo_b, uint32_t* hl, hipStream_t st) {
    if (!ne && !nb) return;
    uint64_t span = std::max<uint64_t>(1ull << (nv - L), 1ull << L);
    hipLaunchKernelGGL(term_tables_kernel, dim3(blocks_of(span), ne + nb), dim3(256), 0, st, d_desc_e, d_desc_b,
                       d_coords, nv, L, ne, hi_e, lo_e, hi_b, lo_b, hl);
}

uint64_t term_limb_bytes(uint32_t nv, uint32_t L, uint32_t terms) {
    return L >= 8 ? (uint64_t)terms * (1ull << (nv - L)) * TERM_LIMB_WORDS * 4 : 0;
}

void term_accumulate_launch(uint64_t* d_acc, uint64_t n, uint32_t L, const uint64_t* hi_e, const uint64_t* lo_e,
                            uint32_t te, const uint64_t* hi_b, const uint64_t* lo_b, uint32_t tb, const uint32_t* hl,
                            bool accumulate, hipStream_t st) {
    const uint64_t LO = 1ull << L, H = n >> L;
    if (LO < 256) {
        hipLaunchKernelGGL(term_accumulate_kernel, dim3(blocks_of(n)), dim3(256), 0, st, d_acc, n, L, hi_e, lo_e, te,
                           hi_b, lo_b, tb, accumulate ? 1 : 0);
        return;
    }
    hipLaunchKernelGGL(term_sweep_kernel, dim3((unsigned)((LO / 256) * H)), dim3(256), 0, st, d_acc, H, L, hl, lo_e,
                       te, lo_b, tb, accumulate ? 1 : 0);
}

unsigned mle_eval_launch(const uint64_t* f, bool f_base, uint64_t n, uint32_t L, const uint64_t* hi,
                         const uint64_t* lo, uint32_t points, uint64_t* partials, hipStream_t st) {
    if (points < 1 || points > MLE_MAX_POINTS) return 0;
    // 3 * points totals per block: 2048 blocks for one or two points, 1024 for three or four
    unsigned g = std::min<unsigned>(blocks_of(n), points > 2 ? SUMCHECK_MAX_BLOCKS / 2 : SUMCHECK_MAX_BLOCKS);
    // the kernel needs a stride (256 g) that is a multiple of 2^L
    const unsigned gq = L > 8 ? (1u << (L - 8)) : 1u;
    g = std::max(gq, g / gq * gq);
    const int fb = f_base ? 1 : 0;
    switch (points) {
        case 1: hipLaunchKernelGGL(mle_eval_kernel<1>, dim3(g), dim3(256), 0, st, f, fb, n, L, hi, lo, partials); break;
        case 2: hipLaunchKernelGGL(mle_eval_kernel<2>, dim3(g), dim3(256), 0, st, f, fb, n, L, hi, lo, partials); break;
        case 3: hipLaunchKernelGGL(mle_eval_kernel<3>, dim3(g), dim3(256), 0, st, f, fb, n, L, hi, lo, partials); break;
        default: hipLaunchKernelGGL(mle_eval_kernel<4>, dim3(g), dim3(256), 0, st, f, fb, n, L, hi, lo, partials);
    }
    return g;
}

unsigned fold_compute_launch(const uint64_t* a, bool a_base, const uint64_t* b, uint64_t n, const uint64_t r[3],
                             uint64_t* a_out, uint64_t* b_out, uint64_t* partials, hipStream_t st, bool device_sum) {
    unsigned g = std::min<unsigned>(blocks_of(n / 4), SUMCHECK_MAX_BLOCKS);
    e3 rr = {r[0], r[1], r[2]};
    hipLaunchKernelGGL(fold_compute_kernel, dim3(g), dim3(256), 0, st, a, a_base ? 1 : 0, b, n / 4, rr, a_out, b_out,
                       partials);
    if (device_sum) hipLaunchKernelGGL(red::sum_blocks_kernel<2>, dim3(1), dim3(256), 0, st, partials, g, partials);
    return g;
}

void gather_rows_launch(const uint64_t* matrix, uint64_t row_u64, const uint64_t* d_idx, uint64_t count, uint64_t* out,
                        hipStream_t st) {
    if (!count || !row_u64) return;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_of(count * row_u64)), dim3(256), 0, st, matrix, row_u64, d_idx,
                       count, out);
}

void gather_open_launch(const uint64_t* matrix, uint64_t row_u64, const uint64_t* d_ridx, uint64_t count,
                        uint64_t* rows, const uint8_t* nodes, const uint64_t* d_hidx, uint64_t nh, uint8_t* hashes,
                        hipStream_t st) {
    const uint64_t span = std::max<uint64_t>(count * row_u64, 2 * nh);
    if (!span) return;
    hipLaunchKernelGGL(gather_open_kernel, dim3(blocks_of(span), 2), dim3(256), 0, st, matrix, row_u64, d_ridx, count,
                       rows, nodes, d_hidx, nh, hashes);
}

void rlc_combine_launch(const uint64_t* f, uint64_t n, uint32_t nvec, const uint64_t* d_coef, uint64_t* out,
                        hipStream_t st) {
    hipLaunchKernelGGL(rlc_combine_kernel, dim3(blocks_of(n)), dim3(256), 0, st, f, n, nvec, d_coef, out);
}

}  // namespace whir_gpu
