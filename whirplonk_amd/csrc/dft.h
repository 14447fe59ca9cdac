// dft.h — small Goldilocks DFTs for gfx950: in registers (size <= 64, all
// twiddles are powers of two) and two-step in LDS (size <= 1024 = 32 x 32).
//
// Definition (the reference's NTT, algebra/ntt/cooley_tukey.hpp:200-259):
//   X[k] = sum_j x[j] * w_N^{jk},   w_N = (7^eps)^(2^32/N),  natural order.
// The inverse direction uses w_N^{-1} with no 1/N scaling, matching
// NttEngine::intt (cooley_tukey.hpp:754-776).
#pragma once
#include "gl.h"

namespace dft {

// w_64 = 2^39 for the reference's roots; w_64^k = 2^(39k mod 192).
__host__ __device__ constexpr int w64_exp(int k, bool inv) {
    int e = (39 * (k & 63)) % 192;
    return inv ? (192 - e) % 192 : e;
}

// x * 2^e mod p for 0 <= e < 96, e a constant after unrolling (2^96 = -1 is
// folded into the butterfly by swapping its operands instead).
__device__ __forceinline__ uint64_t mul_pow2_lt96(uint64_t x, int e) {
    if (e == 0) return x;
    if (e < 64) return gl::reduce128(x << e, x >> (64 - e));
    // bits [64, 96) of x * 2^e weigh eps, bits [96, 160) weigh -1 (and are < 2^63 < p)
    const int k = e - 64;
    uint32_t mid = (uint32_t)x << k;
    uint64_t hi = x >> (32 - k);
    uint32_t b1, b2;
    uint32_t t0 = __builtin_subc(0u, mid, 0u, &b1);
    uint32_t t1 = __builtin_subc(mid, 0u, b1, &b2);
    return gl::sub(((uint64_t)t1 << 32) | t0, hi);
}

// (a - b) * 2^e for any e in [0, 192): a factor 2^96 = -1 becomes (b - a).
__device__ __forceinline__ uint64_t sub_mul_pow2(uint64_t a, uint64_t b, int e) {
    return e >= 96 ? mul_pow2_lt96(gl::sub(b, a), e - 96) : mul_pow2_lt96(gl::sub(a, b), e);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

constexpr int bitrev(int k, int bits) {
    int r = 0;
    for (int i = 0; i < bits; ++i) r |= ((k >> i) & 1) << (bits - 1 - i);
    return r;
}

// In-register DFT of size R (power of two, R <= 64), natural order in and out.
// Radix-2 decimation in frequency; the bit-reversed result is permuted back
// at compile time (register renaming only).
template <int R, bool INV>
__device__ __forceinline__ void dft_reg(uint64_t* v) {
    static_assert(R >= 1 && R <= 64 && (R & (R - 1)) == 0, "R must be a power of two <= 64");
    if constexpr (R == 1) {
        return;
    } else {
#pragma unroll
        for (int len = R; len >= 2; len >>= 1) {
#pragma unroll
            for (int g = 0; g < R; g += len) {
#pragma unroll
                for (int j = 0; j < len / 2; ++j) {
                    uint64_t a = v[g + j], b = v[g + j + len / 2];
                    v[g + j] = gl::add(a, b);
                    v[g + j + len / 2] = sub_mul_pow2(a, b, w64_exp(j * (64 / len), INV));
                }
            }
        }
        uint64_t t[R];
#pragma unroll
        for (int k = 0; k < R; ++k) t[k] = v[bitrev(k, ilog2(R))];
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = t[k];
    }
}

// Split of an LDS DFT size N into R1 x R2 register transforms.
template <int N>
struct Split {
    static constexpr int LOG = ilog2(N);
    static constexpr int R1 = 1 << (LOG / 2);          // first-step size
    static constexpr int R2 = N / R1;                  // second-step size (>= R1)
    static constexpr int LOG_R2 = ilog2(R2);
    static constexpr int PADDED = N + N / R2;          // one pad slot per R2 elements
};

// padded LDS index of element q in a DFT whose second step reads R2
// contiguous elements: breaks the bank aliasing of the stride-R2 reads
template <int N>
__device__ __forceinline__ int pidx(int q) {
    return q + (q >> Split<N>::LOG_R2);
}

// Step 1 of the two-step DFT of size N = R1*R2 on LDS buffer `buf` (padded
// layout): work item n2 < R2 transforms x[R2*n1 + n2] over n1, multiplies
// by w_N^{n2*k1} (table tw1024 holds w_1024^t), and writes back in place.
// Caller synchronizes before and after.
template <int N, bool INV>
__device__ __forceinline__ void lds_dft_step1(uint64_t* buf, int n2, const uint64_t* tw1024) {
    constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
    uint64_t v[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) v[n1] = buf[pidx<N>(R2 * n1 + n2)];
    dft_reg<R1, INV>(v);
    if constexpr (R2 > 1) {
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) {
            int t = ((n2 * k1) * (1024 / N)) & 1023;
            v[k1] = gl::mul(v[k1], tw1024[t]);
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) buf[pidx<N>(R2 * k1 + n2)] = v[k1];
}

// Step 2: work item k1 < R1 transforms the contiguous run x[R2*k1 + n2] over
// n2 and leaves X[k1 + R1*k2] in out[k2] (registers), k2 < R2.
template <int N, bool INV>
__device__ __forceinline__ void lds_dft_step2(const uint64_t* buf, int k1, uint64_t* out) {
    constexpr int R2 = Split<N>::R2;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) out[n2] = buf[pidx<N>(R2 * k1 + n2)];
    dft_reg<R2, INV>(out);
}

}  // namespace dft
