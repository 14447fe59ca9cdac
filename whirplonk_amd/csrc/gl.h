// gl.h — Goldilocks (p = 2^64 - 2^32 + 1) and its cubic extension on gfx950.
//
// Exact integer arithmetic only: 32-bit VALU mul_lo/mul_hi/mad_u64_u32, adds
// with carry, shifts.  No floating point, no MFMA.  Every value leaving a
// function is canonical in [0, p), so results are bit-identical to the
// reference's Goldilocks (algebra/goldilocks.hpp:54-202) and GoldilocksExt3
// (algebra/goldilocks_ext3.hpp:33-137) on the same inputs.
//
// The reduction uses the field's structure: for x = hi_hi*2^96 + hi_lo*2^64 + lo,
// 2^64 = eps (eps = 2^32 - 1) and 2^96 = -1 mod p, so x = lo - hi_hi + hi_lo*eps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GL_P 0xFFFFFFFF00000001ull
#define GL_EPS 0xFFFFFFFFull

namespace gl {

__host__ __device__ __forceinline__ uint64_t canon(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

#if defined(__HIP_DEVICE_COMPILE__)
#define GL_FORM __device__ __forceinline__
#include "gl_forms.h"
#undef GL_FORM
#else
__host__ __forceinline__ uint64_t add(uint64_t a, uint64_t b) {
    // a, b canonical: a + b < 2p; on carry the wrapped sum minus p is exact
    uint64_t s = a + b;
    bool over = (s < a) | (s >= GL_P);
    return over ? s - GL_P : s;
}

__host__ __forceinline__ uint64_t sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    return (a < b) ? d + GL_P : d;
}

__host__ __forceinline__ uint64_t neg(uint64_t a) { return a ? GL_P - a : 0; }

// x = hi*2^64 + lo, any hi, lo < 2^64; returns x mod p, canonical.
__host__ __forceinline__ uint64_t reduce128(uint64_t lo, uint64_t hi) {
    uint64_t hi_hi = hi >> 32;
    uint64_t hi_lo = hi & 0xFFFFFFFFull;
    uint64_t t0 = lo - hi_hi;
    if (lo < hi_hi) t0 -= GL_EPS;           // borrow: -2^64 = -eps
    uint64_t t1 = (hi_lo << 32) - hi_lo;    // hi_lo * eps, < p
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;                // carry: +2^64 = +eps (cannot carry again)
    return canon(r);
}
#endif

__host__ __device__ __forceinline__ uint64_t dbl(uint64_t a) { return add(a, a); }

#if defined(__HIP_DEVICE_COMPILE__)
// 64x64 -> 128 from four 32x32 mad_u64_u32 (every partial sum fits in 64 bits)
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t p00 = (uint64_t)a0 * b0;
    uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
    uint64_t u = (uint64_t)a1 * b0 + (t & 0xFFFFFFFFull);
    lo = (u << 32) | (p00 & 0xFFFFFFFFull);
    hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
}
#else
__host__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    unsigned __int128 p = (unsigned __int128)a * b;
    lo = (uint64_t)p;
    hi = (uint64_t)(p >> 64);
}
#endif

__host__ __device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128(lo, hi);
}

__host__ __device__ __forceinline__ uint64_t pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mul(r, a);
        a = mul(a, a);
        e >>= 1;
    }
    return r;
}

__host__ __device__ __forceinline__ uint64_t inv(uint64_t a) { return pow(a, GL_P - 2); }

// x * 2^E mod p for a compile-time E in [0, 192) (2^192 = 1): the twiddles of
// every sub-transform of size <= 64 are powers of two, since the reference's
// primitive 64th root is 2^39.
template <int E>
__host__ __device__ __forceinline__ uint64_t mul_pow2(uint64_t x) {
    static_assert(E >= 0 && E < 192, "exponent out of range");
    if constexpr (E == 0) {
        return x;
    } else if constexpr (E >= 96) {
        return neg(mul_pow2<E - 96>(x));    // 2^96 = -1
    } else if constexpr (E < 64) {
        return reduce128(x << E, x >> (64 - E));
    } else {
        // bits [64,96) of x*2^E (mid) weigh eps, bits [96,160) (hi < 2^63) weigh -1
        uint64_t mid = (x << (E - 64)) & 0xFFFFFFFFull;
        uint64_t hi = x >> (96 - E);
        uint64_t t = (mid << 32) - mid;
        return (t >= hi) ? t - hi : t - hi + GL_P;
    }
}

// primitive 2^k-th root of unity as the reference's NTT engine picks it:
// (7^eps)^(2^32 / 2^k)  (goldilocks.hpp:155-159, cooley_tukey.hpp:166-177)
__host__ __forceinline__ uint64_t root_of_unity(uint32_t log_n) {
    uint64_t w32 = pow(7, GL_EPS);
    return pow(w32, 1ull << (32 - log_n));
}

// ------------------------------------------------------------------ Ext3
struct e3 {
    uint64_t c0, c1, c2;
};

__host__ __device__ __forceinline__ e3 e3_add(e3 a, e3 b) { return {add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)}; }
__host__ __device__ __forceinline__ e3 e3_sub(e3 a, e3 b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)}; }

// X^3 = 2: c0 = a0b0 + 2(a1b2 + a2b1), c1 = a0b1 + a1b0 + 2 a2b2, c2 = a0b2 + a1b1 + a2b0
__host__ __device__ __forceinline__ e3 e3_mul(e3 a, e3 b);

// ---- exact accumulation of 128-bit products, reduced once ----
// value = top * 2^128 + hi * 2^64 + lo; 2^128 = -2^32 (mod p), so the value
// is reduce128(lo, hi) - top * 2^32 (top stays far below 2^31 here).
struct acc128 {
    uint64_t lo, hi, top;
};

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void acc_add(acc128& a, uint64_t plo, uint64_t phi) {
    uint32_t c1, c2, c3, c4;
    uint32_t l0 = __builtin_addc(lo32(a.lo), lo32(plo), 0u, &c1);
    uint32_t l1 = __builtin_addc(hi32(a.lo), hi32(plo), c1, &c2);
    uint32_t h0 = __builtin_addc(lo32(a.hi), lo32(phi), c2, &c3);
    uint32_t h1 = __builtin_addc(hi32(a.hi), hi32(phi), c3, &c4);
    a.lo = mk64(l0, l1);
    a.hi = mk64(h0, h1);
    a.top += c4;
}
#else
__host__ __forceinline__ void acc_add(acc128& a, uint64_t plo, uint64_t phi) {
    unsigned __int128 v = ((unsigned __int128)a.hi << 64) | a.lo;
    unsigned __int128 p = ((unsigned __int128)phi << 64) | plo;
    unsigned __int128 s = v + p;
    a.top += s < p;
    a.lo = (uint64_t)s;
    a.hi = (uint64_t)(s >> 64);
}
#endif

// acc += x * y (exact)
__host__ __device__ __forceinline__ void acc_mac(acc128& a, uint64_t x, uint64_t y) {
    uint64_t lo, hi;
    mul128(x, y, lo, hi);
    acc_add(a, lo, hi);
}

__host__ __device__ __forceinline__ uint64_t acc_reduce(const acc128& a) {
    return sub(reduce128(a.lo, a.hi), a.top << 32);
}

// Ext3 accumulator: three exact component sums.
struct e3acc {
    acc128 c[3];
};

__host__ __device__ __forceinline__ e3acc e3acc_zero() { return {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}}; }

// acc += a * b over Ext3 (X^3 = 2), with no reduction
__host__ __device__ __forceinline__ void e3_mac(e3acc& acc, e3 a, e3 b) {
    const uint64_t a1d = dbl(a.c1), a2d = dbl(a.c2);
    acc_mac(acc.c[0], a.c0, b.c0);
    acc_mac(acc.c[0], a1d, b.c2);
    acc_mac(acc.c[0], a2d, b.c1);
    acc_mac(acc.c[1], a.c0, b.c1);
    acc_mac(acc.c[1], a.c1, b.c0);
    acc_mac(acc.c[1], a2d, b.c2);
    acc_mac(acc.c[2], a.c0, b.c2);
    acc_mac(acc.c[2], a.c1, b.c1);
    acc_mac(acc.c[2], a.c2, b.c0);
}

// acc += a * b with b in the base field (component-wise)
__host__ __device__ __forceinline__ void e3_mac_base(e3acc& acc, e3 a, uint64_t b) {
    acc_mac(acc.c[0], a.c0, b);
    acc_mac(acc.c[1], a.c1, b);
    acc_mac(acc.c[2], a.c2, b);
}

__host__ __device__ __forceinline__ e3 e3acc_reduce(const e3acc& acc) {
    return {acc_reduce(acc.c[0]), acc_reduce(acc.c[1]), acc_reduce(acc.c[2])};
}

// base x Ext3, component-wise (Basefield::mixed_mul, embedding.hpp:75-77)
__host__ __device__ __forceinline__ e3 e3_mul_base(e3 a, uint64_t b) { return {mul(a.c0, b), mul(a.c1, b), mul(a.c2, b)}; }

// each component is a sum of three products reduced once (e3_mac)
__host__ __device__ __forceinline__ e3 e3_mul(e3 a, e3 b) {
    e3acc acc = e3acc_zero();
    e3_mac(acc, a, b);
    return e3acc_reduce(acc);
}

__host__ __device__ __forceinline__ e3 e3_pow(e3 a, uint64_t e) {
    e3 r = {1, 0, 0};
    while (e) {
        if (e & 1) r = e3_mul(r, a);
        a = e3_mul(a, a);
        e >>= 1;
    }
    return r;
}

}  // namespace gl
