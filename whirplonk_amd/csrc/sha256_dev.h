// sha256_dev.h — SHA-256 (FIPS 180-4) for one message per lane: the second
// hash engine of the Merkle trees, row leaves and proof of work
// (hash/sha2_engine.hpp:47-81 hashes every message with the plain SHA-256 of
// its bytes; EngineId ENGINE_ID_SHA2).
//
// Messages here are rows of little-endian u64 limbs (row leaves), two 32-byte
// child hashes (tree nodes) or challenge || LE64(nonce) || 0^24 (proof of
// work).  SHA-256 reads big-endian 32-bit words, so a little-endian word of
// the byte string enters the schedule byte-swapped, and the digest's words
// are byte-swapped back into bytes.  Integer VALU only: add, xor, rotate
// (v_alignbit_b32), three-input logic (v_bitop3_b32), bitfield select (v_bfi_b32),
// byte permute (v_perm_b32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace s256 {

// ------------------------------------------------------------ constants
// K[i]: the first 32 fractional bits of the cube roots of the first 64 primes;
// H0[i]: of the square roots of the first 8 (FIPS 180-4, 4.2.2 and 5.3.3).
// Computed here from integer roots: floor(cbrt(p * 2^96)) mod 2^32 and
// floor(sqrt(p * 2^64)) mod 2^32.
struct Tables {
    uint32_t k[64];
    uint32_t h0[8];
};

constexpr uint64_t icbrt(unsigned __int128 n) {
    uint64_t lo = 0, hi = 1ull << 40;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if ((unsigned __int128)mid * mid * mid <= n)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

constexpr uint64_t isqrt(unsigned __int128 n) {
    uint64_t lo = 0, hi = 1ull << 40;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if ((unsigned __int128)mid * mid <= n)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

constexpr Tables make_tables() {
    Tables t{};
    int found = 0;
    for (uint32_t p = 2; found < 64; ++p) {
        bool prime = true;
        for (uint32_t d = 2; d * d <= p; ++d)
            if (p % d == 0) prime = false;
        if (!prime) continue;
        t.k[found] = (uint32_t)icbrt((unsigned __int128)p << 96);
        if (found < 8) t.h0[found] = (uint32_t)isqrt((unsigned __int128)p << 64);
        ++found;
    }
    return t;
}

constexpr Tables T = make_tables();
static_assert(T.h0[0] == 0x6A09E667u && T.k[0] == 0x428A2F98u && T.k[63] == 0xC67178F2u,
              "SHA-256 constants (FIPS 180-4)");

// ------------------------------------------------------------ compression
__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return __builtin_rotateright32(x, n); }
__device__ __forceinline__ uint32_t bswap(uint32_t x) { return __builtin_bswap32(x); }
// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96, symmetric in its inputs);
// the compiler otherwise spends two v_xor_b32 on each sigma function.  The
// builtin is opaque to constant folding, so a value known at compile time
// (the padding block's message schedule) takes the plain form and folds away.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    if (__builtin_constant_p(a ^ b ^ c)) return a ^ b ^ c;
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ void init(uint32_t s[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] = T.h0[i];
}

// One 64-byte block w[0..16) (big-endian words, already byte-swapped) into s.
__device__ __forceinline__ void compress(uint32_t s[8], const uint32_t* win) {
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = win[i];
    uint32_t a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6], h = s[7];
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        if (i >= 16) {
            const uint32_t w15 = w[(i - 15) & 15], w2 = w[(i - 2) & 15];
            const uint32_t s0 = xor3(rotr(w15, 7), rotr(w15, 18), w15 >> 3);
            const uint32_t s1 = xor3(rotr(w2, 17), rotr(w2, 19), w2 >> 10);
            w[i & 15] = w[i & 15] + s0 + w[(i - 7) & 15] + s1;
        }
        const uint32_t S1 = xor3(rotr(e, 6), rotr(e, 11), rotr(e, 25));
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t t1 = h + S1 + ch + T.k[i] + w[i & 15];
        const uint32_t S0 = xor3(rotr(a, 2), rotr(a, 13), rotr(a, 22));
        const uint32_t mj = (a & b) | (c & (a | b));
        const uint32_t t2 = S0 + mj;
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    s[0] += a;
    s[1] += b;
    s[2] += c;
    s[3] += d;
    s[4] += e;
    s[5] += f;
    s[6] += g;
    s[7] += h;
}

// digest words as stored little-endian: the 32 digest bytes in order
__device__ __forceinline__ void digest(const uint32_t s[8], uint32_t out[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = bswap(s[i]);
}

// SHA-256 of a 64-byte message given as 16 little-endian words (tree nodes:
// left || right; proof of work: challenge || nonce || zeros): the data block,
// then the padding block (0x80, zeros, bit length 512).
__device__ __forceinline__ void hash64(const uint32_t* m_le, uint32_t out[8]) {
    uint32_t s[8], w[16];
    init(s);
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = bswap(m_le[i]);
    compress(s, w);
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = 0;
    w[0] = 0x80000000u;
    w[15] = 512;
    compress(s, w);
    digest(s, out);
}

// SHA-256 of n little-endian u64 limbs (8n bytes): a row leaf.
__device__ __forceinline__ void hash_limbs(const uint64_t* row, uint32_t n, uint32_t out[8]) {
    uint32_t s[8], w[16];
    init(s);
    uint32_t done = 0;
    for (; done + 8 <= n; done += 8) {  // full 64-byte blocks
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t x = row[done + k];
            w[2 * k] = bswap((uint32_t)x);
            w[2 * k + 1] = bswap((uint32_t)(x >> 32));
        }
        compress(s, w);
    }
    // the last 0-7 limbs, 0x80, zeros and the 64-bit bit length; a second
    // block when the length does not fit behind 7 limbs
    const uint32_t rest = n - done;
    const uint64_t bits = (uint64_t)n * 64;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint64_t x = (uint32_t)k < rest ? row[done + k] : 0;
        w[2 * k] = bswap((uint32_t)x);
        w[2 * k + 1] = bswap((uint32_t)(x >> 32));
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((uint32_t)k == rest) w[2 * k] = 0x80000000u;
    if (rest < 7) {
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        compress(s, w);
    } else {
        compress(s, w);
#pragma unroll
        for (int i = 0; i < 14; ++i) w[i] = 0;
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        compress(s, w);
    }
    digest(s, out);
}

}  // namespace s256
