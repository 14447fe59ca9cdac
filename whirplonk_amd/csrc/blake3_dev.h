// blake3_dev.h — BLAKE3 (hash mode, 32-byte output) for one message per lane.
//
// Standard BLAKE3 as the reference uses it through its vendored C library
// (hash/blake3_engine.hpp:90-199, third_party/blake3 1.5.4): unkeyed, single
// chunk for messages <= 1024 bytes (CHUNK_START on the first block,
// CHUNK_END|ROOT on the last, counter 0); longer messages are split into
// 1024-byte chunks whose chaining values merge in a binary tree (left
// subtree = largest power of two of chunks), the root node carrying ROOT.
// All 32-bit integer VALU work: add, xor, rotate (v_alignbit_b32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

__host__ __device__ constexpr uint32_t iv(int i) {
    return i == 0 ? 0x6A09E667u : i == 1 ? 0xBB67AE85u : i == 2 ? 0x3C6EF372u : i == 3 ? 0xA54FF53Au
         : i == 4 ? 0x510E527Fu : i == 5 ? 0x9B05688Cu : i == 6 ? 0x1F83D9ABu : 0x5BE0CD19u;
}

// message word used at position i of round r: the spec permutes the message
// words between rounds with perm; here the permutation is resolved at compile
// time so no data moves at run time.
__host__ __device__ constexpr int perm(int i) {
    constexpr int P[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
    return P[i];
}
__host__ __device__ constexpr int sched(int r, int i) {
    for (int k = 0; k < r; ++k) i = perm(i);
    return i;
}

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return __builtin_rotateright32(x, n); }

#define B3_G(a, b, c, d, x, y)          \
    do {                                \
        s[a] = s[a] + s[b] + (x);       \
        s[d] = rotr(s[d] ^ s[a], 16);   \
        s[c] = s[c] + s[d];             \
        s[b] = rotr(s[b] ^ s[c], 12);   \
        s[a] = s[a] + s[b] + (y);       \
        s[d] = rotr(s[d] ^ s[a], 8);    \
        s[c] = s[c] + s[d];             \
        s[b] = rotr(s[b] ^ s[c], 7);    \
    } while (0)

// Compression: cv (8 words) and one 64-byte block m (16 LE words).  Writes the
// first 8 output words (the chaining value, or the hash when flags has ROOT).
__device__ __forceinline__ void compress(const uint32_t* cv, const uint32_t* m, uint64_t counter,
                                         uint32_t block_len, uint32_t flags, uint32_t* out) {
    uint32_t s[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] = cv[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) s[8 + i] = iv(i);
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        B3_G(0, 4, 8, 12, m[sched(r, 0)], m[sched(r, 1)]);
        B3_G(1, 5, 9, 13, m[sched(r, 2)], m[sched(r, 3)]);
        B3_G(2, 6, 10, 14, m[sched(r, 4)], m[sched(r, 5)]);
        B3_G(3, 7, 11, 15, m[sched(r, 6)], m[sched(r, 7)]);
        B3_G(0, 5, 10, 15, m[sched(r, 8)], m[sched(r, 9)]);
        B3_G(1, 6, 11, 12, m[sched(r, 10)], m[sched(r, 11)]);
        B3_G(2, 7, 8, 13, m[sched(r, 12)], m[sched(r, 13)]);
        B3_G(3, 4, 9, 14, m[sched(r, 14)], m[sched(r, 15)]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = s[i] ^ s[i + 8];
}
#undef B3_G

// BLAKE3 of a 64-byte message (a Merkle node: left || right).
__device__ __forceinline__ void hash64(const uint32_t* m, uint32_t* out) {
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = iv(i);
    compress(cv, m, 0, 64, CHUNK_START | CHUNK_END | ROOT, out);
}

// Lane k of a quad (lanes 4q..4q+3) reads the value of lane (k + R) % 4.
template <int R>
__device__ __forceinline__ uint32_t quad_rot(uint32_t x) {
    constexpr int ctrl = (R & 3) | (((R + 1) & 3) << 2) | (((R + 2) & 3) << 4) | (((R + 3) & 3) << 6);
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, ctrl, 0xF, 0xF, false);
}

// hash64 shared by the four lanes of a quad (all four active): lane k holds
// column k of the 4x4 state and runs the column G of index k; for the
// diagonal step rows 1-3 rotate across the quad (DPP quad permutes) and back.
// Each lane does 2 of the 8 G functions per round, so a node's latency is
// about a third of hash64's.  m: the 16 message words (the same in the four
// lanes).  Returns words k and 4 + k of the hash.
__device__ __forceinline__ void hash64_quad(const uint32_t* m, uint32_t k, uint32_t& w_lo, uint32_t& w_hi) {
    const bool k1 = k & 1, k2 = k & 2;
    auto sel = [&](uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) { return k2 ? (k1 ? x3 : x2) : (k1 ? x1 : x0); };
    uint32_t a = sel(iv(0), iv(1), iv(2), iv(3));
    uint32_t b = sel(iv(4), iv(5), iv(6), iv(7));
    uint32_t c = a;  // v[8 + k] = iv(k)
    uint32_t d = sel(0u, 0u, 64u, CHUNK_START | CHUNK_END | ROOT);  // counter, block length, flags
    auto G = [](uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d, uint32_t x, uint32_t y) {
        a = a + b + x;
        d = rotr(d ^ a, 16);
        c = c + d;
        b = rotr(b ^ c, 12);
        a = a + b + y;
        d = rotr(d ^ a, 8);
        c = c + d;
        b = rotr(b ^ c, 7);
    };
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        G(a, b, c, d, sel(m[sched(r, 0)], m[sched(r, 2)], m[sched(r, 4)], m[sched(r, 6)]),
          sel(m[sched(r, 1)], m[sched(r, 3)], m[sched(r, 5)], m[sched(r, 7)]));
        b = quad_rot<1>(b);
        c = quad_rot<2>(c);
        d = quad_rot<3>(d);
        G(a, b, c, d, sel(m[sched(r, 8)], m[sched(r, 10)], m[sched(r, 12)], m[sched(r, 14)]),
          sel(m[sched(r, 9)], m[sched(r, 11)], m[sched(r, 13)], m[sched(r, 15)]));
        b = quad_rot<3>(b);
        c = quad_rot<2>(c);
        d = quad_rot<1>(d);
    }
    w_lo = a ^ c;
    w_hi = b ^ d;
}

// Chaining value of one chunk (<= 1024 bytes) whose bytes are produced by
// the functor load(word_index) -> uint32 (little-endian words; bytes past len
// must read as zero).  is_root marks the single-chunk root case.
template <typename Load>
__device__ __forceinline__ void chunk_cv(Load load, uint32_t len, uint64_t chunk_idx, bool is_root, uint32_t* cv) {
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = iv(i);
    uint32_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (uint32_t b = 0; b < nblocks; ++b) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = load(b * 16 + i);
        uint32_t bl = len == 0 ? 0 : ((len - 64 * b) < 64 ? (len - 64 * b) : 64);
        uint32_t flags = (b == 0 ? CHUNK_START : 0);
        if (b + 1 == nblocks) flags |= CHUNK_END | (is_root ? ROOT : 0);
        compress(cv, m, chunk_idx, bl, flags, cv);
    }
}

__device__ __forceinline__ void parent_cv(const uint32_t* l, const uint32_t* r, bool is_root, uint32_t* out) {
    uint32_t m[16], key[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        m[i] = l[i];
        m[8 + i] = r[i];
        key[i] = iv(i);
    }
    compress(key, m, 0, 64, PARENT | (is_root ? ROOT : 0), out);
}

// Full BLAKE3 of a message of any length.  Chunk CVs are merged with the
// standard stack discipline (a chunk count's trailing zero bits decide the
// merges), which builds exactly the left-balanced tree of the spec.
template <typename Load>
__device__ void hash_any(Load load, uint64_t len, uint32_t* out) {
    uint64_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
    if (nchunks == 1) {
        chunk_cv(load, (uint32_t)len, 0, true, out);
        return;
    }
    uint32_t stack[16][8];  // up to 2^16 chunks (64 MiB messages)
    int depth = 0;
    for (uint64_t c = 0; c < nchunks; ++c) {
        uint32_t clen = (uint32_t)((len - c * 1024) < 1024 ? (len - c * 1024) : 1024);
        uint64_t base = c * 256;  // words
        uint32_t cv[8];
        chunk_cv([&](uint32_t w) { return load(base + w); }, clen, c, false, cv);
        if (c + 1 == nchunks) {
            // final: fold the stack from the top, the last merge is the root
            for (int k = depth - 1; k >= 0; --k) parent_cv(stack[k], cv, k == 0, cv);
#pragma unroll
            for (int i = 0; i < 8; ++i) out[i] = cv[i];
            return;
        }
        uint64_t total = c + 1;
        while ((total & 1) == 0) {
            parent_cv(stack[depth - 1], cv, false, cv);
            --depth;
            total >>= 1;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) stack[depth][i] = cv[i];
        ++depth;
    }
}

}  // namespace b3
