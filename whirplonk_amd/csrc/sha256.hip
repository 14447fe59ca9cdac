// sha256.hip — the SHA-256 engine's kernels on gfx950: row leaves, the flat
// Merkle tree and hash_many (hash/sha2_engine.hpp:47-81, merkle_tree.hpp:
// 104-147 with parent = SHA-256(left || right)).  One message per lane; the
// tree is built one level per launch (the engine serves the SHA-256
// configurations of the protocol, which the reference's own runs use at small
// sizes; the BLAKE3 engine is the fast path).
#include "kernels.h"
#include "sha256_dev.h"

namespace whir_gpu {
namespace {

unsigned blocks_256(uint64_t n) { return (unsigned)((n + 255) / 256); }

__device__ __forceinline__ void store32(uint8_t* p, const uint32_t h[8]) {
    uint4* d = reinterpret_cast<uint4*>(p);
    d[0] = make_uint4(h[0], h[1], h[2], h[3]);
    d[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

__global__ __launch_bounds__(256) void sha256_leaf_kernel(const uint64_t* __restrict__ m, uint64_t rows,
                                                          uint32_t row_u64, uint8_t* __restrict__ leaves) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    uint32_t h[8];
    s256::hash_limbs(m + r * row_u64, row_u64, h);
    store32(leaves + 32 * r, h);
}

// the `count` parents of the level whose 2 * count children start at child_off
__global__ __launch_bounds__(256) void sha256_level_kernel(uint8_t* __restrict__ nodes, uint64_t child_off,
                                                           uint64_t count, uint8_t* root_copy) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4* src = reinterpret_cast<const uint4*>(nodes + 32 * (child_off + 2 * i));
    uint32_t m[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = src[q];
        m[4 * q] = v.x;
        m[4 * q + 1] = v.y;
        m[4 * q + 2] = v.z;
        m[4 * q + 3] = v.w;
    }
    uint32_t h[8];
    s256::hash64(m, h);
    store32(nodes + 32 * (child_off + 2 * count + i), h);
    if (root_copy && count == 1) store32(root_copy, h);
}

// The last levels of a tree in one workgroup: the level of 2 * parents
// children at child_off (parents <= 256, a power of two) up to the root, each
// level's hashes kept in LDS for the next and stored to the node array.  A
// level this small is bound by the latency of one node hash (one wave per
// SIMD), so merging the levels saves each one's launch and global reload.
__global__ __launch_bounds__(256) void sha256_top_kernel(uint8_t* __restrict__ nodes, uint64_t child_off,
                                                         uint32_t parents, uint8_t* root_copy) {
    __shared__ uint32_t lvl[2][256 * 8];
    const uint32_t t = threadIdx.x;
    uint64_t off = child_off;
    uint32_t width = 2 * parents;
    int cur = 0;
    for (bool first = true; width > 1; first = false) {
        const uint32_t np = width / 2;
        if (t < np) {
            uint32_t m[16];
            if (first) {
                const uint4* src = reinterpret_cast<const uint4*>(nodes + 32 * (off + 2 * t));
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint4 v = src[q];
                    m[4 * q] = v.x;
                    m[4 * q + 1] = v.y;
                    m[4 * q + 2] = v.z;
                    m[4 * q + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 16; ++k) m[k] = lvl[cur][16 * t + k];
            }
            uint32_t h[8];
            s256::hash64(m, h);
            store32(nodes + 32 * (off + width + t), h);
#pragma unroll
            for (int k = 0; k < 8; ++k) lvl[cur ^ 1][8 * t + k] = h[k];
            if (root_copy && np == 1) store32(root_copy, h);
        }
        __syncthreads();
        cur ^= 1;
        off += width;
        width = np;
    }
}

// big-endian word k of the 64-byte block at p (bytes past `avail` are zero,
// and the byte at `avail` is the 0x80 end marker when `mark`)
__device__ __forceinline__ uint32_t block_word(const uint8_t* p, int k, uint64_t avail, bool mark) {
    uint32_t w = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint64_t j = 4 * k + b;
        uint32_t byte = 0;
        if (j < avail)
            byte = p[j];
        else if (mark && j == avail)
            byte = 0x80;
        w |= byte << (24 - 8 * b);
    }
    return w;
}

// out[i] = SHA-256(in[i*size .. (i+1)*size]) for any byte size
__global__ __launch_bounds__(256) void sha256_many_kernel(const uint8_t* __restrict__ in, uint64_t size,
                                                          uint64_t count, uint8_t* __restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const uint8_t* p = in + r * size;
    uint32_t s[8], w[16];
    s256::init(s);
    const uint64_t full = size / 64, rest = size - 64 * full, bits = size * 8;
    for (uint64_t b = 0; b < full; ++b) {
#pragma unroll
        for (int k = 0; k < 16; ++k) w[k] = block_word(p + 64 * b, k, 64, false);
        s256::compress(s, w);
    }
    const uint8_t* q = p + 64 * full;
#pragma unroll
    for (int k = 0; k < 16; ++k) w[k] = block_word(q, k, rest, true);
    if (rest >= 56) {  // no room for the length: one more block
        s256::compress(s, w);
#pragma unroll
        for (int k = 0; k < 16; ++k) w[k] = 0;
    }
    w[14] = (uint32_t)(bits >> 32);
    w[15] = (uint32_t)bits;
    s256::compress(s, w);
    uint32_t h[8];
    s256::digest(s, h);
    store32(out + 32 * r, h);
}

}  // namespace

void sha256_leaves_launch(const uint64_t* matrix, uint64_t rows, uint32_t row_u64, uint8_t* leaves,
                          hipStream_t st) {
    if (!rows) return;
    hipLaunchKernelGGL(sha256_leaf_kernel, dim3(blocks_256(rows)), dim3(256), 0, st, matrix, rows, row_u64, leaves);
}

void sha256_merkle_build_launch(uint8_t* nodes, uint64_t n_leaves, hipStream_t st, uint8_t* root_copy) {
    uint64_t width = 1;
    while (width < n_leaves) width <<= 1;
    if (n_leaves < width) fill_zero_launch(nodes + 32 * n_leaves, 32 * (width - n_leaves), st);
    if (n_leaves == 0) {
        fill_zero_launch(nodes, 32, st);  // empty tree: root is the zero hash
        return;
    }
    uint64_t off = 0;
    while (width > 512) {  // one level per launch while a level fills more than one CU
        const uint64_t parents = width / 2;
        hipLaunchKernelGGL(sha256_level_kernel, dim3(blocks_256(parents)), dim3(256), 0, st, nodes, off, parents,
                           root_copy);
        off += width;
        width = parents;
    }
    if (width > 1)
        hipLaunchKernelGGL(sha256_top_kernel, dim3(1), dim3(256), 0, st, nodes, off, (uint32_t)(width / 2), root_copy);
}

void sha256_hash_many_launch(const uint8_t* in, uint64_t size, uint64_t count, uint8_t* out, hipStream_t st) {
    if (!count) return;
    hipLaunchKernelGGL(sha256_many_kernel, dim3(blocks_256(count)), dim3(256), 0, st, in, size, count, out);
}

}  // namespace whir_gpu
