// merkle.hip — codeword row leaves and the binary Merkle tree, BLAKE3.
//
// Layout follows the reference's witness (protocols/merkle_tree.hpp:104-147):
// nodes = [leaves padded to 2^L with all-zero hashes | level L-1 | ... | root],
// 32 bytes per node, parent = BLAKE3(left || right) over 64 bytes.  One
// node (or leaf) per lane; levels wider than TOP_WIDTH are one launch each,
// the narrow top of the tree is finished by one workgroup.
#include "blake3_dev.h"
#include "kernels.h"

namespace whir_gpu {
namespace {

constexpr uint64_t TOP_WIDTH = 2048;  // levels at most this wide run in one workgroup of TOP_WIDTH/2 lanes

__device__ __forceinline__ void store_hash(uint8_t* dst, const uint32_t* h) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    d[0] = make_uint4(h[0], h[1], h[2], h[3]);
    d[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

__device__ __forceinline__ void load64(const uint8_t* src, uint32_t* m) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint4 v = s[i];
        m[4 * i] = v.x;
        m[4 * i + 1] = v.y;
        m[4 * i + 2] = v.z;
        m[4 * i + 3] = v.w;
    }
}

// leaves[r] = BLAKE3(LE u64 words of row r) — matrix_commit.hpp:169-195 with
// encode_into (:73-136): each canonical limb is written as 8 LE bytes, which
// is the native in-memory layout, so the row bytes are the row itself.
__global__ __launch_bounds__(256) void leaf_hash_kernel(const uint64_t* __restrict__ mat, uint64_t rows,
                                                        uint32_t row_u64, uint8_t* __restrict__ leaves) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint64_t* row = mat + r * row_u64;
    const uint64_t len = (uint64_t)row_u64 * 8;
    uint32_t out[8];
    if (len <= 1024) {
        uint32_t cv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) cv[i] = b3::iv(i);
        uint32_t nblocks = len == 0 ? 1 : (uint32_t)((len + 63) / 64);
        for (uint32_t b = 0; b < nblocks; ++b) {
            uint32_t m[16];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                uint32_t w = b * 8 + i;
                uint64_t x = w < row_u64 ? row[w] : 0;
                m[2 * i] = (uint32_t)x;
                m[2 * i + 1] = (uint32_t)(x >> 32);
            }
            uint32_t bl = (uint32_t)((len - 64 * b) < 64 ? (len - 64 * b) : 64);
            uint32_t flags = (b == 0 ? b3::CHUNK_START : 0);
            if (b + 1 == nblocks) flags |= b3::CHUNK_END | b3::ROOT;
            b3::compress(cv, m, 0, bl, flags, cv);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) out[i] = cv[i];
    } else {
        b3::hash_any(
            [&](uint64_t w) -> uint32_t {
                uint64_t q = w >> 1;
                if (q >= row_u64) return 0;
                uint64_t x = row[q];
                return (w & 1) ? (uint32_t)(x >> 32) : (uint32_t)x;
            },
            len, out);
    }
    store_hash(leaves + 32 * r, out);
}

// leaf_hash_kernel for rows of NB whole 64-byte blocks (row_u64 = 8 NB <= 128):
// each block's 16 message words come in four 16-byte loads.
template <int NB>
__global__ __launch_bounds__(256) void leaf_hash_blocks_kernel(const uint64_t* __restrict__ mat, uint64_t rows,
                                                               uint8_t* __restrict__ leaves) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint4* row = reinterpret_cast<const uint4*>(mat + r * 8 * NB);
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = b3::iv(i);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        uint32_t m[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = row[4 * b + q];
            m[4 * q] = v.x;
            m[4 * q + 1] = v.y;
            m[4 * q + 2] = v.z;
            m[4 * q + 3] = v.w;
        }
        uint32_t flags = (b == 0 ? b3::CHUNK_START : 0) | (b + 1 == NB ? (b3::CHUNK_END | b3::ROOT) : 0);
        b3::compress(cv, m, 0, 64, flags, cv);
    }
    store_hash(leaves + 32 * r, cv);
}

// out[i] = BLAKE3(in[i*size .. (i+1)*size]) for any byte size (hash_many,
// hash/blake3_engine.hpp:90-199; both of its branches compute standard BLAKE3)
__global__ __launch_bounds__(256) void hash_many_kernel(const uint8_t* __restrict__ in, uint64_t size,
                                                        uint64_t count, uint8_t* __restrict__ out) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const uint8_t* msg = in + r * size;
    uint32_t h[8];
    b3::hash_any(
        [&](uint64_t w) -> uint32_t {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint64_t b = 4 * w + k;
                if (b < size) v |= (uint32_t)msg[b] << (8 * k);
            }
            return v;
        },
        size, h);
    store_hash(out + 32 * r, h);
}

__global__ __launch_bounds__(256) void merkle_level_kernel(uint8_t* __restrict__ nodes, uint64_t child_off,
                                                           uint64_t parent_off, uint64_t width) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width) return;
    uint32_t m[16], h[8];
    load64(nodes + 32 * (child_off + 2 * i), m);
    b3::hash64(m, h);
    store_hash(nodes + 32 * (parent_off + i), h);
}

// Two levels per launch: lane i hashes children 4i..4i+3 into parents 2i, 2i+1
// (written out) and their parent i, so the middle level is never re-read.
__global__ __launch_bounds__(256) void merkle_2level_kernel(uint8_t* __restrict__ nodes, uint64_t child_off,
                                                            uint64_t child_width) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= child_width / 4) return;
    uint32_t m[16], h0[8], h1[8], g[8];
    load64(nodes + 32 * (child_off + 4 * i), m);
    b3::hash64(m, h0);
    load64(nodes + 32 * (child_off + 4 * i + 2), m);
    b3::hash64(m, h1);
    const uint64_t parent_off = child_off + child_width;
    store_hash(nodes + 32 * (parent_off + 2 * i), h0);
    store_hash(nodes + 32 * (parent_off + 2 * i + 1), h1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        m[k] = h0[k];
        m[8 + k] = h1[k];
    }
    b3::hash64(m, g);
    store_hash(nodes + 32 * (parent_off + child_width / 2 + i), g);
}

// The top of the tree (child level of width <= TOP_WIDTH), one workgroup: the
// level being reduced lives in LDS, so each level costs one compression per
// lane plus two barriers instead of a round trip through the cache; every
// parent is also written to the flat node array.
__global__ __launch_bounds__(1024) void merkle_top_kernel(uint8_t* __restrict__ nodes, uint64_t child_off,
                                                          uint64_t child_width, uint8_t* root_copy) {
    __shared__ uint32_t lvl[TOP_WIDTH * 8];
    for (uint64_t i = threadIdx.x; i < child_width; i += blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(nodes + 32 * (child_off + i));
        uint4 a = src[0], b = src[1];
        uint32_t* d = lvl + 8 * i;
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    }
    __syncthreads();
    while (child_width > 1) {
        const uint64_t width = child_width / 2, parent_off = child_off + child_width;
        if (width <= blockDim.x / 4) {
            // narrow levels: one parent per quad of lanes (latency, not throughput)
            const uint32_t i = threadIdx.x >> 2, k = threadIdx.x & 3;
            uint32_t lo = 0, hi = 0;
            if (i < width) {
                uint32_t m[16];
#pragma unroll
                for (int w = 0; w < 16; ++w) m[w] = lvl[16 * i + w];
                b3::hash64_quad(m, k, lo, hi);
                uint32_t* dst = reinterpret_cast<uint32_t*>(nodes + 32 * (parent_off + i));
                dst[k] = lo;
                dst[4 + k] = hi;
            }
            __syncthreads();  // every child read before any parent overwrites the level
            if (i < width) {
                lvl[8 * i + k] = lo;
                lvl[8 * i + 4 + k] = hi;
            }
            __syncthreads();
            child_off = parent_off;
            child_width = width;
            continue;
        }
        // width <= TOP_WIDTH / 2 = blockDim.x: one parent per lane
        const uint32_t i = threadIdx.x;
        uint32_t h[8];
        if (i < width) {
            uint32_t m[16];
#pragma unroll
            for (int w = 0; w < 16; ++w) m[w] = lvl[16 * i + w];
            b3::hash64(m, h);
            store_hash(nodes + 32 * (parent_off + i), h);
        }
        __syncthreads();  // every child read before any parent overwrites the level
        if (i < width) {
#pragma unroll
            for (int w = 0; w < 8; ++w) lvl[8 * i + w] = h[w];
        }
        __syncthreads();
        child_off = parent_off;
        child_width = width;
    }
    if (root_copy && threadIdx.x == 0) store_hash(root_copy, lvl);
}

// The three levels below the top when the child level has 512 to TOP_WIDTH
// nodes: workgroup w reduces children [256 w, 256 w + 256) to 32 parents, one
// parent per quad of lanes, with each level staged in LDS and written to the
// flat node array.  Spread over width / 256 CUs, these levels no longer queue
// on the single workgroup of merkle_top_kernel, which then starts from at
// most 256 children, all in quad mode.
__global__ __launch_bounds__(512) void merkle_top3_kernel(uint8_t* __restrict__ nodes, uint64_t child_off,
                                                          uint64_t child_width) {
    __shared__ uint32_t lvl[256 * 8];
    const uint64_t w = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(nodes + 32 * (child_off + 256 * w + i));
        uint4 a = src[0], b = src[1];
        uint32_t* d = lvl + 8 * i;
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    }
    __syncthreads();
    uint64_t off = child_off, width = child_width;
    uint32_t seg = 256;  // this workgroup's nodes at the level being reduced
    for (int lev = 0; lev < 3; ++lev) {
        const uint64_t parent_off = off + width;
        const uint32_t np = seg / 2;
        const uint32_t i = threadIdx.x >> 2, k = threadIdx.x & 3;
        uint32_t lo = 0, hi = 0;
        if (i < np) {
            uint32_t m[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) m[q] = lvl[16 * i + q];
            b3::hash64_quad(m, k, lo, hi);
            uint32_t* dst = reinterpret_cast<uint32_t*>(nodes + 32 * (parent_off + w * np + i));
            dst[k] = lo;
            dst[4 + k] = hi;
        }
        __syncthreads();  // every child read before any parent overwrites the level
        if (i < np) {
            lvl[8 * i + k] = lo;
            lvl[8 * i + 4 + k] = hi;
        }
        __syncthreads();
        off = parent_off;
        width /= 2;
        seg = np;
    }
}

// ------------------------------------------------ many trees per launch
// The trees of a batch built level by level together: each launch covers one
// step (the same kernel body as above) of every tree that has it, so a batch
// of T trees costs a handful of launches instead of T times as many.  A
// launch's segment table lists, per tree, its node array (byte offset from
// `base`), the child level's offset and width, and the launch's first
// workgroup for it; a workgroup finds its tree by binary search over the
// table (workgroup-uniform, scalar loads).
__device__ __forceinline__ MerkleSeg find_seg(const MerkleSeg* segs, uint32_t nseg) {
    uint32_t lo = 0, hi = nseg;  // segs[lo].block0 <= blockIdx.x < segs[hi].block0
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (segs[mid].block0 <= blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    return segs[lo];
}

__global__ __launch_bounds__(256) void merkle_2level_multi_kernel(uint8_t* __restrict__ base,
                                                                  const MerkleSeg* __restrict__ segs, uint32_t nseg) {
    const MerkleSeg g = find_seg(segs, nseg);
    uint8_t* nodes = base + g.nodes;
    const uint64_t i = (uint64_t)(blockIdx.x - g.block0) * blockDim.x + threadIdx.x;
    if (i >= g.width / 4) return;
    uint32_t m[16], h0[8], h1[8], p[8];
    load64(nodes + 32 * (g.child_off + 4 * i), m);
    b3::hash64(m, h0);
    load64(nodes + 32 * (g.child_off + 4 * i + 2), m);
    b3::hash64(m, h1);
    const uint64_t parent_off = g.child_off + g.width;
    store_hash(nodes + 32 * (parent_off + 2 * i), h0);
    store_hash(nodes + 32 * (parent_off + 2 * i + 1), h1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        m[k] = h0[k];
        m[8 + k] = h1[k];
    }
    b3::hash64(m, p);
    store_hash(nodes + 32 * (parent_off + g.width / 2 + i), p);
}

__global__ __launch_bounds__(256) void merkle_level_multi_kernel(uint8_t* __restrict__ base,
                                                                 const MerkleSeg* __restrict__ segs, uint32_t nseg) {
    const MerkleSeg g = find_seg(segs, nseg);
    uint8_t* nodes = base + g.nodes;
    const uint64_t i = (uint64_t)(blockIdx.x - g.block0) * blockDim.x + threadIdx.x;
    if (i >= g.width / 2) return;
    uint32_t m[16], h[8];
    load64(nodes + 32 * (g.child_off + 2 * i), m);
    b3::hash64(m, h);
    store_hash(nodes + 32 * (g.child_off + g.width + i), h);
}

// merkle_top3_kernel's three levels for each tree: workgroup w of a tree
// reduces its children [256 w, 256 w + 256) to 32 parents
__global__ __launch_bounds__(512) void merkle_top3_multi_kernel(uint8_t* __restrict__ base,
                                                                const MerkleSeg* __restrict__ segs, uint32_t nseg) {
    __shared__ uint32_t lvl[256 * 8];
    const MerkleSeg g = find_seg(segs, nseg);
    uint8_t* nodes = base + g.nodes;
    const uint64_t w = blockIdx.x - g.block0;
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(nodes + 32 * (g.child_off + 256 * w + i));
        uint4 a = src[0], b = src[1];
        uint32_t* d = lvl + 8 * i;
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    }
    __syncthreads();
    uint64_t off = g.child_off, width = g.width;
    uint32_t seg = 256;
    for (int lev = 0; lev < 3; ++lev) {
        const uint64_t parent_off = off + width;
        const uint32_t np = seg / 2;
        const uint32_t i = threadIdx.x >> 2, k = threadIdx.x & 3;
        uint32_t lo = 0, hi = 0;
        if (i < np) {
            uint32_t m[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) m[q] = lvl[16 * i + q];
            b3::hash64_quad(m, k, lo, hi);
            uint32_t* dst = reinterpret_cast<uint32_t*>(nodes + 32 * (parent_off + w * np + i));
            dst[k] = lo;
            dst[4 + k] = hi;
        }
        __syncthreads();
        if (i < np) {
            lvl[8 * i + k] = lo;
            lvl[8 * i + 4 + k] = hi;
        }
        __syncthreads();
        off = parent_off;
        width /= 2;
        seg = np;
    }
}

// the top (child level of at most 256 nodes) of every tree: one workgroup
// per tree, quad mode, each level staged in LDS
__global__ __launch_bounds__(1024) void merkle_top_multi_kernel(uint8_t* __restrict__ base,
                                                                const MerkleSeg* __restrict__ segs, uint32_t nseg) {
    __shared__ uint32_t lvl[256 * 8];
    const MerkleSeg g = find_seg(segs, nseg);
    uint8_t* nodes = base + g.nodes;
    uint64_t child_off = g.child_off, child_width = g.width;
    for (uint64_t i = threadIdx.x; i < child_width; i += blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(nodes + 32 * (child_off + i));
        uint4 a = src[0], b = src[1];
        uint32_t* d = lvl + 8 * i;
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
        d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    }
    __syncthreads();
    while (child_width > 1) {
        const uint64_t width = child_width / 2, parent_off = child_off + child_width;
        const uint32_t i = threadIdx.x >> 2, k = threadIdx.x & 3;
        uint32_t lo = 0, hi = 0;
        if (i < width) {
            uint32_t m[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) m[q] = lvl[16 * i + q];
            b3::hash64_quad(m, k, lo, hi);
            uint32_t* dst = reinterpret_cast<uint32_t*>(nodes + 32 * (parent_off + i));
            dst[k] = lo;
            dst[4 + k] = hi;
        }
        __syncthreads();
        if (i < width) {
            lvl[8 * i + k] = lo;
            lvl[8 * i + 4 + k] = hi;
        }
        __syncthreads();
        child_off = parent_off;
        child_width = width;
    }
}

__global__ __launch_bounds__(256) void gather_hashes_kernel(const uint8_t* __restrict__ nodes,
                                                            const uint64_t* __restrict__ idx, uint64_t count,
                                                            uint8_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4* s = reinterpret_cast<const uint4*>(nodes + 32 * idx[i]);
    uint4* d = reinterpret_cast<uint4*>(out + 32 * i);
    d[0] = s[0];
    d[1] = s[1];
}

__global__ __launch_bounds__(256) void copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                     uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void fill_zero_kernel(uint4* p, uint64_t n16) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) p[i] = make_uint4(0, 0, 0, 0);
}

unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void leaf_hash_launch(const uint64_t* matrix, uint64_t rows, uint32_t row_u64, uint8_t* leaves, hipStream_t st) {
    if (!rows) return;
    const dim3 g(blocks_for(rows, 256));
    switch (row_u64) {  // the row shapes of the prover's commits: base 16, Ext3 48
        case 16: hipLaunchKernelGGL(leaf_hash_blocks_kernel<2>, g, dim3(256), 0, st, matrix, rows, leaves); return;
        case 48: hipLaunchKernelGGL(leaf_hash_blocks_kernel<6>, g, dim3(256), 0, st, matrix, rows, leaves); return;
    }
    hipLaunchKernelGGL(leaf_hash_kernel, g, dim3(256), 0, st, matrix, rows, row_u64, leaves);
}

void hash_many_launch(const uint8_t* in, uint64_t size, uint64_t count, uint8_t* out, hipStream_t st) {
    if (!count) return;
    hipLaunchKernelGGL(hash_many_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, in, size, count, out);
}

void fill_zero_launch(uint8_t* p, uint64_t bytes, hipStream_t st) {
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks_for(n16, 256)), dim3(256), 0, st, reinterpret_cast<uint4*>(p), n16);
}

void copy_launch(void* dst, const void* src, uint64_t bytes, hipStream_t st) {
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    unsigned g = blocks_for(n16, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(copy16_kernel, dim3(g), dim3(256), 0, st, reinterpret_cast<uint4*>(dst),
                       reinterpret_cast<const uint4*>(src), n16);
}

void merkle_build_launch(uint8_t* nodes, uint64_t n_leaves, hipStream_t st, uint8_t* root_copy) {
    uint64_t width = 1;
    while (width < n_leaves) width <<= 1;
    if (n_leaves < width) fill_zero_launch(nodes + 32 * n_leaves, 32 * (width - n_leaves), st);
    if (n_leaves == 0) {
        fill_zero_launch(nodes, 32, st);  // empty tree: root is the zero hash
        return;
    }
    uint64_t off = 0;
    while (width >= 4 * TOP_WIDTH) {
        hipLaunchKernelGGL(merkle_2level_kernel, dim3(blocks_for(width / 4, 256)), dim3(256), 0, st, nodes, off, width);
        off += width + width / 2;
        width /= 4;
    }
    while (width > TOP_WIDTH) {
        uint64_t parents = width / 2;
        hipLaunchKernelGGL(merkle_level_kernel, dim3(blocks_for(parents, 256)), dim3(256), 0, st, nodes, off,
                           off + width, parents);
        off += width;
        width = parents;
    }
    if (width > 256) {  // 512 <= width <= TOP_WIDTH
        hipLaunchKernelGGL(merkle_top3_kernel, dim3((unsigned)(width / 256)), dim3(512), 0, st, nodes, off, width);
        off += width + width / 2 + width / 4;
        width /= 8;
    }
    if (width > 1)
        hipLaunchKernelGGL(merkle_top_kernel, dim3(1), dim3(TOP_WIDTH / 2), 0, st, nodes, off, width, root_copy);
}

int merkle_multi_plan(const uint64_t* nodes_off, const uint64_t* leaves, size_t count,
                      std::vector<MerkleSeg>& segs, std::vector<MerkleMultiLaunch>& launches) {
    segs.clear();
    launches.clear();
    // per tree: its current child level (offset, width); every tree runs the
    // steps of merkle_build_launch, one launch per step across the trees
    std::vector<uint64_t> off(count, 0), width(count);
    for (size_t k = 0; k < count; ++k) {
        if (!leaves[k] || (leaves[k] & (leaves[k] - 1))) return -1;  // powers of two only (no padding)
        width[k] = leaves[k];
    }
    auto add_launch = [&](uint32_t kind, auto take, auto blocks, auto advance) {
        MerkleMultiLaunch L{kind, (uint32_t)segs.size(), 0, 0};
        for (size_t k = 0; k < count; ++k) {
            if (!take(width[k])) continue;
            segs.push_back(MerkleSeg{nodes_off[k], off[k], width[k], L.blocks});
            L.blocks += blocks(width[k]);
            ++L.nseg;
            advance(off[k], width[k]);
        }
        if (L.nseg) launches.push_back(L);
        return L.nseg;
    };
    // two levels per step while the child level has at least 4 * TOP_WIDTH nodes
    while (add_launch(
        0, [](uint64_t w) { return w >= 4 * TOP_WIDTH; }, [](uint64_t w) { return (uint32_t)blocks_for(w / 4, 256); },
        [](uint64_t& o, uint64_t& w) {
            o += w + w / 2;
            w /= 4;
        })) {
    }
    while (add_launch(
        1, [](uint64_t w) { return w > TOP_WIDTH; }, [](uint64_t w) { return (uint32_t)blocks_for(w / 2, 256); },
        [](uint64_t& o, uint64_t& w) {
            o += w;
            w /= 2;
        })) {
    }
    add_launch(
        2, [](uint64_t w) { return w > 256; }, [](uint64_t w) { return (uint32_t)(w / 256); },
        [](uint64_t& o, uint64_t& w) {
            o += w + w / 2 + w / 4;
            w /= 8;
        });
    add_launch(
        3, [](uint64_t w) { return w > 1; }, [](uint64_t) { return 1u; }, [](uint64_t&, uint64_t& w) { w = 1; });
    return 0;
}

void merkle_multi_launch(uint8_t* base, const MerkleSeg* d_segs, const std::vector<MerkleMultiLaunch>& launches,
                         hipStream_t st) {
    for (const MerkleMultiLaunch& L : launches) {
        const MerkleSeg* t = d_segs + L.seg0;
        switch (L.kind) {
            case 0:
                hipLaunchKernelGGL(merkle_2level_multi_kernel, dim3(L.blocks), dim3(256), 0, st, base, t, L.nseg);
                break;
            case 1:
                hipLaunchKernelGGL(merkle_level_multi_kernel, dim3(L.blocks), dim3(256), 0, st, base, t, L.nseg);
                break;
            case 2:
                hipLaunchKernelGGL(merkle_top3_multi_kernel, dim3(L.blocks), dim3(512), 0, st, base, t, L.nseg);
                break;
            default:
                hipLaunchKernelGGL(merkle_top_multi_kernel, dim3(L.blocks), dim3(1024), 0, st, base, t, L.nseg);
                break;
        }
    }
}

void gather_hashes_launch(const uint8_t* nodes, const uint64_t* idx, uint64_t count, uint8_t* out, hipStream_t st) {
    if (!count) return;
    hipLaunchKernelGGL(gather_hashes_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, nodes, idx, count, out);
}

}  // namespace whir_gpu
