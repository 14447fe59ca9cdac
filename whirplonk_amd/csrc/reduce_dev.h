// reduce_dev.h — workgroup sums of Ext3 values and the second-stage sum of
// per-workgroup totals (shared by the sumcheck and prove kernels).
#pragma once
#include "gl.h"

namespace whir_gpu {
namespace red {

using gl::e3;

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor((uint32_t)v, m, 64);
    uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ e3 wave_sum(e3 v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        v.c0 = gl::add(v.c0, shfl_xor64(v.c0, m));
        v.c1 = gl::add(v.c1, shfl_xor64(v.c1, m));
        v.c2 = gl::add(v.c2, shfl_xor64(v.c2, m));
    }
    return v;
}

// Sum of NV Ext3 values per thread over the workgroup (<= 1024 threads);
// thread 0 gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(e3* v) {
    __shared__ e3 sh[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            e3 s = {0, 0, 0};
            for (int w = 0; w < nw; ++w) s = gl::e3_add(s, sh[k][w]);
            v[k] = s;
        }
    }
}

// After block_sum: thread 0 writes its block's totals to partials[NV*block + v].
template <int NV>
__device__ __forceinline__ void store_block_totals(const e3* acc, uint64_t* partials) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            uint64_t* q = partials + 3 * ((uint64_t)NV * blockIdx.x + v);
            q[0] = acc[v].c0;
            q[1] = acc[v].c1;
            q[2] = acc[v].c2;
        }
    }
}

// One workgroup: sums the NV-tuples of `count` blocks into out[0..3*NV)
// (out may be partials itself: every read is done before thread 0 writes).
// (A grid-wide "last workgroup reduces" variant needs a device-scope fence
// per workgroup, which costs more than this second launch across the XCDs'
// separate L2 caches.)
template <int NV>
__global__ __launch_bounds__(256) void sum_blocks_kernel(const uint64_t* partials, uint32_t count, uint64_t* out) {
    e3 tot[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) tot[v] = {0, 0, 0};
    for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const uint64_t* q = partials + 3 * ((uint64_t)NV * k + v);
            tot[v] = gl::e3_add(tot[v], e3{q[0], q[1], q[2]});
        }
    }
    block_sum<NV>(tot);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            out[3 * v] = tot[v].c0;
            out[3 * v + 1] = tot[v].c1;
            out[3 * v + 2] = tot[v].c2;
        }
    }
}

}  // namespace red
}  // namespace whir_gpu
