// lde.hip — interleaved Reed-Solomon encode (low-degree extension) and batched
// NTT for Goldilocks on gfx950.
//
// Computes, for every column c (one interleaved block of one polynomial,
// algebra/ntt/mod_ntt.hpp:49-130),
//     Y_c[i] = sum_{j<m} f_c[j] * w_n^{ij},   i < n = E*m,
// i.e. the size-n NTT of the zero-padded block, without touching the zeros.
//
// Two passes through HBM per coset of the output (four-step NTT), with the
// register/LDS sub-transforms of dft.h:
//   j = j2 + M2*j1 (j1 < M1, j2 < M2, m = M1*M2),  i = i1 + L*i2 (L = E*M1)
//   Y[i1 + L i2] = sum_{j2} w_M2^{i2 j2} * [ w_n^{i1 j2} * sum_{j1} f[j2 + M2 j1] w_L^{i1 j1} ]
// and i1 = s + E*k splits the inner length-L transform into E cosets:
//   sum_{j1} f[..] w_L^{i1 j1} = DFT_M1( f[j2 + M2 j1] * w_L^{s j1} )[k].
// Pass A (per coset s): column tiles of TA consecutive j2 (128-byte coalesced
// reads), DFT_M1 in LDS, inter-pass twiddle, write Z_s[k][c][j2].
// Pass B (per coset s): for each k, the C*M2 contiguous values of Z_s,
// DFT_M2 in LDS, write rows i = s + E k + L i2 of the output, all C columns
// of a row together (128-byte rows for C = 16).
// Only the m nonzero inputs are ever read; the scratch Z_s holds m*C values
// per coset in flight.
// Twiddles shared by a whole workgroup live in LDS: pass A's coset twist
// w_L^{s j1} depends on j1 only (shared by its TA columns), and when a pass-B
// workgroup owns a single i1 the inter-pass twiddle w_n^{i1 j2} is applied
// there (shared by all C columns) instead of at the end of pass A.
// With C = 16 columns the pass-B workgroup holds complete 128-byte rows, so
// the BLAKE3 row leaves (matrix_commit.hpp:169-195) are hashed right there,
// from LDS, before the rows are written out.
// Large shapes run persistent forms of both passes (lde_pass_a_regs,
// lde_pass_b_rows, lde_pass_b_cols): one workgroup per CU walks its tiles and
// keeps the next tile's loads in flight in registers while it transforms the
// current one, so the CU's VALU is not idle while a tile loads.
#include <algorithm>
#include <cstdlib>

#include "blake3_dev.h"
#include "dft.h"
#include "kernels.h"
#include "sha256_dev.h"

#include <map>
#include <mutex>

namespace whir_gpu {
namespace {

constexpr int THREADS = 512;
#ifndef LDE_THREADS_A
#define LDE_THREADS_A 512
#endif
#ifndef LDE_THREADS_B_SMALL
#define LDE_THREADS_B_SMALL 512
#endif
constexpr int THREADS_A = LDE_THREADS_A;

struct LdeArgs {
    const uint64_t* in;
    uint64_t* out;
    uint64_t* z;
    uint8_t* leaves;   // fused row leaves (pass B), or null
    const uint64_t* tw_lo;
    const uint64_t* tw_hi;
    const uint64_t* t1024;
    uint64_t C, m, n, M1, M2, E, L, zstride;
    uint32_t in_el, in_comp, out_el, out_comp;
    uint32_t TA, TB, CB, row_major;
    uint32_t tw_in_b;  // inter-pass twiddle applied in pass B (TB == 1)
    // split3: column c is component c % 3 of Ext3 column c / 3 (AoS input);
    // direct: M1 == 1, pass B reads the input itself (no pass A, no scratch)
    uint32_t split3, direct;
    // pass A: cosets per block (the input tile is reused for each) and the
    // number of cosets in this launch
    uint32_t cpb, ncos;
    // log2 of the power-of-two tile sizes: every index split is a shift/mask
    uint32_t lgTA, lgTB, lgCB, lgNJ, lgNK;
};

constexpr int FUSE_COLS = 16;  // row width (u64) of the fused leaf hashing

// Leaf hash of one 16-limb (128-byte) row whose limb c is ld(c), with the
// context's engine: BLAKE3 (one chunk of two blocks, CHUNK_START then
// CHUNK_END | ROOT, matrix_commit.hpp:169-195) or SHA-256 (two data blocks
// and the constant padding block of a 128-byte message, sha2_engine.hpp:47-81).
// out: the 32 digest bytes as 8 little-endian words.
template <uint32_t HASH, typename Load>
__device__ __forceinline__ void row_leaf(Load ld, uint32_t out[8]) {
    if constexpr (HASH == HASH_SHA2) {
        uint32_t st[8], w[16];
        s256::init(st);
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) {
                const uint64_t x = ld(8 * blk + cc);
                w[2 * cc] = s256::bswap((uint32_t)x);
                w[2 * cc + 1] = s256::bswap((uint32_t)(x >> 32));
            }
            s256::compress(st, w);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) w[i] = 0;
        w[0] = 0x80000000u;
        w[15] = 128 * 8;  // message length in bits
        s256::compress(st, w);
        s256::digest(st, out);
    } else {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 8; ++i) out[i] = b3::iv(i);
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) {
                const uint64_t x = ld(8 * blk + cc);
                m[2 * cc] = (uint32_t)x;
                m[2 * cc + 1] = (uint32_t)(x >> 32);
            }
            b3::compress(out, m, 0, 64, blk == 0 ? b3::CHUNK_START : (b3::CHUNK_END | b3::ROOT), out);
        }
    }
}

// swizzled position of (row, col) in the 16-wide LDS row stage: rows are
// read one per lane, so columns rotate with the row to spread the banks
__device__ __forceinline__ int stage_idx(int row, int col) { return row * FUSE_COLS + (col ^ (row & 15)); }

// input index of coefficient j of column c
__device__ __forceinline__ uint64_t in_index(const LdeArgs& a, uint64_t c, uint64_t j) {
    if (a.split3) return ((c / 3) * a.m + j) * 3 + (c % 3);
    return (c * a.m + j) * a.in_el + a.in_comp;
}

// w_n^e for e < n from the two-level table
__device__ __forceinline__ uint64_t twid(const LdeArgs& a, uint64_t e) {
    uint64_t x = a.tw_lo[e & 2047];
    uint64_t h = e >> 11;
    return h ? gl::mul(x, a.tw_hi[h]) : x;
}

template <int M1, bool INV>
__global__ __launch_bounds__(THREADS_A) void lde_pass_a(LdeArgs a, uint64_t s0) {
    using S = dft::Split<M1>;
    constexpr int R1 = S::R1, R2 = S::R2, PAD = S::PADDED + 1;
    extern __shared__ uint64_t lds[];
    const uint32_t TA = a.TA, lgTA = a.lgTA;
    const uint32_t c = blockIdx.x >> a.lgNJ;
    const uint32_t J = (blockIdx.x & ((1u << a.lgNJ) - 1)) << lgTA;
    const uint64_t* col = a.in + in_index(a, c, J);
    const uint32_t in_el = a.in_el, M2 = (uint32_t)a.M2, zrow = (uint32_t)(a.C * a.M2);
    uint64_t* twist = lds + TA * PAD;  // w_L^{s j1} = w_n^{s j1 M2}, shared by the TA columns
    // The block transforms its input tile (TA columns x M1 values) for each of
    // its a.cpb cosets; the re-reads of the tile hit the cache.
    const uint32_t nq = TA * M1;
    for (uint32_t g = 0; g < a.cpb; ++g) {
        const uint32_t zi = blockIdx.y * a.cpb + g;  // coset slot of this launch
        if (zi >= a.ncos) break;                    // uniform over the block
        const uint64_t s = s0 + zi;
        uint64_t* z = a.z + zi * a.zstride + (uint64_t)c * a.M2 + J;  // Z[k][c][J + t] = z[k*C*M2 + t]
        if (s != 0) {
            for (uint32_t j1 = threadIdx.x; j1 < M1; j1 += THREADS_A) twist[j1] = twid(a, (s * j1 * a.M2) & (a.n - 1));
            __syncthreads();
        }
        for (uint32_t q = threadIdx.x; q < nq; q += THREADS_A) {
            uint32_t t = q & (TA - 1), j1 = q >> lgTA;
            uint64_t x = col[(t + M2 * j1) * in_el];
            if (s != 0) x = gl::mul(x, twist[j1]);
            lds[t * PAD + dft::pidx<M1>(j1)] = x;
        }
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < TA * R2; w += THREADS_A) {
            uint32_t n2 = w % R2, t = w / R2;
            dft::lds_dft_step1<M1, INV>(lds + t * PAD, n2, a.t1024);
        }
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < TA * R1; w += THREADS_A) {
            uint32_t t = w & (TA - 1), k1 = w >> lgTA;
            uint64_t v[R2];
            dft::lds_dft_step2<M1, INV>(lds + t * PAD, k1, v);
            const uint64_t j2 = J + t;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                uint32_t k = k1 + R1 * k2;
                uint64_t zv = v[k2];
                if (!a.tw_in_b) {
                    uint64_t e = ((s + a.E * k) * j2) & (a.n - 1);
                    if (e) zv = gl::mul(zv, twid(a, e));
                }
                z[k * zrow + t] = zv;
            }
        }
        __syncthreads();  // the next coset rewrites the LDS tile and twist table
    }
}

#ifndef LDE_PA_PERSIST
#define LDE_PA_PERSIST 1
#endif

// Pass A for M1 >= 16 when the tile fills the workgroup (TA * R2 threads).
// Persistent: workgroup b walks the tiles b, b + gridDim.x, ... (tile = column
// c, first j2 J, as blockIdx.x of lde_pass_a) and runs every coset of the
// launch on each.  Lane (t, n2) = (tid % TA, tid / TA) holds its step-1
// inputs x[t][R2*n1 + n2] in registers for all cosets (the tile is read from
// HBM once, with no LDS staging pass), and issues the next tile's loads once
// the last coset's step 1 is done, so they are in flight during its step 2.
template <int M1, bool INV>
__global__ __launch_bounds__(THREADS_A) void lde_pass_a_regs(LdeArgs a, uint64_t s0, uint32_t ntiles) {
    using S = dft::Split<M1>;
    constexpr int R1 = S::R1, R2 = S::R2, PAD = S::PADDED + 1, TA = THREADS_A / R2;
    static_assert(R1 <= R2 && TA * R2 == THREADS_A, "step-2 items must fit the workgroup");
    extern __shared__ uint64_t lds[];
    uint64_t* twist = lds + TA * PAD;  // w_L^{s j1}, shared by the TA columns
    const uint32_t M2 = (uint32_t)a.M2, in_el = a.in_el, zrow = (uint32_t)(a.C * a.M2);
    const uint32_t jmask = (1u << a.lgNJ) - 1;
    uint64_t pre[R1];
    auto fetch = [&](uint32_t tile, uint32_t tid) {
        const uint64_t* col = a.in + in_index(a, tile >> a.lgNJ, (tile & jmask) * TA);  // uniform
        const uint32_t t = tid % TA, n2 = tid / TA;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) pre[n1] = col[(t + M2 * (R2 * n1 + n2)) * in_el];
    };
    uint32_t tile = blockIdx.x;
    if (tile < ntiles) fetch(tile, threadIdx.x);
    for (; tile < ntiles; tile += gridDim.x) {
        const uint32_t c = tile >> a.lgNJ, J = (tile & jmask) * TA;
        const bool more = tile + gridDim.x < ntiles;
        for (uint32_t g = 0; g < a.ncos; ++g) {
            // lane indices from an opaque copy of the thread id, every coset
            // (see lde_pass_b_rows)
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            const uint32_t t = tid % TA, n2 = tid / TA;
            const uint64_t s = s0 + g;
            uint64_t* z = a.z + g * a.zstride + (uint64_t)c * M2 + J;  // Z[k][c][J + t] = z[k*C*M2 + t]
            if (s != 0) {
                for (uint32_t j1 = tid; j1 < M1; j1 += THREADS_A) twist[j1] = twid(a, (s * j1 * M2) & (a.n - 1));
                __syncthreads();
            }
            uint64_t v[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) v[n1] = s != 0 ? gl::mul(pre[n1], twist[R2 * n1 + n2]) : pre[n1];
            dft::dft_reg<R1, INV>(v);
            if constexpr (R2 > 1) {
#pragma unroll
                for (int k1 = 1; k1 < R1; ++k1) v[k1] = gl::mul(v[k1], a.t1024[((n2 * k1) * (1024 / M1)) & 1023]);
            }
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) lds[t * PAD + dft::pidx<M1>(R2 * k1 + n2)] = v[k1];
            if (g + 1 == a.ncos && more) fetch(tile + gridDim.x, tid);
            __syncthreads();
            for (uint32_t w = tid; w < TA * R1; w += THREADS_A) {
                const uint32_t t2 = w % TA, k1 = w / TA;
                uint64_t u[R2];
                dft::lds_dft_step2<M1, INV>(lds + t2 * PAD, k1, u);
                const uint64_t j2 = J + t2;
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {
                    uint32_t k = k1 + R1 * k2;
                    uint64_t zv = u[k2];
                    if (!a.tw_in_b) {
                        uint64_t e = ((s + a.E * k) * j2) & (a.n - 1);
                        if (e) zv = gl::mul(zv, twid(a, e));
                    }
                    z[k * zrow + t2] = zv;
                }
            }
            __syncthreads();  // the next coset rewrites the LDS tile and twist table
        }
    }
}

template <int M2, bool INV>
__global__ __launch_bounds__(THREADS) void lde_pass_b(LdeArgs a, uint64_t s0) {
    using S = dft::Split<M2>;
    constexpr int R1 = S::R1, R2 = S::R2, PAD = S::PADDED + 1;
    constexpr int DFT_WORDS = 16 * PAD, STAGE_WORDS = FUSE_COLS * M2;
    extern __shared__ uint64_t lds[];
    const uint32_t TB = a.TB, CB = a.CB, DB = TB * CB, lgTB = a.lgTB, lgCB = a.lgCB;
    const uint32_t K0 = (blockIdx.x & ((1u << a.lgNK) - 1)) << lgTB;
    const uint32_t C0 = (blockIdx.x >> a.lgNK) << lgCB;
    const uint64_t s = s0 + blockIdx.y;
    const uint32_t CM2 = (uint32_t)(a.C * M2);
    const uint64_t* z = a.z + blockIdx.y * a.zstride + (uint64_t)K0 * CM2 + (uint64_t)C0 * M2;
    const bool fuse = a.leaves != nullptr;
    // inter-pass twiddles w_n^{i1 j2} of this workgroup's single i1
    uint64_t* twt = lds + (fuse ? (DFT_WORDS > STAGE_WORDS ? DFT_WORDS : STAGE_WORDS) : DB * PAD);

    if (a.tw_in_b) {
        const uint64_t i1 = s + a.E * K0;
        for (uint32_t j2 = threadIdx.x; j2 < M2; j2 += blockDim.x) twt[j2] = twid(a, (i1 * j2) & (a.n - 1));
        __syncthreads();
    }
    for (uint32_t q = threadIdx.x; q < DB * M2; q += blockDim.x) {
        uint32_t j2 = q % M2, d = q / M2;
        uint32_t tb = a.row_major ? d >> lgCB : d & (TB - 1);
        uint32_t cb = a.row_major ? d & (CB - 1) : d >> lgTB;
        uint64_t v = a.direct ? a.in[in_index(a, C0 + cb, j2)] : z[tb * CM2 + cb * M2 + j2];
        if (a.tw_in_b) v = gl::mul(v, twt[j2]);
        lds[d * PAD + dft::pidx<M2>(j2)] = v;
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < DB * R2; w += blockDim.x) {
        uint32_t n2 = w % R2, d = w / R2;
        dft::lds_dft_step1<M2, INV>(lds + d * PAD, n2, a.t1024);
    }
    __syncthreads();
    if (!fuse) {
        for (uint32_t w = threadIdx.x; w < DB * R1; w += blockDim.x) {
            uint32_t d = w & (DB - 1), k1 = w >> (lgTB + lgCB);
            uint32_t tb = a.row_major ? d >> lgCB : d & (TB - 1);
            uint32_t cb = a.row_major ? d & (CB - 1) : d >> lgTB;
            uint64_t v[R2];
            dft::lds_dft_step2<M2, INV>(lds + d * PAD, k1, v);
            const uint64_t i1 = s + a.E * (K0 + tb);
            const uint64_t c = C0 + cb;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                uint64_t i = i1 + a.L * (k1 + (uint64_t)R1 * k2);
                uint64_t idx = a.row_major ? (i * a.C + c) : (c * a.n + i);
                a.out[idx * a.out_el + a.out_comp] = v[k2];
            }
        }
        return;
    }
    // Fused leaves: this workgroup owns rows i = i1 + L*i2 (i2 < M2) with all
    // 16 columns (TB = 1, CB = C = 16).  Stage them row-major in LDS, hash one
    // row per lane, then write the rows out coalesced.
    constexpr int NW = (16 * R1 + THREADS - 1) / THREADS;
    uint64_t v[NW][R2];
#pragma unroll
    for (int it = 0; it < NW; ++it) {
        uint32_t w = threadIdx.x + it * THREADS;
        if (w < 16 * R1) dft::lds_dft_step2<M2, INV>(lds + (w % 16) * PAD, w / 16, v[it]);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NW; ++it) {
        uint32_t w = threadIdx.x + it * THREADS;
        if (w < 16 * R1) {
            const int cb = w % 16, k1 = w / 16;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                int row = k1 + R1 * k2;
                lds[stage_idx(row, cb)] = v[it][k2];
            }
        }
    }
    __syncthreads();
    const uint64_t i1 = s + a.E * K0;
    for (uint32_t r = threadIdx.x; r < M2; r += blockDim.x) {
        uint32_t cv[8], m[16];
#pragma unroll
        for (int i = 0; i < 8; ++i) cv[i] = b3::iv(i);
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                uint64_t x = lds[stage_idx(r, 8 * blk + c)];
                m[2 * c] = (uint32_t)x;
                m[2 * c + 1] = (uint32_t)(x >> 32);
            }
            b3::compress(cv, m, 0, 64, blk == 0 ? b3::CHUNK_START : (b3::CHUNK_END | b3::ROOT), cv);
        }
        uint64_t i = i1 + a.L * r;
        uint4* dst = reinterpret_cast<uint4*>(a.leaves + 32 * i);
        dst[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
        dst[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
    }
    for (uint32_t q = threadIdx.x; q < 16 * M2; q += blockDim.x) {
        uint32_t r = q / 16, c = q % 16;
        a.out[(i1 + a.L * r) * 16 + c] = lds[stage_idx(r, c)];
    }
}

#ifndef LDE_PB_PERSIST
#define LDE_PB_PERSIST 1
#endif

// Pass B of a 16-column base codeword with M2 = 1024 and the row leaves fused
// (every commit of a polynomial with m >= 2048 coefficients per column).
// Persistent: workgroup b walks the tiles b, b + gridDim.x, ... (tile = coset
// slot g * M1 + i1 index k) and, while it transforms and hashes one tile,
// holds the loads of its next tile in flight in registers.  Thread
// (c, n2) = (tid / 32, tid % 32) loads the step-1 inputs of column c straight
// into registers, so the input needs no LDS staging pass; LDS carries the
// step-1 -> step-2 transposition and the row-major staging of the leaves.
// LDS slot of (row r, column c) of a 16-column, 1024-row tile in the padded
// DFT buffer (column stride 1057): the slot that step-2 item (c, r % 32)
// read its input n2 = r / 32 from.  A wave reading 64 consecutive rows of one
// column hits every bank pair four times (1057 and 33 are 1 mod 16).
__device__ __forceinline__ int rows_idx(int r, int c) { return c * 1057 + 33 * (r & 31) + (r >> 5); }

template <bool INV, uint32_t HASH = HASH_BLAKE3>
__global__ __launch_bounds__(THREADS) void lde_pass_b_rows(LdeArgs a, uint64_t s0, uint32_t ntiles) {
    constexpr int M2 = 1024, R = 32, PAD = M2 + M2 / R + 1;
    static_assert(PAD == 1057, "rows_idx assumes the padded 1024-point layout");
    static_assert(THREADS == 16 * R, "one thread per (column, n2)");
    extern __shared__ uint64_t lds[];
    // inter-pass twiddles of this tile and the next (two tables past the DFT
    // buffer): the next tile's are computed while this one is transformed
    uint64_t* twt2 = lds + 16 * PAD;
    const uint32_t kmask = (1u << a.lgNK) - 1;
    uint64_t pre[R];
    uint32_t off = (threadIdx.x / R) * M2 + threadIdx.x % R;
    auto fetch = [&](uint32_t t) {
        // uniform tile base + 32-bit lane offset: scalar-base addressing
        const uint64_t* zb = a.z + (uint64_t)(t >> a.lgNK) * a.zstride + (uint64_t)(t & kmask) * (16 * M2);
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) pre[n1] = zb[off + R * n1];
    };
    auto tile_i1 = [&](uint32_t t) { return s0 + (t >> a.lgNK) + a.E * (t & kmask); };
    auto twiddles = [&](uint32_t t, uint64_t* dst, uint32_t tid) {
        const uint64_t i1 = tile_i1(t);
        for (uint32_t j2 = tid; j2 < M2; j2 += THREADS) dst[j2] = twid(a, (i1 * j2) & (a.n - 1));
    };
    uint32_t tile = blockIdx.x;
    if (tile < ntiles) {
        fetch(tile);
        twiddles(tile, twt2, threadIdx.x);
    }
    __syncthreads();
    for (uint32_t it = 0; tile < ntiles; tile += gridDim.x, ++it) {
        // The lane indices are re-derived every tile from an opaque copy of
        // the thread id: hoisted out of the loop, the per-lane twiddle and
        // stage addresses they feed would stay live and spill.
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const uint32_t c = tid / R, n2 = tid % R;
        const uint64_t i1 = tile_i1(tile);
        const uint64_t* twt = twt2 + (it & 1) * M2;
        uint64_t v[R];
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = gl::mul(pre[n1], twt[n2 + R * n1]);
        dft::dft_reg<R, INV>(v);
#pragma unroll
        for (int k1 = 1; k1 < R; ++k1) v[k1] = gl::mul(v[k1], a.t1024[(n2 * k1) & 1023]);
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) lds[c * PAD + dft::pidx<M2>(R * k1 + n2)] = v[k1];
        // the next tile's loads stay in flight through step 2 and the hashing,
        // and its twiddles go to the other table (last read two tiles ago)
        if (tile + gridDim.x < ntiles) {
            fetch(tile + gridDim.x);
            twiddles(tile + gridDim.x, twt2 + ((it + 1) & 1) * M2, tid);
        }
        __syncthreads();
        // Step 2: item (cb, k1) reads its 32 values buf[cb][33 k1 + n2] and
        // writes row k1 + 32 k2 of column cb back into the same 32 slots
        // (rows_idx), so no other item's input is touched and no barrier is
        // needed between the reads and the writes.
        const uint32_t cb = tid % 16, k1 = tid / 16;
        dft::lds_dft_step2<M2, INV>(lds + cb * PAD, k1, v);
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) lds[rows_idx(k1 + R * k2, cb)] = v[k2];
        __syncthreads();
#pragma unroll
        for (int h = 0; h < M2 / THREADS; ++h) {
            const uint32_t r = tid + h * THREADS;
            uint32_t cv[8];
            row_leaf<HASH>([&](int c) { return lds[rows_idx(r, c)]; }, cv);
            uint4* dst = reinterpret_cast<uint4*>(a.leaves + 32 * (i1 + a.L * r));
            dst[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
            dst[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
        }
        // rows out, two columns (16 bytes) per store: 8 lanes per 128-byte row
#pragma unroll
        for (int h = 0; h < 16 * M2 / (2 * THREADS); ++h) {
            const uint32_t q = tid + h * THREADS;
            const uint32_t r = q / 8, cp = 2 * (q % 8);
            const uint64_t x0 = lds[rows_idx(r, cp)], x1 = lds[rows_idx(r, cp + 1)];
            *reinterpret_cast<ulonglong2*>(a.out + (i1 + a.L * r) * 16 + cp) = make_ulonglong2(x0, x1);
        }
        __syncthreads();  // the next tile rewrites the twiddles and the stage
    }
}

// Pass B of a 16-column base codeword with M2 = 1024 and the row leaves fused,
// at four waves per SIMD: two 512-lane workgroups share a CU, each with a
// 66 KB LDS buffer (half a tile) and at most 128 VGPRs, instead of one
// workgroup holding the whole 16 x 1024 tile in LDS.  The step-1 -> step-2
// transposition runs in two rounds by the top bit of n2, and the rows are
// staged for the leaf hashes and the stores in two halves of 512 rows.  While
// one workgroup waits at a barrier or for its loads, the other computes (no
// register prefetch of the next tile).
//   step 1: lane (c, n2), c = (tid >> 4) & 15, n2 = (tid & 15) | ((tid >> 8) << 4)
//   step 2: lane (cb, k1) = (tid % 16, tid / 16)
// Exchange slot of (column c, k1, n2 mod 16): c * 519 + 16 k1 + n2 % 16 (at
// most 3-way bank aliasing for both the writes and the reads).
__device__ __forceinline__ int xch_idx(int c, int k1, int n2l) { return c * 519 + 16 * k1 + n2l; }
constexpr int XCH_WORDS = 16 * 519;
static_assert(XCH_WORDS >= 512 * FUSE_COLS, "the row stage reuses the exchange buffer");

template <bool INV, uint32_t HASH = HASH_BLAKE3>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void lde_pass_b_rows2(
    LdeArgs a, uint64_t s0, uint32_t ntiles) {
    constexpr int M2 = 1024, R = 32;
    static_assert(THREADS == 16 * R, "one thread per (column, n2)");
    extern __shared__ uint64_t lds[];
    uint64_t* buf = lds;               // exchange, then row stage
    uint64_t* twt = lds + XCH_WORDS;   // inter-pass twiddles of the tile
    const uint32_t kmask = (1u << a.lgNK) - 1;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // lane indices re-derived per tile (see lde_pass_b_rows)
        const uint64_t i1 = s0 + (tile >> a.lgNK) + a.E * (tile & kmask);
        const uint32_t c = (tid >> 4) & 15, hw = tid >> 8, n2 = (tid & 15) | (hw << 4);
        const uint64_t* zb = a.z + (uint64_t)(tile >> a.lgNK) * a.zstride + (uint64_t)(tile & kmask) * (16 * M2);
        uint64_t v[R];
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = zb[c * M2 + n2 + R * n1];
        // the previous tile's last reads of twt and buf ended at its final barrier
        for (uint32_t j2 = tid; j2 < M2; j2 += THREADS) twt[j2] = twid(a, (i1 * j2) & (a.n - 1));
        __syncthreads();
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = gl::mul(v[n1], twt[n2 + R * n1]);
        dft::dft_reg<R, INV>(v);
#pragma unroll
        for (int k1 = 1; k1 < R; ++k1) v[k1] = gl::mul(v[k1], a.t1024[(n2 * k1) & 1023]);
        // transposition: in round h the lanes with n2 >> 4 == h (waves 4h..4h+3)
        // write all 32 of their values and every lane reads its 16 inputs with
        // n2 >> 4 == h
        const uint32_t cb = tid % 16, k1r = tid / 16;
        uint64_t u[R];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (hw == (uint32_t)h) {  // uniform per wave
#pragma unroll
                for (int k1 = 0; k1 < R; ++k1) buf[xch_idx(c, k1, n2 & 15)] = v[k1];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) u[16 * h + q] = buf[xch_idx(cb, k1r, q)];
            __syncthreads();
        }
        dft::dft_reg<R, INV>(u);  // u[k2] = row k1r + 32 k2 of column cb
        // rows 512 h .. 512 h + 511 staged row-major, hashed one per lane, stored
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) buf[stage_idx(k1r + R * kk, cb)] = u[16 * h + kk];
            __syncthreads();
            {
                const uint32_t r = tid;  // local row
                uint32_t cv[8];
                row_leaf<HASH>([&](int c) { return buf[stage_idx(r, c)]; }, cv);
                uint4* dst = reinterpret_cast<uint4*>(a.leaves + 32 * (i1 + a.L * (512 * h + r)));
                dst[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
                dst[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
            }
#pragma unroll
            for (int q2 = 0; q2 < 512 * 16 / (2 * THREADS); ++q2) {
                const uint32_t q = tid + q2 * THREADS;
                const uint32_t r = q / 8, cp = 2 * (q % 8);
                const uint64_t x0 = buf[stage_idx(r, cp)], x1 = buf[stage_idx(r, cp + 1)];
                *reinterpret_cast<ulonglong2*>(a.out + (i1 + a.L * (512 * h + r)) * 16 + cp) = make_ulonglong2(x0, x1);
            }
            __syncthreads();  // the next half (or tile) rewrites the stage
        }
    }
}

// Pass B of an Ext3 codeword of 16 Ext3 columns (split3: 48 base columns, a
// row is 384 bytes = six BLAKE3 blocks) with M2 = 1024 and the row leaves
// fused.  A tile (coset slot, i1) is transformed as three sub-tiles of 16
// base columns each, in order, like lde_pass_b_rows's tile; sub-tile q's
// columns 16q .. 16q + 15 are bytes 128q .. 128q + 127 of every row, i.e.
// blocks 2q and 2q + 1 of its leaf, so each lane carries the chaining values
// of its two rows in registers from one sub-tile to the next (BLAKE3: one
// chunk, CHUNK_START on block 0, CHUNK_END | ROOT on block 5, counter 0;
// SHA-256: the state over the six data blocks, then the padding block).  The
// next sub-tile's loads are in flight during the current one's step 2,
// hashing and stores; the inter-pass twiddles are computed once per tile.
template <bool INV, uint32_t HASH = HASH_BLAKE3>
__global__ __launch_bounds__(THREADS) void lde_pass_b_rows_e3(LdeArgs a, uint64_t s0, uint32_t ntiles) {
    constexpr int M2 = 1024, R = 32, PAD = M2 + M2 / R + 1, SUB = 3, COLS = 3 * FUSE_COLS;
    static_assert(PAD == 1057, "rows_idx assumes the padded 1024-point layout");
    extern __shared__ uint64_t lds[];
    uint64_t* twt2 = lds + 16 * PAD;  // inter-pass twiddles of this tile and the next
    const uint32_t kmask = (1u << a.lgNK) - 1;
    uint64_t pre[R];
    const uint32_t nwork = ntiles * SUB;  // (tile, sub-tile) steps of the whole launch
    auto tile_of = [&](uint32_t w) { return blockIdx.x + (w / SUB) * gridDim.x; };
    auto fetch = [&](uint32_t tile, uint32_t sub, uint32_t tid) {
        const uint64_t* zb = a.z + (uint64_t)(tile >> a.lgNK) * a.zstride + (uint64_t)(tile & kmask) * (COLS * M2) +
                             (uint64_t)sub * (FUSE_COLS * M2);
        const uint32_t off = (tid / R) * M2 + tid % R;
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) pre[n1] = zb[off + R * n1];
    };
    auto tile_i1 = [&](uint32_t t) { return s0 + (t >> a.lgNK) + a.E * (t & kmask); };
    auto twiddles = [&](uint32_t t, uint64_t* dst, uint32_t tid) {
        const uint64_t i1 = tile_i1(t);
        for (uint32_t j2 = tid; j2 < M2; j2 += THREADS) dst[j2] = twid(a, (i1 * j2) & (a.n - 1));
    };
    // this workgroup's steps: w = SUB * (local tile number) + sub-tile
    uint32_t w = 0;
    if (tile_of(w) < ntiles) {
        fetch(tile_of(w), 0, threadIdx.x);
        twiddles(tile_of(w), twt2, threadIdx.x);
    }
    __syncthreads();
    uint32_t cv[2][8];
    for (; tile_of(w) < ntiles && w < nwork; ++w) {
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // lane indices re-derived per step (see lde_pass_b_rows)
        const uint32_t tile = tile_of(w), sub = w % SUB, it = w / SUB;
        const uint32_t c = tid / R, n2 = tid % R;
        const uint64_t i1 = tile_i1(tile);
        const uint64_t* twt = twt2 + (it & 1) * M2;
        uint64_t v[R];
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = gl::mul(pre[n1], twt[n2 + R * n1]);
        dft::dft_reg<R, INV>(v);
#pragma unroll
        for (int k1 = 1; k1 < R; ++k1) v[k1] = gl::mul(v[k1], a.t1024[(n2 * k1) & 1023]);
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) lds[c * PAD + dft::pidx<M2>(R * k1 + n2)] = v[k1];
        // the next step's loads (and, starting a new tile, its twiddles into
        // the other table, last read two tiles ago)
        if (tile_of(w + 1) < ntiles) {
            fetch(tile_of(w + 1), (w + 1) % SUB, tid);
            if (sub == SUB - 1) twiddles(tile_of(w + 1), twt2 + ((it + 1) & 1) * M2, tid);
        }
        __syncthreads();
        {
            const uint32_t cb = tid % 16, k1 = tid / 16;
            dft::lds_dft_step2<M2, INV>(lds + cb * PAD, k1, v);
#pragma unroll
            for (int k2 = 0; k2 < R; ++k2) lds[rows_idx(k1 + R * k2, cb)] = v[k2];
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < M2 / THREADS; ++h) {
            const uint32_t r = tid + h * THREADS;
            if constexpr (HASH == HASH_SHA2) {
                // SHA-256 of the 384-byte row: six data blocks (two per
                // sub-tile, the limbs' bytes as big-endian words), then the
                // constant padding block of a 3072-bit message
                if (sub == 0) s256::init(cv[h]);
#pragma unroll
                for (int blk = 0; blk < 2; ++blk) {
                    uint32_t wd[16];
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        const uint64_t x = lds[rows_idx(r, 8 * blk + cc)];
                        wd[2 * cc] = s256::bswap((uint32_t)x);
                        wd[2 * cc + 1] = s256::bswap((uint32_t)(x >> 32));
                    }
                    s256::compress(cv[h], wd);
                }
                if (sub == SUB - 1) {
                    uint32_t wd[16];
#pragma unroll
                    for (int i = 0; i < 16; ++i) wd[i] = 0;
                    wd[0] = 0x80000000u;
                    wd[15] = 384 * 8;
                    s256::compress(cv[h], wd);
                    uint32_t out[8];
                    s256::digest(cv[h], out);
                    uint4* dst = reinterpret_cast<uint4*>(a.leaves + 32 * (i1 + a.L * r));
                    dst[0] = make_uint4(out[0], out[1], out[2], out[3]);
                    dst[1] = make_uint4(out[4], out[5], out[6], out[7]);
                }
            } else {
                if (sub == 0) {
#pragma unroll
                    for (int i = 0; i < 8; ++i) cv[h][i] = b3::iv(i);
                }
#pragma unroll
                for (int blk = 0; blk < 2; ++blk) {
                    uint32_t m[16];
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        const uint64_t x = lds[rows_idx(r, 8 * blk + cc)];
                        m[2 * cc] = (uint32_t)x;
                        m[2 * cc + 1] = (uint32_t)(x >> 32);
                    }
                    const uint32_t flags = (sub == 0 && blk == 0)
                                               ? b3::CHUNK_START
                                               : ((sub == SUB - 1 && blk == 1) ? (b3::CHUNK_END | b3::ROOT) : 0u);
                    b3::compress(cv[h], m, 0, 64, flags, cv[h]);
                }
                if (sub == SUB - 1) {
                    uint4* dst = reinterpret_cast<uint4*>(a.leaves + 32 * (i1 + a.L * r));
                    dst[0] = make_uint4(cv[h][0], cv[h][1], cv[h][2], cv[h][3]);
                    dst[1] = make_uint4(cv[h][4], cv[h][5], cv[h][6], cv[h][7]);
                }
            }
        }
        // this sub-tile's 16 columns of every row: 8 lanes per 128-byte segment
#pragma unroll
        for (int h = 0; h < 16 * M2 / (2 * THREADS); ++h) {
            const uint32_t q = tid + h * THREADS;
            const uint32_t r = q / 8, cp = 2 * (q % 8);
            const uint64_t x0 = lds[rows_idx(r, cp)], x1 = lds[rows_idx(r, cp + 1)];
            *reinterpret_cast<ulonglong2*>(a.out + (i1 + a.L * r) * COLS + FUSE_COLS * sub + cp) =
                make_ulonglong2(x0, x1);
        }
        __syncthreads();  // the next step rewrites the LDS tile (and a twiddle table)
    }
}

// Pass B with M2 = 1024 and 16 Z columns per tile (TB * CB = 16), no fused
// leaves: Ext3 codewords (split3, 48 base columns), column-major output and
// the plain encode.  Persistent with the next tile's loads in flight, as
// lde_pass_b_rows; tile = coset slot g * nb + block index bx of lde_pass_b.
template <bool INV>
__global__ __launch_bounds__(THREADS) void lde_pass_b_cols(LdeArgs a, uint64_t s0, uint32_t nb, uint32_t ntiles) {
    constexpr int M2 = 1024, R = 32, PAD = M2 + M2 / R + 1;
    static_assert(THREADS == 16 * R, "one thread per (Z column, n2)");
    extern __shared__ uint64_t lds[];
    uint64_t* twt = lds + 16 * PAD;
    const uint32_t TB = a.TB, CB = a.CB, lgTB = a.lgTB, lgCB = a.lgCB;
    const uint32_t CM2 = (uint32_t)(a.C * M2);
    uint64_t pre[R];
    auto split = [&](uint32_t d, uint32_t& tb, uint32_t& cb) {
        tb = a.row_major ? d >> lgCB : d & (TB - 1);
        cb = a.row_major ? d & (CB - 1) : d >> lgTB;
    };
    auto fetch = [&](uint32_t t, uint32_t tid) {
        const uint32_t g = t / nb, bx = t - g * nb;
        const uint32_t K0 = (bx & ((1u << a.lgNK) - 1)) << lgTB, C0 = (bx >> a.lgNK) << lgCB;
        const uint64_t* zb = a.z + (uint64_t)g * a.zstride + (uint64_t)K0 * CM2 + (uint64_t)C0 * M2;
        uint32_t tb, cb;
        split(tid / R, tb, cb);
        const uint32_t off = tb * CM2 + cb * M2 + tid % R;
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) pre[n1] = zb[off + R * n1];
    };
    uint32_t tile = blockIdx.x;
    if (tile < ntiles) fetch(tile, threadIdx.x);
    for (; tile < ntiles; tile += gridDim.x) {
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // lane indices re-derived per tile (see lde_pass_b_rows)
        const uint32_t d = tid / R, n2 = tid % R;
        const uint32_t g = tile / nb, bx = tile - g * nb;
        const uint32_t K0 = (bx & ((1u << a.lgNK) - 1)) << lgTB, C0 = (bx >> a.lgNK) << lgCB;
        const uint64_t s = s0 + g;
        if (a.tw_in_b) {
            const uint64_t i1 = s + a.E * K0;
            for (uint32_t j2 = tid; j2 < M2; j2 += THREADS) twt[j2] = twid(a, (i1 * j2) & (a.n - 1));
            __syncthreads();
        }
        uint64_t v[R];
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = a.tw_in_b ? gl::mul(pre[n1], twt[n2 + R * n1]) : pre[n1];
        dft::dft_reg<R, INV>(v);
#pragma unroll
        for (int k1 = 1; k1 < R; ++k1) v[k1] = gl::mul(v[k1], a.t1024[(n2 * k1) & 1023]);
#pragma unroll
        for (int k1 = 0; k1 < R; ++k1) lds[d * PAD + dft::pidx<M2>(R * k1 + n2)] = v[k1];
        if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x, tid);
        __syncthreads();
        {
            const uint32_t d2 = tid % 16, k1 = tid / 16;
            dft::lds_dft_step2<M2, INV>(lds + d2 * PAD, k1, v);
            uint32_t tb, cb;
            split(d2, tb, cb);
            const uint64_t i1 = s + a.E * (K0 + tb);
            const uint64_t c = C0 + cb;
#pragma unroll
            for (int k2 = 0; k2 < R; ++k2) {
                const uint64_t i = i1 + a.L * (k1 + (uint64_t)R * k2);
                const uint64_t idx = a.row_major ? (i * a.C + c) : (c * a.n + i);
                a.out[idx * a.out_el + a.out_comp] = v[k2];
            }
        }
        __syncthreads();  // the next tile rewrites the LDS buffer and twiddles
    }
}

// One-pass encode of a shape with m = M <= 64 coefficients per column and
// many output cosets (the last WHIR rounds: m = 4 .. 64, thousands of
// cosets): lane (s, c) computes the whole coset s of column c in registers,
//   Y[s + E k][c] = DFT_M( f_c[j] * w_n^{s j} )[k],   k < M,
// the workgroup's SPB = 256 / C cosets (at most 2048 / M) sharing one LDS
// table of their twist factors w_n^{s j}, built one entry per lane, so a
// table entry serves the C columns of its coset.  A coset's M output rows are
// written C contiguous values at a time (row-major) or one column run
// (column-major).
template <int M>
__global__ __launch_bounds__(256) void lde_small_direct(LdeArgs a, uint32_t spb) {
    __shared__ uint64_t tw[2048];
    const uint32_t C = (uint32_t)a.C;
    const uint64_t s0 = (uint64_t)blockIdx.x * spb;
    for (uint32_t q = threadIdx.x; q < spb * M; q += blockDim.x) {
        const uint64_t s = s0 + q / M, j = q % M;
        tw[q] = s < a.E ? twid(a, (s * j) & (a.n - 1)) : 1;
    }
    __syncthreads();
    const uint32_t c = threadIdx.x % C, sl = threadIdx.x / C;
    const uint64_t s = s0 + sl;
    if (sl >= spb || s >= a.E) return;
    uint64_t x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = a.in[in_index(a, c, j)];
#pragma unroll
    for (int j = 1; j < M; ++j) x[j] = gl::mul(x[j], tw[sl * M + j]);
    dft::dft_reg<M, false>(x);
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const uint64_t i = s + a.E * k;
        const uint64_t idx = a.row_major ? (i * C + c) : (c * a.n + i);
        a.out[idx * a.out_el + a.out_comp] = x[k];
    }
}

// m == 1: every codeword entry of column c is its single coefficient.
__global__ __launch_bounds__(256) void lde_broadcast_kernel(LdeArgs a) {
    const uint64_t total = a.n * a.C;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t i, c;
        if (a.row_major) {
            i = q / a.C;
            c = q - i * a.C;
        } else {
            c = q / a.n;
            i = q - c * a.n;
        }
        a.out[q * a.out_el + a.out_comp] = a.in[in_index(a, c, 0)];
    }
}

// workgroups of `fn` resident per CU (registers and LDS), cached per kernel and
// LDS size; at least 1
int resident_blocks(const void* fn, int threads, size_t lds) {
    static std::mutex mu;
    static std::map<std::pair<const void*, size_t>, int> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({fn, lds});
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, lds) != hipSuccess || n < 1) n = 1;
    cache[{fn, lds}] = n;
    return n;
}

// compute units of the current device (cached per device)
int cu_count() {
    static int cache[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cache[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

// NTTs longer than 2^20 (n = m * E, m = 2^20, 2 <= E <= 64), step 1: for
// column c and j0 < m, the E-point DFT over f[j0 + m j1] (j1 < E) in
// registers, output s twisted by w_n^(s j0), into g[c][s][j0].  Step 2 is the
// m-point NTT of every g[c][s] written to rows i = s + E k (row-major over E
// columns):  Y[s + E k] = sum_j0 w_m^(k j0) w_n^(s j0) sum_j1 f[j0 + m j1] w_E^(s j1).
template <int E, bool INV>
__global__ __launch_bounds__(256) void ntt_long_step1(const uint64_t* __restrict__ in, uint64_t* __restrict__ g,
                                                      uint32_t lg_n, uint32_t lg_m, uint64_t C,
                                                      const uint64_t* __restrict__ tw_lo,
                                                      const uint64_t* __restrict__ tw_hi) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (C << lg_m)) return;
    const uint64_t m = 1ull << lg_m, c = q >> lg_m, j0 = q & (m - 1);
    const uint64_t* f = in + (c << lg_n) + j0;
    uint64_t v[E];
#pragma unroll
    for (int j1 = 0; j1 < E; ++j1) v[j1] = f[m * j1];
    dft::dft_reg<E, INV>(v);
    uint64_t w = tw_lo[j0 & 2047];  // w_n^j0 from the two-level table
    if (j0 >> 11) w = gl::mul(w, tw_hi[j0 >> 11]);
    uint64_t* o = g + (c << lg_n) + j0;
    o[0] = v[0];
    uint64_t t = w;
#pragma unroll
    for (int s = 1; s < E; ++s) {
        o[m * s] = gl::mul(v[s], t);
        t = gl::mul(t, w);
    }
}

int ilog2u(uint64_t x) {
    int k = 0;
    while ((1ull << k) < x) ++k;
    return k;
}
bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
uint32_t split_r2(uint64_t M) {
    int lg = ilog2u(M);
    return (uint32_t)(M >> (lg / 2));
}
uint64_t padded(uint64_t M) { return M + M / split_r2(M) + 1; }

// LDS of lde_pass_b_rows2: the exchange / stage buffer and one twiddle table
constexpr size_t ROWS2_LDS = (size_t)(XCH_WORDS + 1024) * 8;

// The fused-rows pass B form: 2 = lde_pass_b_rows2 (two workgroups per CU,
// the default), 1 = lde_pass_b_rows (one workgroup per CU with the next
// tile's loads in registers).  WHIR_LDE_ROWS overrides it for A/B timing.
int rows_variant() {
    static const int v = [] {
        const char* e = std::getenv("WHIR_LDE_ROWS");
        return (e && e[0] == '1') ? 1 : 2;
    }();
    return v;
}

using KernelFn = void (*)(LdeArgs, uint64_t);
using RowsFn = void (*)(LdeArgs, uint64_t, uint32_t);
using ColsFn = void (*)(LdeArgs, uint64_t, uint32_t, uint32_t);

template <bool INV>
KernelFn pick_a(uint64_t M) {
    switch (M) {
        case 1: return lde_pass_a<1, INV>;
        case 2: return lde_pass_a<2, INV>;
        case 4: return lde_pass_a<4, INV>;
        case 8: return lde_pass_a<8, INV>;
        case 16: return lde_pass_a<16, INV>;
        case 32: return lde_pass_a<32, INV>;
        case 64: return lde_pass_a<64, INV>;
        case 128: return lde_pass_a<128, INV>;
        case 256: return lde_pass_a<256, INV>;
        case 512: return lde_pass_a<512, INV>;
        case 1024: return lde_pass_a<1024, INV>;
    }
    return nullptr;
}

template <bool INV>
KernelFn pick_b(uint64_t M) {
    switch (M) {
        case 1: return lde_pass_b<1, INV>;
        case 2: return lde_pass_b<2, INV>;
        case 4: return lde_pass_b<4, INV>;
        case 8: return lde_pass_b<8, INV>;
        case 16: return lde_pass_b<16, INV>;
        case 32: return lde_pass_b<32, INV>;
        case 64: return lde_pass_b<64, INV>;
        case 128: return lde_pass_b<128, INV>;
        case 256: return lde_pass_b<256, INV>;
        case 512: return lde_pass_b<512, INV>;
        case 1024: return lde_pass_b<1024, INV>;
    }
    return nullptr;
}

struct Plan {
    uint64_t M1, M2, E, L, G;  // G cosets per launch
    uint32_t TA, TB, CB;
    uint32_t threads_b;  // pass-B workgroup size
    bool direct;
};

// Largest power of two dividing x (capped).
uint64_t pow2_div(uint64_t x, uint64_t cap) {
    uint64_t d = 1;
    while ((x % (d * 2)) == 0 && d * 2 <= cap) d *= 2;
    return d;
}

#ifndef LDE_SCRATCH_LOG
#define LDE_SCRATCH_LOG 26
#endif
#ifndef LDE_BLOCKS_TARGET
#define LDE_BLOCKS_TARGET 1024
#endif
constexpr uint64_t SCRATCH_TARGET = 1ull << LDE_SCRATCH_LOG;  // u64 per launch group for small problems

bool make_plan(const LdeShape& s, Plan& p) {
    if (s.C == 0 || !is_pow2(s.m) || !is_pow2(s.n) || s.n < s.m || s.n > (1ull << 32)) return false;
    p.M2 = std::min<uint64_t>(s.m, 1024);
    if (s.lg_m2) {
        uint64_t m2 = 1ull << s.lg_m2;
        if (m2 > s.m || m2 > 1024) return false;
        p.M2 = m2;
    }
    p.M1 = s.m / p.M2;
    if (p.M1 > 1024) return false;
    if (s.m * s.C * (s.split3 ? 1 : s.in_el) >= (1ull << 31)) return false;  // per-launch offsets are 32-bit
    if (s.split3 && (s.C % 3 || s.in_el != 3 || s.out_el != 1)) return false;
    p.E = s.n / s.m;
    p.L = p.E * p.M1;
    p.TA = (uint32_t)std::min<uint64_t>(p.M2, std::max<uint64_t>(1, THREADS_A / split_r2(p.M1)));
    // 16-column base codewords keep 512 threads (the fused row leaves need all
    // 16 columns of a row in one workgroup); other shapes may use fewer
    const bool fusable = s.C == FUSE_COLS && s.row_major && s.out_el == 1 && !s.split3;
    p.threads_b = fusable ? THREADS : LDE_THREADS_B_SMALL;
    uint64_t DB = std::min<uint64_t>(std::max<uint64_t>(1, p.threads_b / split_r2(p.M2)), s.C * p.M1);
    if (s.row_major) {
        p.CB = (uint32_t)pow2_div(s.C, DB);
        p.TB = (uint32_t)std::min<uint64_t>(DB / p.CB, p.M1);
    } else {
        p.TB = (uint32_t)std::min<uint64_t>(DB, p.M1);
        p.CB = (uint32_t)pow2_div(s.C, DB / p.TB);
    }
    uint64_t per = s.m * s.C;
    p.direct = p.M1 == 1;
    p.G = p.direct ? std::min<uint64_t>(p.E, 32768)
                   : std::max<uint64_t>(1, std::min<uint64_t>({p.E, SCRATCH_TARGET / per, 32768}));
    return true;
}

// ------------------------------------------------- batches (LdeMultiSeg)
// Level-synchronous encodes of many 16-column polynomials: every first-pass
// tile of every polynomial in one launch, then every second-pass tile in one
// launch.  The same transform as lde_pass_a_regs + lde_pass_b_rows2 (M2 =
// 1024, inter-pass twiddle in the second pass, all E cosets of a polynomial
// in one scratch region), with the tile's polynomial looked up in a segment
// table: a workgroup walks global tiles t, t + gridDim.x, ... and finds the
// segment of t by binary search over the segments' first tiles (uniform, so
// scalar loads).  Both kernels run two 512-lane workgroups per CU (at most
// 128 VGPRs) and keep no register prefetch: the other workgroup computes
// while one waits for its loads.

__device__ __forceinline__ uint64_t twid_seg(const LdeMultiSeg& g, uint64_t e) {
    uint64_t x = ((__attribute__((address_space(1))) const uint64_t*)g.tw_lo)[e & 2047];
    uint64_t h = e >> 11;
    return h ? gl::mul(x, ((__attribute__((address_space(1))) const uint64_t*)g.tw_hi)[h]) : x;
}

// The segment table's pointers are loaded from memory, so the compiler
// cannot tell they address global memory and would use flat instructions
// (which also count against the LDS wait counter); these casts state it.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ gptr<T> global_ptr(T* p) {
    return (gptr<T>)p;
}

// the last segment whose first tile (of pass B, or of pass A) is <= t
template <bool PASS_B>
__device__ __forceinline__ uint32_t seg_of(const LdeMultiSeg* segs, uint32_t nseg, uint32_t t) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        const uint32_t first = PASS_B ? segs[mid].tile_b0 : segs[mid].tile_a0;
        if (first <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// first-pass tiles of one polynomial: columns c < 16 split into runs of TA
// consecutive j2 (TA * R2 = 512 lanes)
template <int M1>
constexpr uint32_t multi_a_tiles() {
    return 16 * (1024 / (THREADS / dft::Split<M1>::R2));
}
template <int M1>
constexpr size_t multi_a_lds_words() {
    return (size_t)(THREADS / dft::Split<M1>::R2) * (dft::Split<M1>::PADDED + 1) + M1;
}
constexpr size_t max_sz(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t MULTI_A_LDS =
    8 * max_sz(max_sz(max_sz(multi_a_lds_words<2>(), multi_a_lds_words<4>()),
                      max_sz(multi_a_lds_words<8>(), multi_a_lds_words<16>())),
               max_sz(max_sz(multi_a_lds_words<32>(), multi_a_lds_words<64>()),
                      max_sz(multi_a_lds_words<128>(), multi_a_lds_words<256>())));

// One first-pass tile (column c, j2 in [J, J + TA)) for all E cosets: lane
// (t, n2) holds x[t][R2 n1 + n2] in registers across the cosets; per coset:
// twist w_L^{s j1} (LDS table), DFT_R1 in registers, t1024 twiddles, DFT_R2
// from LDS, Z_s[k][c][j2] written (16 * 1024 values per k).
template <int M1>
__device__ __forceinline__ void multi_a_tile(const LdeMultiSeg& g, uint32_t lt, const uint64_t* __restrict__ t1024,
                                             uint64_t* lds) {
    using S = dft::Split<M1>;
    constexpr int R1 = S::R1, R2 = S::R2, PAD = S::PADDED + 1, TA = THREADS / R2;
    constexpr uint32_t M2 = 1024, NJ = M2 / TA, ZROW = 16 * M2;
    static_assert(R1 <= R2 && TA * R2 == THREADS && TA <= (int)M2, "first-pass tile shape");
    uint64_t* twist = lds + TA * PAD;
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // lane indices re-derived per tile (see lde_pass_b_rows)
    const uint32_t c = lt / NJ, J = (lt % NJ) * TA;
    const uint32_t t = tid % TA, n2 = tid / TA;
    const uint64_t m = 1ull << g.lg_m, nmask = (1ull << g.lg_n) - 1;
    const uint32_t E = 1u << (g.lg_n - g.lg_m);
    const gptr<const uint64_t> col = global_ptr(g.in + c * m + J);  // uniform base, 32-bit lane offsets
    const uint32_t off = t + M2 * n2;
    // the tile's inputs stay in registers across the cosets, except at the
    // largest size, where holding them would spill under the 128-VGPR budget
    // (its later cosets re-read the tile, from the cache)
    constexpr bool KEEP = M1 < 256;
    uint64_t pre[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) pre[n1] = col[off + M2 * R2 * n1];
    for (uint32_t s = 0; s < E; ++s) {
        if (!KEEP && s != 0) {
            uint32_t tid2 = threadIdx.x;
            asm volatile("" : "+v"(tid2));  // no addresses kept from the first load
            const uint32_t off2 = tid2 % TA + M2 * (tid2 / TA);
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) pre[n1] = col[off2 + M2 * R2 * n1];
        }
        if (s != 0) {
            for (uint32_t j1 = tid; j1 < (uint32_t)M1; j1 += THREADS)
                twist[j1] = twid_seg(g, ((uint64_t)s * j1 * M2) & nmask);
            __syncthreads();
        }
        uint64_t v[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = s != 0 ? gl::mul(pre[n1], twist[R2 * n1 + n2]) : pre[n1];
        dft::dft_reg<R1, false>(v);
        if constexpr (R2 > 1) {
#pragma unroll
            for (int k1 = 1; k1 < R1; ++k1) v[k1] = gl::mul(v[k1], t1024[((n2 * k1) * (1024 / M1)) & 1023]);
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) lds[t * PAD + dft::pidx<M1>(R2 * k1 + n2)] = v[k1];
        __syncthreads();
        const gptr<uint64_t> z = global_ptr(g.z + (uint64_t)s * (16 * m) + (uint64_t)c * M2 + J);  // uniform
        for (uint32_t w = tid; w < (uint32_t)(TA * R1); w += THREADS) {
            const uint32_t t2 = w % TA, k1 = w / TA;
            uint64_t u[R2];
            dft::lds_dft_step2<M1, false>(lds + t2 * PAD, k1, u);
            const uint32_t zo = k1 * ZROW + t2;  // < 2^28: 32-bit lane offsets
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) z[zo + (uint32_t)(R1 * k2) * ZROW] = u[k2];
        }
        __syncthreads();  // the next coset (or tile) rewrites the LDS tile and the twist table
    }
}

__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void lde_multi_a(
    const LdeMultiSeg* __restrict__ segs, uint32_t nseg, uint32_t ntiles, const uint64_t* __restrict__ t1024) {
    extern __shared__ uint64_t lds[];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const LdeMultiSeg g = segs[seg_of<false>(segs, nseg, t)];
        const uint32_t lt = t - g.tile_a0;
        switch (g.lg_m) {
            case 11: multi_a_tile<2>(g, lt, t1024, lds); break;
            case 12: multi_a_tile<4>(g, lt, t1024, lds); break;
            case 13: multi_a_tile<8>(g, lt, t1024, lds); break;
            case 14: multi_a_tile<16>(g, lt, t1024, lds); break;
            case 15: multi_a_tile<32>(g, lt, t1024, lds); break;
            case 16: multi_a_tile<64>(g, lt, t1024, lds); break;
            case 17: multi_a_tile<128>(g, lt, t1024, lds); break;
            case 18: multi_a_tile<256>(g, lt, t1024, lds); break;
            default: break;  // lde_multi_prepare admits no other size
        }
    }
}

// Second pass of every polynomial: tile = (coset s, first-pass output index
// k), 16 columns x 1024 rows i = i1 + L r (i1 = s + E k, L = E M1), as
// lde_pass_b_rows2 (two-round transposition, rows staged in two halves for
// the BLAKE3 leaves and the 16-byte stores).  A polynomial with m = 1024 has
// no first pass: its tiles read the coefficients themselves (k = 0, L = E,
// the inter-pass twiddle w_n^{s j2} being the coset twist).
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void lde_multi_b(
    const LdeMultiSeg* __restrict__ segs, uint32_t nseg, uint32_t ntiles, const uint64_t* __restrict__ t1024) {
    constexpr int M2 = 1024, R = 32;
    extern __shared__ uint64_t lds[];
    uint64_t* buf = lds;              // exchange, then row stage
    uint64_t* twt = lds + XCH_WORDS;  // inter-pass twiddles of the tile
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));  // lane indices re-derived per tile (see lde_pass_b_rows)
        const LdeMultiSeg g = segs[seg_of<true>(segs, nseg, tile)];
        const uint32_t lt = tile - g.tile_b0;
        const uint32_t lgM1 = g.lg_m - 10, lgE = g.lg_n - g.lg_m, lgL = lgE + lgM1;
        const uint32_t s = lt >> lgM1, k = lt & ((1u << lgM1) - 1);
        const uint64_t i1 = s + ((uint64_t)k << lgE);
        const uint64_t m = 1ull << g.lg_m, nmask = (1ull << g.lg_n) - 1;
        const uint32_t c = (tid >> 4) & 15, hw = tid >> 8, n2 = (tid & 15) | (hw << 4);
        const gptr<const uint64_t> src =
            global_ptr(lgM1 ? g.z + (uint64_t)s * (16 * m) + (uint64_t)k * (16 * M2) + (uint64_t)c * M2
                            : g.in + (uint64_t)c * m);
        uint64_t v[R];
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = src[n2 + R * n1];
        // the previous tile's last reads of twt and buf ended at its final barrier
        for (uint32_t j2 = tid; j2 < (uint32_t)M2; j2 += THREADS) twt[j2] = twid_seg(g, (i1 * j2) & nmask);
        __syncthreads();
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) v[n1] = gl::mul(v[n1], twt[n2 + R * n1]);
        dft::dft_reg<R, false>(v);
#pragma unroll
        for (int k1 = 1; k1 < R; ++k1) v[k1] = gl::mul(v[k1], t1024[(n2 * k1) & 1023]);
        const uint32_t cb = tid % 16, k1r = tid / 16;
        uint64_t u[R];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (hw == (uint32_t)h) {  // uniform per wave
#pragma unroll
                for (int k1 = 0; k1 < R; ++k1) buf[xch_idx(c, k1, n2 & 15)] = v[k1];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) u[16 * h + q] = buf[xch_idx(cb, k1r, q)];
            __syncthreads();
        }
        dft::dft_reg<R, false>(u);  // u[k2] = row k1r + 32 k2 of column cb
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) buf[stage_idx(k1r + R * kk, cb)] = u[16 * h + kk];
            __syncthreads();
            {
                const uint32_t r = tid;  // local row
                uint32_t cv[8];
                row_leaf<HASH_BLAKE3>([&](int cc) { return buf[stage_idx(r, cc)]; }, cv);
                const uint64_t i = i1 + ((uint64_t)(512 * h + r) << lgL);
                const gptr<u32x4> dst = global_ptr(reinterpret_cast<u32x4*>(g.leaves + 32 * i));
                dst[0] = u32x4{cv[0], cv[1], cv[2], cv[3]};
                dst[1] = u32x4{cv[4], cv[5], cv[6], cv[7]};
            }
#pragma unroll
            for (int q2 = 0; q2 < 512 * 16 / (2 * THREADS); ++q2) {
                const uint32_t q = tid + q2 * THREADS;
                const uint32_t r = q / 8, cp = 2 * (q % 8);
                const uint64_t x0 = buf[stage_idx(r, cp)], x1 = buf[stage_idx(r, cp + 1)];
                const uint64_t i = i1 + ((uint64_t)(512 * h + r) << lgL);
                *global_ptr(reinterpret_cast<u64x2*>(g.out + i * 16 + cp)) = u64x2{x0, x1};
            }
            __syncthreads();  // the next half (or tile) rewrites the stage
        }
    }
}

}  // namespace

int lde_splits(const LdeShape& s, uint32_t* lg_m2, int cap) {
    int k = 0;
    LdeShape t = s;
    t.lg_m2 = 0;
    Plan p;
    if (!make_plan(t, p)) return 0;
    const uint32_t def = (uint32_t)ilog2u(p.M2);
    if (k < cap) lg_m2[k++] = def;
    if (p.direct) return k;  // one pass, no scratch: the only split offered
    for (uint32_t lg = def; lg-- > 0 && k < cap;) {
        t.lg_m2 = lg;
        if (lg == 0 || !make_plan(t, p)) break;
        lg_m2[k++] = lg;
    }
    return k;
}

size_t lde_scratch_elems_any_split(const LdeShape& s) {
    uint32_t cand[11];
    int nc = lde_splits(s, cand, 11);
    size_t need = 0;
    LdeShape t = s;
    for (int c = 0; c < nc; ++c) {
        t.lg_m2 = cand[c];
        need = std::max(need, lde_scratch_elems(t));
    }
    return need;
}

size_t lde_scratch_elems(const LdeShape& s) {
    Plan p;
    if (!make_plan(s, p)) return 0;
    return p.direct ? 0 : (size_t)(p.G * s.m * s.C);
}

int lde_launch(const LdeShape& s, const uint64_t* in, uint64_t* out, uint64_t* scratch,
               const Twiddles& tw, hipStream_t st, uint8_t* leaves, bool* fused) {
    Plan p;
    if (fused) *fused = false;
    if (!make_plan(s, p)) return 1;
    if (tw.log_n != (uint32_t)ilog2u(s.n)) return 2;
    LdeArgs a;
    a.in = in;
    a.out = out;
    a.z = scratch;
    a.tw_in_b = p.TB == 1 ? 1 : 0;
    bool fuse = leaves && s.row_major && s.C == FUSE_COLS && p.TB == 1 && p.CB == FUSE_COLS && s.out_el == 1;
    // SHA-256 row leaves are hashed inside the persistent forward rows forms
    // only; for other shapes the caller runs the separate leaf kernel
    const bool sha2 = s.leaf_hash == HASH_SHA2;
    if (sha2 && !(LDE_PB_PERSIST && p.M2 == 1024 && !p.direct && !s.inverse)) fuse = false;
    // Ext3 rows of 16 Ext3 columns (48 base columns): the leaves (either
    // engine) fused into the three-sub-tile persistent pass B (forward,
    // M2 = 1024, with a first pass)
    const bool fuse3 = leaves && s.split3 && s.row_major && s.C == 3 * FUSE_COLS && p.TB == 1 &&
                       p.CB == FUSE_COLS && s.out_el == 1 && LDE_PB_PERSIST && p.M2 == 1024 && !p.direct &&
                       !s.inverse;
    a.leaves = (fuse || fuse3) ? leaves : nullptr;
    if (fused) *fused = fuse || fuse3;
    a.tw_lo = tw.lo;
    a.tw_hi = tw.hi;
    a.t1024 = tw.t1024;
    a.C = s.C;
    a.m = s.m;
    a.n = s.n;
    a.M1 = p.M1;
    a.M2 = p.M2;
    a.E = p.E;
    a.L = p.L;
    a.zstride = s.m * s.C;
    a.in_el = s.in_el;
    a.in_comp = s.in_comp;
    a.out_el = s.out_el;
    a.out_comp = s.out_comp;
    a.TA = p.TA;
    a.TB = p.TB;
    a.CB = p.CB;
    a.lgTA = (uint32_t)ilog2u(p.TA);
    a.lgTB = (uint32_t)ilog2u(p.TB);
    a.lgCB = (uint32_t)ilog2u(p.CB);
    a.lgNJ = (uint32_t)ilog2u(p.M2 / p.TA);
    a.lgNK = (uint32_t)ilog2u(p.M1 / p.TB);
    a.row_major = s.row_major ? 1 : 0;
    a.split3 = s.split3 ? 1 : 0;
    a.direct = p.direct ? 1 : 0;

    KernelFn ka = s.inverse ? pick_a<true>(p.M1) : pick_a<false>(p.M1);
    KernelFn kb = s.inverse ? pick_b<true>(p.M2) : pick_b<false>(p.M2);
    size_t lds_a = ((size_t)p.TA * padded(p.M1) + p.M1) * 8;
    size_t words_b = fuse ? std::max<size_t>(16 * padded(p.M2), FUSE_COLS * p.M2) : (size_t)p.TB * p.CB * padded(p.M2);
    size_t lds_b = (words_b + (a.tw_in_b ? p.M2 : 0)) * 8;
    static std::mutex attr_mu;
    static bool attr_done[2][2][11] = {};
    int ia = ilog2u(p.M1), ib = ilog2u(p.M2), inv = s.inverse ? 1 : 0;
    {
        std::lock_guard<std::mutex> g(attr_mu);
        if (!attr_done[inv][0][ia]) {
            (void)hipFuncSetAttribute((const void*)ka, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            attr_done[inv][0][ia] = true;
        }
        if (!attr_done[inv][1][ib]) {
            (void)hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            attr_done[inv][1][ib] = true;
        }
    }
    // fused 16-column rows with M2 = 1024: the persistent prefetching pass B
    const bool rows = LDE_PB_PERSIST && fuse && p.M2 == 1024 && !p.direct;
    RowsFn kr = s.inverse ? lde_pass_b_rows<true>
                          : (sha2 ? lde_pass_b_rows<false, HASH_SHA2> : lde_pass_b_rows<false>);
    // (SHA-256 leaves take the one-workgroup form: under the two-workgroup
    // form's 128-VGPR budget their 64-round schedule spills)
    RowsFn kr2 = s.inverse ? lde_pass_b_rows2<true> : lde_pass_b_rows2<false>;
    if (rows) {
        std::lock_guard<std::mutex> g(attr_mu);
        static bool rows_attr[3] = {};  // forward BLAKE3, inverse, forward SHA-256
        const int which = s.inverse ? 1 : (sha2 ? 2 : 0);
        if (!rows_attr[which]) {
            (void)hipFuncSetAttribute((const void*)kr, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            (void)hipFuncSetAttribute((const void*)kr2, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            rows_attr[which] = true;
        }
    }
    // other M2 = 1024 shapes with 16 Z columns per tile: the persistent
    // prefetching pass B without leaves
    RowsFn kr3 = sha2 ? lde_pass_b_rows_e3<false, HASH_SHA2> : lde_pass_b_rows_e3<false>;
    if (fuse3) {
        std::lock_guard<std::mutex> g(attr_mu);
        static bool e3_attr[2] = {};
        if (!e3_attr[sha2]) {
            (void)hipFuncSetAttribute((const void*)kr3, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            e3_attr[sha2] = true;
        }
    }
    const bool cols = LDE_PB_PERSIST && !fuse && !fuse3 && p.M2 == 1024 && !p.direct && p.TB * p.CB == 16 &&
                      p.threads_b == THREADS;
    ColsFn kc = s.inverse ? lde_pass_b_cols<true> : lde_pass_b_cols<false>;
    if (cols) {
        std::lock_guard<std::mutex> g(attr_mu);
        static bool cols_attr[2] = {};
        if (!cols_attr[inv]) {
            (void)hipFuncSetAttribute((const void*)kc, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            cols_attr[inv] = true;
        }
    }
    if (s.m == 1 && !fuse) {  // a broadcast: one memory-bound launch for all cosets
        const uint64_t total = s.n * s.C;
        const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(lde_broadcast_kernel, dim3(blocks), dim3(256), 0, st, a);
        return 0;
    }
    // forward one-pass shapes of at most 64 coefficients per column without
    // fused leaves: whole cosets in registers, one lane per (coset, column)
    if (p.direct && !fuse && !s.inverse && s.m >= 2 && s.m <= 32 && p.E >= 2 && s.C <= 256) {
        const uint32_t spb = (uint32_t)std::min<uint64_t>(256 / s.C, 2048 / s.m);
        const unsigned blocks = (unsigned)((p.E + spb - 1) / spb);
        switch (s.m) {
            case 2: hipLaunchKernelGGL(lde_small_direct<2>, dim3(blocks), dim3(256), 0, st, a, spb); break;
            case 4: hipLaunchKernelGGL(lde_small_direct<4>, dim3(blocks), dim3(256), 0, st, a, spb); break;
            case 8: hipLaunchKernelGGL(lde_small_direct<8>, dim3(blocks), dim3(256), 0, st, a, spb); break;
            case 16: hipLaunchKernelGGL(lde_small_direct<16>, dim3(blocks), dim3(256), 0, st, a, spb); break;
            case 32: hipLaunchKernelGGL(lde_small_direct<32>, dim3(blocks), dim3(256), 0, st, a, spb); break;
            default: hipLaunchKernelGGL(lde_small_direct<64>, dim3(blocks), dim3(256), 0, st, a, spb); break;
        }
        return 0;
    }
    // M1 >= 16 with a full tile: the persistent register-tile pass A
    RowsFn kar = nullptr;
    if (LDE_PA_PERSIST && !p.direct && p.TA * split_r2(p.M1) == THREADS_A) {
        switch (p.M1) {
            case 16: kar = s.inverse ? lde_pass_a_regs<16, true> : lde_pass_a_regs<16, false>; break;
            case 32: kar = s.inverse ? lde_pass_a_regs<32, true> : lde_pass_a_regs<32, false>; break;
            case 64: kar = s.inverse ? lde_pass_a_regs<64, true> : lde_pass_a_regs<64, false>; break;
            case 128: kar = s.inverse ? lde_pass_a_regs<128, true> : lde_pass_a_regs<128, false>; break;
            case 256: kar = s.inverse ? lde_pass_a_regs<256, true> : lde_pass_a_regs<256, false>; break;
            case 512: kar = s.inverse ? lde_pass_a_regs<512, true> : lde_pass_a_regs<512, false>; break;
            case 1024: kar = s.inverse ? lde_pass_a_regs<1024, true> : lde_pass_a_regs<1024, false>; break;
        }
    }
    if (kar) {
        std::lock_guard<std::mutex> g(attr_mu);
        static bool regs_attr[2][11] = {};
        if (!regs_attr[inv][ia]) {
            (void)hipFuncSetAttribute((const void*)kar, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            regs_attr[inv][ia] = true;
        }
    }
    dim3 ga((unsigned)(s.C * (p.M2 / p.TA)), 1, 1);
    if (kar && ga.x < (unsigned)cu_count()) kar = nullptr;  // too few tiles: split cosets over blocks
    dim3 gb((unsigned)((p.M1 / p.TB) * (s.C / p.CB)), 1, 1);
    // cosets per pass-A block: enough blocks to fill the chip, the rest reuse
    // the loaded tile
    const uint64_t bpc = ga.x;
    for (uint64_t s0 = 0; s0 < p.E; s0 += p.G) {
        uint64_t g = std::min<uint64_t>(p.G, p.E - s0);
        uint64_t cpb = std::max<uint64_t>(1, std::min<uint64_t>(g, g * bpc / LDE_BLOCKS_TARGET));
        a.cpb = (uint32_t)cpb;
        a.ncos = (uint32_t)g;
        ga.y = (unsigned)((g + cpb - 1) / cpb);
        gb.y = (unsigned)g;
        if (kar) {
            a.cpb = (uint32_t)g;
            const uint32_t per_cu = (uint32_t)resident_blocks((const void*)kar, THREADS_A, lds_a);
            const uint32_t grid = std::min<uint32_t>(ga.x, per_cu * (uint32_t)cu_count());
            hipLaunchKernelGGL(kar, dim3(grid), dim3(THREADS_A), lds_a, st, a, s0, ga.x);
        } else if (!p.direct) {
            hipLaunchKernelGGL(ka, ga, dim3(THREADS_A), lds_a, st, a, s0);
        }
        if (rows && !sha2 && (s.rows_form ? s.rows_form : rows_variant()) == 2) {
            // two resident workgroups per CU, half a tile of LDS each
            const uint32_t ntiles = (uint32_t)(g * p.M1);
            const uint32_t per_cu = (uint32_t)std::min(2, resident_blocks((const void*)kr2, THREADS, ROWS2_LDS));
            const uint32_t grid = std::min<uint32_t>(ntiles, per_cu * (uint32_t)cu_count());
            hipLaunchKernelGGL(kr2, dim3(grid), dim3(THREADS), ROWS2_LDS, st, a, s0, ntiles);
        } else if (rows) {
            // one resident workgroup per CU (the LDS use allows no more)
            const uint32_t ntiles = (uint32_t)(g * p.M1);
            const uint32_t grid = std::min<uint32_t>(ntiles, (uint32_t)cu_count());
            const size_t lds_rows = (16 * padded(1024) + 2 * 1024) * 8;  // DFT buffer + two twiddle tables
            hipLaunchKernelGGL(kr, dim3(grid), dim3(THREADS), lds_rows, st, a, s0, ntiles);
        } else if (fuse3) {
            // one resident workgroup per CU (the LDS use allows no more)
            const uint32_t ntiles = (uint32_t)(g * p.M1);
            const uint32_t grid = std::min<uint32_t>(ntiles, (uint32_t)cu_count());
            const size_t lds_rows = (16 * padded(1024) + 2 * 1024) * 8;  // DFT buffer + two twiddle tables
            hipLaunchKernelGGL(kr3, dim3(grid), dim3(THREADS), lds_rows, st, a, s0, ntiles);
        } else if (cols) {
            const uint32_t ntiles = (uint32_t)(g * gb.x);
            const uint32_t grid = std::min<uint32_t>(ntiles, (uint32_t)cu_count());
            hipLaunchKernelGGL(kc, dim3(grid), dim3(THREADS), lds_b, st, a, s0, (uint32_t)gb.x, ntiles);
        } else {
            hipLaunchKernelGGL(kb, gb, dim3(p.threads_b), lds_b, st, a, s0);
        }
    }
    return 0;
}

size_t lde_multi_scratch_elems(uint64_t m, uint64_t n) { return m > 1024 ? (size_t)(16 * n) : 0; }

namespace {
// kernel attributes and occupancy, queried once on the host (before any
// graph capture: lde_multi_prepare runs when a batch is planned)
struct MultiGrid {
    uint32_t per_cu_a = 1, per_cu_b = 1;
};
const MultiGrid& multi_grid() {
    static MultiGrid g;
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void*)lde_multi_a, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void*)lde_multi_b, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        g.per_cu_a = (uint32_t)std::min(2, resident_blocks((const void*)lde_multi_a, THREADS, MULTI_A_LDS));
        g.per_cu_b = (uint32_t)std::min(2, resident_blocks((const void*)lde_multi_b, THREADS, ROWS2_LDS));
    });
    return g;
}
}  // namespace

int lde_multi_prepare(std::vector<LdeMultiSeg>& segs, LdeMultiPlan& plan) {
    (void)multi_grid();
    uint64_t ta = 0, tb = 0;
    for (size_t i = 0; i < segs.size(); ++i) {
        LdeMultiSeg& g = segs[i];
        if (g.lg_m < 10 || g.lg_m > 18 || g.lg_n <= g.lg_m || g.lg_n > 26) return 1;
        if (i && g.lg_m > segs[i - 1].lg_m) return 1;  // decreasing m
        if (!g.in || !g.out || !g.leaves || !g.tw_lo || !g.tw_hi || (g.lg_m > 10 && !g.z)) return 1;
        g.tile_a0 = (uint32_t)ta;
        g.tile_b0 = (uint32_t)tb;
        static constexpr uint32_t per_a[9] = {0,
                                              multi_a_tiles<2>(),
                                              multi_a_tiles<4>(),
                                              multi_a_tiles<8>(),
                                              multi_a_tiles<16>(),
                                              multi_a_tiles<32>(),
                                              multi_a_tiles<64>(),
                                              multi_a_tiles<128>(),
                                              multi_a_tiles<256>()};
        ta += per_a[g.lg_m - 10];
        tb += 1ull << (g.lg_n - 10);  // E cosets x M1 first-pass outputs = n / 1024
    }
    if (ta >= (1ull << 31) || tb >= (1ull << 31)) return 1;
    plan.tiles_a = (uint32_t)ta;
    plan.tiles_b = (uint32_t)tb;
    return 0;
}

void lde_multi_launch(const LdeMultiSeg* d_segs, uint32_t nseg, const LdeMultiPlan& plan, const uint64_t* t1024,
                      hipStream_t st) {
    const MultiGrid& mg = multi_grid();
    if (!nseg) return;
    if (plan.tiles_a) {
        const uint32_t grid = std::min<uint32_t>(plan.tiles_a, mg.per_cu_a * (uint32_t)cu_count());
        hipLaunchKernelGGL(lde_multi_a, dim3(grid), dim3(THREADS), MULTI_A_LDS, st, d_segs, nseg, plan.tiles_a, t1024);
    }
    if (plan.tiles_b) {
        const uint32_t grid = std::min<uint32_t>(plan.tiles_b, mg.per_cu_b * (uint32_t)cu_count());
        hipLaunchKernelGGL(lde_multi_b, dim3(grid), dim3(THREADS), ROWS2_LDS, st, d_segs, nseg, plan.tiles_b, t1024);
    }
}

namespace {
// dst[c * rows + r] = src[r * cols + c] for a rows x cols u64 matrix: 32 x 32
// tiles through LDS (one padding word per tile row keeps the column reads
// conflict-free), 8-byte reads along source rows and writes along destination
// rows.  Workgroup b walks the tiles b, b + gridDim.x, ...; ragged edges are
// bounds-checked.
constexpr int TT = 32;
__global__ __launch_bounds__(256) void transpose_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                        uint64_t rows, uint64_t cols, uint64_t tiles_c,
                                                        uint64_t ntiles) {
    __shared__ uint64_t tile[TT][TT + 1];
    const uint32_t tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8 threads
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t r0 = (t / tiles_c) * TT, c0 = (t % tiles_c) * TT;
        for (uint32_t k = ty; k < TT; k += 8) {
            const uint64_t r = r0 + k, c = c0 + tx;
            if (r < rows && c < cols) tile[k][tx] = src[r * cols + c];
        }
        __syncthreads();
        for (uint32_t k = ty; k < TT; k += 8) {
            const uint64_t c = c0 + k, r = r0 + tx;  // destination row c, column r
            if (r < rows && c < cols) dst[c * rows + r] = tile[tx][k];
        }
        __syncthreads();  // the next tile rewrites the LDS tile
    }
}
}  // namespace

void transpose_launch(const uint64_t* src, uint64_t* dst, uint64_t rows, uint64_t cols, hipStream_t st) {
    if (!rows || !cols) return;
    const uint64_t tiles_c = (cols + TT - 1) / TT, ntiles = ((rows + TT - 1) / TT) * tiles_c;
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, 16384);
    hipLaunchKernelGGL(transpose_kernel, dim3(grid), dim3(256), 0, st, src, dst, rows, cols, tiles_c, ntiles);
}

int ntt_long_step1_launch(const uint64_t* in, uint64_t* g, uint64_t n, uint64_t m, uint64_t C, const Twiddles& tw,
                          bool inverse, hipStream_t st) {
    using F = void (*)(const uint64_t*, uint64_t*, uint32_t, uint32_t, uint64_t, const uint64_t*, const uint64_t*);
    F f = nullptr;
    switch (n / m) {
        case 2: f = inverse ? ntt_long_step1<2, true> : ntt_long_step1<2, false>; break;
        case 4: f = inverse ? ntt_long_step1<4, true> : ntt_long_step1<4, false>; break;
        case 8: f = inverse ? ntt_long_step1<8, true> : ntt_long_step1<8, false>; break;
        case 16: f = inverse ? ntt_long_step1<16, true> : ntt_long_step1<16, false>; break;
        case 32: f = inverse ? ntt_long_step1<32, true> : ntt_long_step1<32, false>; break;
        case 64: f = inverse ? ntt_long_step1<64, true> : ntt_long_step1<64, false>; break;
    }
    if (!f || !is_pow2(m) || n % m || tw.log_n != (uint32_t)ilog2u(n)) return 1;
    const uint64_t work = C * m;
    hipLaunchKernelGGL(f, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, in, g, (uint32_t)ilog2u(n),
                       (uint32_t)ilog2u(m), C, tw.lo, tw.hi);
    return 0;
}

}  // namespace whir_gpu
