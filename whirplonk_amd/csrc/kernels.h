// kernels.h — host-side launchers of the HIP kernels (internal to the library).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace whir_gpu {

// hash engines of the leaves, trees and proof-of-work searches
enum : uint32_t { HASH_BLAKE3 = 1, HASH_SHA2 = 2 };

// Twiddle tables of one transform size n = 2^log_n, forward or inverse:
// lo[x] = w^x (x < 2048), hi[y] = w^(2048 y) (y < max(1, n/2048)) so that
// w^e = lo[e & 2047] * hi[e >> 11]; t1024[t] = w_1024^t for the sub-DFTs.
struct Twiddles {
    uint32_t log_n = 0;
    uint64_t* lo = nullptr;
    uint64_t* hi = nullptr;
    uint64_t* t1024 = nullptr;
};

// Geometry of one interleaved LDE / batched NTT:
//   columns c < C, each with m input values at in[(c*m + j)*in_el + in_comp]
//   (Ext3 AoS input: in_el = 3, in_comp = component), output n = E*m values
//   per column, written row-major (out[(i*C + c)*out_el + out_comp]) or
//   column-major (out[((c*n) + i)*out_el + out_comp]).
struct LdeShape {
    uint64_t C = 0, m = 0, n = 0;
    uint32_t in_el = 1, in_comp = 0, out_el = 1, out_comp = 0;
    bool row_major = true;
    bool inverse = false;
    // Ext3 AoS input as 3 base columns per Ext3 column: column c reads
    // component c % 3 of Ext3 column c / 3 (C counts base columns, in_el = 3,
    // out_el = 1), so a row-major output row is the AoS Ext3 row.
    bool split3 = false;
    // log2 of the second-pass DFT size M2 (m = M1 * M2); 0 = the default
    // split M2 = min(m, 1024).  Any valid split gives the same output.
    uint32_t lg_m2 = 0;
    // form of the fused-leaves pass B of 16-column rows (M2 = 1024): 1 = one
    // workgroup per CU with the next tile's loads in registers, 2 = two
    // workgroups per CU with half-tile LDS rounds, 0 = the library default.
    // Both give the same output.
    uint32_t rows_form = 0;
    // engine of the row leaves hashed inside the last pass (lde_launch's
    // `leaves`): BLAKE3 fuses into every 16-column row form, SHA-256 into the
    // persistent forward rows forms (M2 = 1024 with a first pass)
    uint32_t leaf_hash = HASH_BLAKE3;
};

// Step 1 of an NTT longer than 2^20 (n = m * E, 2 <= E <= 64): see lde.hip.
// g receives C * n values; tw is the size-n table of the direction.  Returns
// nonzero for an unsupported shape.
int ntt_long_step1_launch(const uint64_t* in, uint64_t* g, uint64_t n, uint64_t m, uint64_t C, const Twiddles& tw,
                          bool inverse, hipStream_t st);
// dst[c * rows + r] = src[r * cols + c] (u64; src and dst must not overlap)
void transpose_launch(const uint64_t* src, uint64_t* dst, uint64_t rows, uint64_t cols, hipStream_t st);
// Scratch needed by lde_launch (elements of u64); 0 when m <= 1024 (one
// pass reads the input directly).
size_t lde_scratch_elems(const LdeShape& s);
// The most scratch any valid split of the shape needs (for buffers planned
// before the split is known).
size_t lde_scratch_elems_any_split(const LdeShape& s);
// The valid second-pass sizes for a shape (log2), default first.
int lde_splits(const LdeShape& s, uint32_t* lg_m2, int cap);
// Returns 0, or >0 when the shape is not supported (not a power of two, m > 2^20).
// With leaves != null and a row-major 16-column codeword the BLAKE3 row
// leaves are hashed inside the last pass; *fused tells whether that happened.
int lde_launch(const LdeShape& s, const uint64_t* in, uint64_t* out, uint64_t* scratch,
               const Twiddles& tw, hipStream_t st, uint8_t* leaves = nullptr, bool* fused = nullptr);

// Encodes of many 16-column base polynomials (folding factor 4) with their
// BLAKE3 row leaves, level-synchronous: one first-pass launch over every
// polynomial with m >= 2048 coefficients per column, then one second-pass
// launch over all of them (M2 = 1024 rows per tile, m = M1 * 1024,
// 1024 <= m <= 2^18, n = E * m with 2 <= E, n <= 2^26).  A segment is one
// polynomial; lde_multi_prepare fills the tile ranges (segments must be in
// decreasing m) and returns the launch sizes.
struct LdeMultiSeg {
    const uint64_t* in;      // coefficients: column c < 16 at in + c * m
    uint64_t* out;           // codeword, row-major n x 16
    uint8_t* leaves;         // n x 32 bytes (the tree's leaf level)
    uint64_t* z;             // first-pass scratch, E * 16 * m u64 (unused when m == 1024)
    const uint64_t* tw_lo;   // the size-n twiddle tables (Twiddles::lo, ::hi)
    const uint64_t* tw_hi;
    uint32_t lg_m, lg_n;
    uint32_t tile_a0, tile_b0;  // the segment's first tile of each launch
};
struct LdeMultiPlan {
    uint32_t tiles_a = 0, tiles_b = 0;
};
int lde_multi_prepare(std::vector<LdeMultiSeg>& segs, LdeMultiPlan& plan);
// u64 of first-pass scratch a segment of this shape needs
size_t lde_multi_scratch_elems(uint64_t m, uint64_t n);
void lde_multi_launch(const LdeMultiSeg* d_segs, uint32_t nseg, const LdeMultiPlan& plan, const uint64_t* t1024,
                      hipStream_t st);

// leaves[i] = BLAKE3(row i), row = row_u64 little-endian u64 words.
void leaf_hash_launch(const uint64_t* matrix, uint64_t rows, uint32_t row_u64, uint8_t* leaves, hipStream_t st);
// out[i] = BLAKE3(in[i*size .. (i+1)*size]) for arbitrary byte size.
void hash_many_launch(const uint8_t* in, uint64_t size, uint64_t count, uint8_t* out, hipStream_t st);
// Build the flat tree in nodes (leaves already in nodes[0..n), padded to 2^L).
// With n_leaves >= 2 and root_copy set, the root is also written there.
void merkle_build_launch(uint8_t* nodes, uint64_t n_leaves, hipStream_t st, uint8_t* root_copy = nullptr);
// The SHA-256 engine (hash_id 2): the same three operations with SHA-256 of
// the message bytes in place of BLAKE3 (hash/sha2_engine.hpp:47-81).
void sha256_leaves_launch(const uint64_t* matrix, uint64_t rows, uint32_t row_u64, uint8_t* leaves,
                          hipStream_t st);
void sha256_hash_many_launch(const uint8_t* in, uint64_t size, uint64_t count, uint8_t* out, hipStream_t st);
void sha256_merkle_build_launch(uint8_t* nodes, uint64_t n_leaves, hipStream_t st, uint8_t* root_copy = nullptr);
// Many BLAKE3 trees built level by level together (a batch's trees): the
// plan lists, per launch, one segment per tree that takes part (its node
// array as a byte offset from a common base, the child level's node offset
// and width, the launch's first workgroup for it).  Leaf counts must be
// powers of two.  Upload `segs` to the device and pass its address to
// merkle_multi_launch, which issues the plan's launches in order.
struct MerkleSeg {
    uint64_t nodes, child_off, width;
    uint32_t block0, pad;
};
struct MerkleMultiLaunch {
    uint32_t kind, seg0, nseg, blocks;
};
int merkle_multi_plan(const uint64_t* nodes_off, const uint64_t* leaves, size_t count, std::vector<MerkleSeg>& segs,
                      std::vector<MerkleMultiLaunch>& launches);
void merkle_multi_launch(uint8_t* base, const MerkleSeg* d_segs, const std::vector<MerkleMultiLaunch>& launches,
                         hipStream_t st);
// out[i] = nodes[idx[i]] (32 bytes each)
void gather_hashes_launch(const uint8_t* nodes, const uint64_t* idx, uint64_t count, uint8_t* out, hipStream_t st);
void fill_zero_launch(uint8_t* p, uint64_t bytes, hipStream_t st);
// dst[0..bytes) = src[0..bytes) with a streaming kernel (bytes % 16 == 0, 16-byte aligned)
void copy_launch(void* dst, const void* src, uint64_t bytes, hipStream_t st);

// sumcheck over Ext3 (a may be base-field, lifted on the fly)
// partials must hold 2*3*SUMCHECK_MAX_BLOCKS u64.  With device_sum, the
// result c0[3], c2[3] is left in partials[0..6) on the device; without it,
// block b's (c0, c2) is left in partials[6b..6b+6) for the caller to sum.
// Returns the number of blocks.
constexpr int SUMCHECK_MAX_BLOCKS = 2048;

unsigned sumcheck_compute_launch(const uint64_t* a, bool a_base, const uint64_t* b, uint64_t n,
                                 uint64_t* partials, hipStream_t st, bool device_sum = true);
// compute + both folds in one sweep; an Ext3 input may be folded in place
unsigned sumcheck_round_launch(const uint64_t* a, bool a_base, const uint64_t* b, uint64_t n, const uint64_t r[3],
                               uint64_t* a_out, uint64_t* b_out, uint64_t* partials, hipStream_t st,
                               bool device_sum = true);
// the last round of a sumcheck: a (base or Ext3) into a_out and Ext3 b into
// b_out, both folded by r in one launch
void sumcheck_fold2_launch(const uint64_t* a, bool a_base, uint64_t* a_out, const uint64_t* b, uint64_t* b_out,
                           uint64_t n, const uint64_t r[3], hipStream_t st);
void sumcheck_fold_launch(const uint64_t* in, bool in_base, uint64_t n, const uint64_t r[3],
                          uint64_t* out, hipStream_t st);

// sum_i f_i z^i over Ext3 (f base or Ext3 AoS): block b's share lands in
// partials[3b..3b+3) and the caller sums the returned number of blocks.
// zp.w[3l..3l+3) = z^(2^l) for l < HORNER_POWS, passed by value.
constexpr int HORNER_MAX_BLOCKS = 1024;
constexpr uint32_t HORNER_MAX_LOGT = 18;
constexpr uint32_t HORNER_POWS = 32;
struct HornerPowers {
    uint64_t w[3 * HORNER_POWS];
};
unsigned horner_launch(const uint64_t* f, bool f_ext3, uint64_t n, const HornerPowers& zp, uint64_t* partials,
                       hipStream_t st);

// ------------------------------------------------------------- prove loop
// Up to POW_BATCH_MAX nonce searches in one launch.  Job k scans nonces
// base + i, i < (block0_{k+1} - block0_k) * 256 * per_thread, for the smallest
// one whose BLAKE3(ch || LE64(nonce) || 0^24) read as LE64 is <= threshold.
// best[k] (~0 on entry) is job k's running minimum that lets lanes stop early;
// reset[0..POW_BATCH_MAX) (the other slot set) is set to ~0 for the next
// launch; block_min[b] gets block b's smallest hit or ~0.
constexpr uint32_t POW_BATCH_MAX = 16;
struct PowJobDesc {
    uint32_t ch[8];
    uint64_t threshold, base;
    uint32_t block0, per_thread;
};
struct PowBatchArgs {
    PowJobDesc job[POW_BATCH_MAX];
    uint32_t njobs, total_blocks;
    uint32_t hash;  // HASH_BLAKE3 or HASH_SHA2, for every job of the launch
};
void pow_batch_launch(const PowBatchArgs& a, unsigned long long* d_best, unsigned long long* d_reset,
                      uint64_t* block_min, hipStream_t st);
void fill_u64_launch(uint64_t* p, uint64_t n, uint64_t v, hipStream_t st);
// One lane stores seq into *flag_d (mapped host memory) after all earlier work
// of the stream, with a system-scope release.
void signal_launch(uint64_t* flag_d, uint64_t seq, hipStream_t st);

// Looks up, with `device` current, the device address of the file-scope
// scratch array of whir_dev_sumcheck_round_noalloc (once per device).
int noalloc_scratch_init(int device);

// Weight terms sum_t s_t w_t(i) over i < 2^nv, factored as hi_t[i >> L] * lo_t[i & (2^L-1)].
// Descriptor of a term, TERM_WORDS u64: kind (0 eq, 1 power), s (Ext3),
// z (Ext3, power terms), z^(2^L) (Ext3, power terms), offset of the eq point
// in the coordinate array (Ext3 elements, nv coordinates, first = top bit).
constexpr uint32_t TERM_WORDS = 11;
// Both groups in one launch.  Ext3 group: hi_e ne * 2^(nv-L) Ext3, lo_e
// ne * 2^L Ext3; base group (power terms with z in the base field): hi_b
// nb * 2^(nv-L) Ext3, lo_b nb * 2^L u64.
void term_tables_launch(const uint64_t* d_desc_e, uint32_t ne, const uint64_t* d_desc_b, uint32_t nb,
                        const uint64_t* d_coords, uint32_t nv, uint32_t L, uint64_t* hi_e, uint64_t* lo_e,
                        uint64_t* hi_b, uint64_t* lo_b, uint32_t* hl, hipStream_t st);
// bytes of the limb form of the hi tables of `terms` terms (hl above; 0 when
// the sweep does not use it, 2^L < 256): written by term_tables_launch when
// hl is not null, read by term_accumulate_launch
uint64_t term_limb_bytes(uint32_t nv, uint32_t L, uint32_t terms);
// acc[i] = (accumulate ? acc[i] : 0) + sum over both term groups
void term_accumulate_launch(uint64_t* d_acc, uint64_t n, uint32_t L, const uint64_t* hi_e, const uint64_t* lo_e,
                            uint32_t te, const uint64_t* hi_b, const uint64_t* lo_b, uint32_t tb, const uint32_t* hl,
                            bool accumulate, hipStream_t st);
// For each of `points` (1..MLE_MAX_POINTS) eq terms p, whose tables are
// hi + 3 (n >> L) p and lo + 3 2^L p: sum_i f_i hi_p[i >> L] lo_p[i & mask],
// per block into partials[3 (points b + p) ..] (at most 3 * points *
// SUMCHECK_MAX_BLOCKS / (points > 2 ? 2 : 1) words); returns the number of
// blocks (0 for a bad point count).
constexpr uint32_t MLE_MAX_POINTS = 4;
unsigned mle_eval_launch(const uint64_t* f, bool f_base, uint64_t n, uint32_t L, const uint64_t* hi,
                         const uint64_t* lo, uint32_t points, uint64_t* partials, hipStream_t st);
// fold (a, b) of length n (>= 4) by r to length n/2, and (c0, c2) of the folded
// pair into partials[0..6) (device_sum), or per block as sumcheck_compute_launch;
// Ext3 inputs may be folded in place.  Returns the number of blocks.
unsigned fold_compute_launch(const uint64_t* a, bool a_base, const uint64_t* b, uint64_t n, const uint64_t r[3],
                             uint64_t* a_out, uint64_t* b_out, uint64_t* partials, hipStream_t st,
                             bool device_sum = true);
// One opening: rows[k*row_u64 + c] = matrix[ridx[k]*row_u64 + c] and the node
// hashes hashes[32h..] = nodes[32*hidx[h]..] (16-byte aligned), one launch
void gather_open_launch(const uint64_t* matrix, uint64_t row_u64, const uint64_t* d_ridx, uint64_t count,
                        uint64_t* rows, const uint8_t* nodes, const uint64_t* d_hidx, uint64_t nh, uint8_t* hashes,
                        hipStream_t st);
// out[k*row_u64 + c] = matrix[idx[k]*row_u64 + c]
void gather_rows_launch(const uint64_t* matrix, uint64_t row_u64, const uint64_t* d_idx, uint64_t count,
                        uint64_t* out, hipStream_t st);
// out = f_0 + sum_{k>=1} coef_k f_k (Ext3 out, base f_k of length n, coef Ext3 on device)
void rlc_combine_launch(const uint64_t* f, uint64_t n, uint32_t nvec, const uint64_t* d_coef, uint64_t* out,
                        hipStream_t st);

}  // namespace whir_gpu
