/*
 * whir_oracle.h — CPU restatement of the WHIR commit-and-fold path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check and the CPU-baseline leg of bench.py call it to compute expected
 * values.  The product library (whirplonk_amd, libwhir_gpu.so) never links,
 * loads or calls it.
 *
 * Every function restates one reference function (cited file:line, relative
 * to the reference repository's cpp/include/whir/ unless noted).  All
 * arithmetic is exact integer arithmetic in canonical form [0, p), so any
 * correct implementation is bit-identical to the reference.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   field ops        golden_fields (base, ext3) + big-int brute force
 *   roots / NTT      closed form  A_k = sum_j a_j w^{jk}, brute-force DFT,
 *                    NTT(INTT(x)) = N x, the reference build (oracle/_ref)
 *   RS encode        golden_whir witness_matrix + closed form + oracle/_ref
 *   BLAKE3           golden_hash (64..1024 B) + golden_matrix_commit leaves
 *   Merkle tree      golden_merkle_tree roots and hints + oracle/_ref
 *   sumcheck         golden_utils (compute, fold) + golden_sumcheck replay
 *   Horner / OOD     golden_whir ood_matrix, golden_utils univariate_evaluate
 *   SHAKE-128 / SHA3 FIPS-202 known answers, golden_whir proof bytes
 */
#ifndef WHIR_ORACLE_H
#define WHIR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P 0xFFFFFFFF00000001ULL

/* ---- Goldilocks base field: algebra/goldilocks.hpp:54-202 ---- */
uint64_t or_gl_from_u64(uint64_t v);             /* :70-72  */
uint64_t or_gl_add(uint64_t a, uint64_t b);       /* :95-101 */
uint64_t or_gl_sub(uint64_t a, uint64_t b);       /* :108-113 */
uint64_t or_gl_neg(uint64_t a);                   /* :103-105 */
uint64_t or_gl_mul(uint64_t a, uint64_t b);       /* :116-124, reduce :174-191 */
uint64_t or_gl_pow(uint64_t a, uint64_t e);       /* :133-142 */
uint64_t or_gl_inv(uint64_t a);                   /* :146-148 */
/* primitive n-th root of unity as NttEngine picks it:
 * (7^eps)^(2^32/n), cooley_tukey.hpp:131-187, cooley_tukey_goldilocks.hpp:24-29 */
uint64_t or_gl_root(uint64_t n);

/* ---- Cubic extension Fp[X]/(X^3-2), AoS (c0,c1,c2): goldilocks_ext3.hpp:33-137 ---- */
void or_e3_add(const uint64_t* a, const uint64_t* b, uint64_t* out);
void or_e3_sub(const uint64_t* a, const uint64_t* b, uint64_t* out);
void or_e3_mul(const uint64_t* a, const uint64_t* b, uint64_t* out);   /* :83-89 */
void or_e3_pow(const uint64_t* a, uint64_t e, uint64_t* out);
void or_e3_inv(const uint64_t* a, uint64_t* out);                      /* :111-119 */

/* ---- NTT: algebra/ntt/cooley_tukey.hpp ---- */
/* forward NTT of each contiguous block of n (n a power of two), natural
 * order in and out: DIT radix-2, ntt_power_of_two :840-872 */
void or_ntt_batch(uint64_t* values, size_t total, size_t n);
/* inverse without 1/N scaling: reverse [1..n-1] then forward, :754-776 */
void or_intt_batch(uint64_t* values, size_t total, size_t n);
/* O(n^2) DFT straight from the definition A_k = sum_j a_j w_n^{jk} */
void or_dft_naive(const uint64_t* in, uint64_t* out, size_t n);

/* ---- Interleaved RS encode: algebra/ntt/mod_ntt.hpp:49-141 ----
 * coeffs: num_polys * poly_size, poly-major.  out: row-major cw x (num_polys*depth),
 * out[i*cols + k*depth + b] = sum_j f_k[b*m + j] * w_cw^{ij}, m = poly_size/depth */
void or_rs_encode(const uint64_t* coeffs, size_t num_polys, size_t poly_size,
                  size_t cw, size_t depth, uint64_t* out);
/* same over Ext3, AoS in and out (the reference's Ext3 engine, mod_ntt.hpp +
 * cooley_tukey_goldilocks.hpp:41-46: base roots embedded, so component-wise) */
void or_rs_encode_ext3(const uint64_t* coeffs, size_t num_polys, size_t poly_size,
                       size_t cw, size_t depth, uint64_t* out);

/* ---- BLAKE3 (standard, unkeyed, any length): hash/blake3_engine.hpp:90-199,
 * third_party/blake3 1.5.4 ---- */
void or_blake3(const uint8_t* in, size_t len, uint8_t out[32]);
/* hash_many(size, input, output): out[i] = BLAKE3(input[i*size..]) */
void or_blake3_many(size_t size, const uint8_t* in, size_t count, uint8_t* out);

/* ---- matrix commit: protocols/matrix_commit.hpp:73-195 ----
 * leaves[i] = BLAKE3(LE64 limbs of row i) */
void or_commit_leaves_base(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves);
void or_commit_leaves_ext3(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves);

/* ---- Merkle tree: protocols/merkle_tree.hpp ---- */
size_t or_merkle_layers(size_t n);                 /* layers_for_size :79-85 */
size_t or_merkle_num_nodes(size_t n);              /* Config::num_nodes :58-60 */
/* build_tree :104-147: nodes = [leaves padded to 2^L with zero hashes | L-1 | ... | root] */
void or_merkle_build(const uint8_t* leaves, size_t n, uint8_t* nodes);
/* open_path :165-211: returns the number of hints; writes them when hints != NULL */
size_t or_merkle_open(const uint8_t* nodes, size_t n_leaves,
                      const uint64_t* idx, size_t n_idx, uint8_t* hints);
/* verify_path :220-316 (used by tests as an independent consistency check) */
int or_merkle_verify(const uint8_t root[32], size_t n_leaves, const uint64_t* idx,
                     const uint8_t* leaf_hashes, size_t n_idx,
                     const uint8_t* hints, size_t n_hints);

/* ---- sumcheck algebra over Ext3: algebra/sumcheck.hpp ---- */
/* compute_sumcheck_polynomial :43-98, half-split */
void or_sumcheck_compute_ext3(const uint64_t* a, const uint64_t* b, size_t n,
                              uint64_t c0[3], uint64_t c2[3]);
/* same with a given in the base field and lifted (Basefield<Ext3>::map, embedding.hpp:43-48) */
void or_sumcheck_compute_base_ext3(const uint64_t* a, const uint64_t* b, size_t n,
                                   uint64_t c0[3], uint64_t c2[3]);
/* fold :103-116: v[i] = v[i] + (v[i+n/2]-v[i])*r, length halves (in place) */
void or_sumcheck_fold_ext3(uint64_t* v, size_t n, const uint64_t r[3]);
/* fold of a base-field vector by an Ext3 weight into an Ext3 output (lift then fold) */
void or_sumcheck_fold_base_ext3(const uint64_t* v, size_t n, const uint64_t r[3], uint64_t* out);

/* ---- Horner: algebra/utilities.hpp:195-246 (mixed_univariate_evaluate) ---- */
void or_horner_base_ext3(const uint64_t* coeffs, size_t n, const uint64_t point[3], uint64_t out[3]);
void or_horner_ext3_ext3(const uint64_t* coeffs, size_t n, const uint64_t point[3], uint64_t out[3]);
uint64_t or_horner_base(const uint64_t* coeffs, size_t n, uint64_t point);

/* ---- prove-path algebra over Ext3 (AoS) ---- */
/* eval_eq, algebra/multilinear.hpp:120-133 (recursive halving: point[0] splits
 * the accumulator's top half); acc has 2^k entries */
void or_eval_eq_ext3(uint64_t* acc, const uint64_t* point, size_t k, const uint64_t scalar[3]);
/* geometric_accumulate, algebra/utilities.hpp:260-307 (serial branch :300-306):
 * acc[i] += sum_j scalars[j] * points[j]^i */
void or_geometric_accumulate_ext3(uint64_t* acc, size_t n, const uint64_t* points, const uint64_t* scalars,
                                  size_t count);
/* mixed_multilinear_extend, algebra/multilinear.hpp:41-103 (the recursive
 * default branch for every k): evals base (is_base) or Ext3, 2^k entries */
void or_mle_evaluate(const uint64_t* evals, int is_base, const uint64_t* point, size_t k, uint64_t out[3]);
/* find_nonce, protocols/proof_of_work.hpp:95-193 (serial scan): the first n
 * with LE64(BLAKE3(ch || LE64(n) || 0^24)[0..8]) <= threshold; UINT64_MAX -> 0 */
uint64_t or_pow_find(const uint8_t ch[32], uint64_t threshold);

/* ---- seeded inputs: the LCG of cpp/tests/dump_*.cpp (e.g. dump_whir.cpp:43-50) ----
 * x <- x*6364136223846793005 + 1442695040888963407.  mode 0: from_u64 (x>=p -> x-p,
 * as the reference's dumpers); mode 1: reject x >= p (exactly uniform). */
void or_lcg_fill(uint64_t* state, uint64_t* out, size_t n, int mode);

#ifdef __cplusplus
}
#endif
#endif
