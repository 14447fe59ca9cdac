/*
 * whir_gpu.h — C ABI of the MI355X (gfx950) WHIR commit-and-fold library
 * (libwhir_gpu.so).
 *
 * Plain C types only.  Field elements are canonical Goldilocks values in
 * [0, p), p = 2^64 - 2^32 + 1, one uint64_t each; cubic-extension elements
 * are three uint64_t in the reference's AoS order (c0, c1, c2); hashes are 32
 * raw bytes.  Matrices are row-major codeword_length x (num_polys * depth)
 * exactly as the reference's interleaved_rs_encode returns them.
 *
 * Status codes (every entry point returning int):
 *    0  done
 *   >0  not handled (unsupported shape, dispatch disabled): the caller runs
 *       its own CPU path, as the reference does when its try_gpu_* returns false
 *   <0  device error (HIP failure); whir_last_error() describes it
 *
 * Two families:
 *  - whir_gpu_*: host-pointer drop-ins for the reference's whir::cuda::gpu_*
 *    functions (cpp/cuda/cuda_integration.hpp); each copies inputs to the
 *    device, runs, and copies results back, like the reference.  They use a
 *    per-process default context on the current HIP device.
 *  - whir_dev_*: the device-resident forms.  Pointers are device addresses,
 *    work is ordered on the context's stream and is asynchronous unless the
 *    function returns host results (then it synchronizes the stream).
 *
 * Citations are relative to the reference repository (hxd77/WhirPlonk).
 */
#ifndef WHIR_GPU_H
#define WHIR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct whir_ctx whir_ctx;

/* ---------------------------------------------------------------- library */

/* Message of the last error on this thread ("" if none). */
const char* whir_last_error(void);
/* 1 when a gfx950 device is visible and usable. */
int whir_device_available(void);

/* Runtime gates, cpp/cuda/cuda_integration.hpp:51-84.  Dispatch starts
 * enabled unless WHIR_GPU_DISABLE (or the reference's WHIR_CUDA_DISABLE) is set
 * to a non-zero value; the threshold (default 65536 elements, :39) applies to
 * whir_gpu_ntt_batch only, as in the reference's NttEngine::ntt_batch. */
int whir_gpu_dispatch_enabled(void);
void whir_set_gpu_dispatch_enabled(int enabled);
size_t whir_gpu_ntt_threshold(void);
void whir_set_gpu_ntt_threshold(size_t threshold);
/* cuda_warmup, cuda_integration.hpp:90-92: create the default context and
 * load every kernel module. */
int whir_gpu_warmup(void);

/* ---------------------------------------------------------------- context */

/* An explicit context: device, stream, grow-only scratch pools and cached
 * twiddle tables (the reference's GpuPool singleton, :142-282, made explicit). */
int whir_ctx_create(int device, whir_ctx** out);
int whir_ctx_destroy(whir_ctx* ctx);
/* Run on a caller-provided stream (hipStream_t passed as void*); NULL restores
 * the context's own stream. */
int whir_ctx_set_stream(whir_ctx* ctx, void* hip_stream);
void* whir_ctx_stream(whir_ctx* ctx);
int whir_ctx_synchronize(whir_ctx* ctx);
/* Hash engine of this context's commits, leaves, trees and PoW searches
 * (whir_dev_commit_*, whir_dev_commit_leaves, whir_dev_merkle_build,
 * whir_dev_pow_search): WHIR_HASH_BLAKE3 (the default) or WHIR_HASH_SHA2.  The
 * reference selects it per tree with merkle_tree::make_config(EngineId, n)
 * (irs_commit.hpp:62-73); provers take it from whir_params.hash_id. */
int whir_ctx_set_hash(whir_ctx* ctx, uint32_t hash_id);

/* Device memory helpers for callers without their own allocator. */
int whir_dev_malloc(whir_ctx* ctx, size_t bytes, void** out);
int whir_dev_free(whir_ctx* ctx, void* p);
int whir_memcpy_h2d(whir_ctx* ctx, void* dst, const void* src, size_t bytes);
int whir_memcpy_d2h(whir_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a streaming kernel on the context's stream
 * (bytes a multiple of 16, 16-byte aligned pointers); asynchronous. */
int whir_dev_copy(whir_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------- host-pointer drop-in functions */

/* NttEngine::ntt_batch GPU branch, cooley_tukey.hpp:200-232 -> gpu_ntt_batch
 * (cuda_integration.hpp:346-386): forward NTT in place of every contiguous
 * block of `size` in values[0..total).  Natural order in and out. */
int whir_gpu_ntt_batch(uint64_t* values, size_t total, size_t size);
/* NttEngine::intt_batch (cooley_tukey.hpp:767-776): inverse without 1/N. */
int whir_gpu_intt_batch(uint64_t* values, size_t total, size_t size);

/* gpu_ntt_batch_transpose, cuda_integration.hpp:388-430 (the GPU branch of
 * NttEngine::try_gpu_ntt_batch_transpose, cooley_tukey.hpp:262-307, which
 * ark_ntt calls on the zero-padded interleaved matrix, mod_ntt.hpp:98):
 * forward NTT of every contiguous block of ntt_size in values[0..total),
 * then the rows x cols row-major result transposed in place of values
 * (values[c * rows + r] = NTT'd[r * cols + c]).  total = rows * cols, a
 * multiple of ntt_size (a power of two <= 2^26).  The reference's `roots`
 * argument, which its implementation ignores, is dropped.  Returns 1 below
 * the NTT threshold or with dispatch disabled, as the reference's try_ does. */
int whir_gpu_ntt_batch_transpose(uint64_t* values, size_t total, size_t ntt_size, size_t rows, size_t cols);

/* gpu_interleaved_rs_encode, cuda_integration.hpp:432-485 (and the CPU
 * interleaved_rs_encode, mod_ntt.hpp:133-141): coeffs num_polys x poly_size,
 * out row-major codeword_length x (num_polys * interleaving_depth). */
int whir_gpu_interleaved_rs_encode(const uint64_t* coeffs, uint64_t* out, size_t num_polys, size_t poly_size,
                                   size_t codeword_length, size_t interleaving_depth);

/* gpu_interleaved_rs_encode_blake3_matrix_leaves, cuda_integration.hpp:555-621:
 * as above plus out_hashes[i] = BLAKE3(row i) (32 bytes each).  The reference
 * refuses rows that are not a multiple of 64 bytes or exceed 1024 bytes; this
 * library hashes any row length (standard BLAKE3, as the reference's CPU path). */
int whir_gpu_interleaved_rs_encode_blake3_matrix_leaves(const uint64_t* coeffs, uint64_t* out_matrix,
                                                        uint8_t* out_hashes, size_t num_polys, size_t poly_size,
                                                        size_t codeword_length, size_t interleaving_depth);

/* gpu_interleaved_rs_encode_blake3_ext3_matrix_leaves, cuda_integration.hpp:702-788:
 * Ext3 vectors given as three SoA component arrays, codeword returned as
 * three SoA component matrices, leaves hashed over the AoS 24-byte limbs. */
int whir_gpu_interleaved_rs_encode_blake3_ext3_matrix_leaves(const uint64_t* coeff0, const uint64_t* coeff1,
                                                             const uint64_t* coeff2, uint64_t* out0, uint64_t* out1,
                                                             uint64_t* out2, uint8_t* out_hashes, size_t num_polys,
                                                             size_t poly_size, size_t codeword_length,
                                                             size_t interleaving_depth);

/* gpu_blake3_merkle_tree, cuda_integration.hpp:942-990: zero-padded binary
 * tree over num_leaves 32-byte leaves; out_nodes (optional) receives the full
 * witness, (2^(L+1) - 1) * 32 bytes, L = ceil(log2 num_leaves). */
int whir_gpu_blake3_merkle_tree(const uint8_t* leaves, size_t num_leaves, uint8_t* out_root, uint8_t* out_nodes);

/* gpu_sha256_merkle_tree, cuda_integration.hpp:888-925: the same tree with
 * parent = SHA-256(left || right). */
int whir_gpu_sha256_merkle_tree(const uint8_t* leaves, size_t num_leaves, uint8_t* out_root, uint8_t* out_nodes);

/* merkle_hint_node_indices, cuda_integration.hpp:854-885: node indices of the
 * open_path hints (merkle_tree.hpp:165-211) in order.  Returns the count;
 * writes them when out != NULL. */
size_t whir_merkle_hint_node_indices(size_t num_leaves, const uint64_t* indices, size_t num_indices, uint64_t* out);

/* gpu_blake3_merkle_open_path, cuda_integration.hpp:1047-1100: root and the
 * sibling hints of open_path for the given (unsorted, possibly repeated) leaf
 * indices.  out_hints must hold whir_merkle_hint_node_indices(...) * 32 bytes. */
int whir_gpu_blake3_merkle_open_path(const uint8_t* leaves, size_t num_leaves, const uint64_t* indices,
                                     size_t num_indices, uint8_t* out_root, uint8_t* out_hints);
/* gpu_sha256_merkle_open_path, cuda_integration.hpp:995-1045: the same over
 * the SHA-256 tree. */
int whir_gpu_sha256_merkle_open_path(const uint8_t* leaves, size_t num_leaves, const uint64_t* indices,
                                     size_t num_indices, uint8_t* out_root, uint8_t* out_hints);

/* gpu_ood_evaluate_ext3_from_base, cuda_integration.hpp:1103-1149: for each
 * Ext3 point p (3 u64) and vector v, out[(p*num_vectors + v)*3 ..] =
 * sum_i coeff0[v*vector_size + i] * point_p^i  (mixed_univariate_evaluate). */
int whir_gpu_ood_evaluate_ext3_from_base(const uint64_t* coeff0, const uint64_t* points, uint64_t* out,
                                         size_t num_vectors, size_t vector_size, size_t num_points);
/* gpu_ood_evaluate_ext3_from_ext3, cuda_integration.hpp:1151-1205: SoA Ext3
 * coefficients. */
int whir_gpu_ood_evaluate_ext3_from_ext3(const uint64_t* coeff0, const uint64_t* coeff1, const uint64_t* coeff2,
                                         const uint64_t* points, uint64_t* out, size_t num_vectors,
                                         size_t vector_size, size_t num_points);

/* gpu_encode_to_bytes, cuda_integration.hpp:1208-1217: u64 -> 8 LE bytes. */
int whir_gpu_encode_to_bytes(const uint64_t* values, uint8_t* out, size_t count);

/* Blake3::hash_many, hash/blake3_engine.hpp:90-199: out[i] = BLAKE3 of the
 * i-th `size`-byte message (any size, including 0 and > 1024). */
int whir_gpu_blake3_hash_many(size_t size, const uint8_t* input, size_t count, uint8_t* output);
/* Sha2::hash_many, hash/sha2_engine.hpp:62-79: out[i] = SHA-256 of the i-th
 * `size`-byte message (any size, including 0). */
int whir_gpu_sha256_hash_many(size_t size, const uint8_t* input, size_t count, uint8_t* output);

/* compute_sumcheck_polynomial over Ext3, algebra/sumcheck.hpp:43-98 (no GPU
 * counterpart in the reference): a, b are n Ext3 AoS values (a_is_base: a is
 * n base-field values, lifted).  Writes c0[3], c2[3]. */
int whir_gpu_sumcheck_compute(const uint64_t* a, int a_is_base, const uint64_t* b, size_t n, uint64_t* out_c0,
                              uint64_t* out_c2);
/* fold, algebra/sumcheck.hpp:103-116: out[i] = v[i] + (v[i+n/2] - v[i]) * r,
 * i < n/2, Ext3 out (v base when v_is_base). */
int whir_gpu_sumcheck_fold(const uint64_t* v, int v_is_base, size_t n, const uint64_t* r, uint64_t* out);

/* ----------------------------------------------- device-resident functions */

/* Interleaved RS encode on device buffers (row-major output). */
int whir_dev_rs_encode(whir_ctx* ctx, const uint64_t* d_coeffs, size_t num_polys, size_t poly_size,
                       size_t codeword_length, size_t interleaving_depth, uint64_t* d_out);
/* Ext3 AoS in and out. */
int whir_dev_rs_encode_ext3(whir_ctx* ctx, const uint64_t* d_coeffs, size_t num_polys, size_t poly_size,
                            size_t codeword_length, size_t interleaving_depth, uint64_t* d_out);
/* In-place forward (inverse != 0: inverse, unscaled) NTT of blocks of size
 * (a power of two <= 2^26). */
int whir_dev_ntt_batch(whir_ctx* ctx, uint64_t* d_values, size_t total, size_t size, int inverse);
/* whir_gpu_ntt_batch_transpose on device buffers, out of place: d_values is
 * not modified and must not overlap d_out (rows * cols values each). */
int whir_dev_ntt_batch_transpose(whir_ctx* ctx, const uint64_t* d_values, uint64_t* d_out, size_t total,
                                 size_t ntt_size, size_t rows, size_t cols);
/* d_leaves[i] = BLAKE3 of row i of a rows x row_u64 u64 matrix. */
int whir_dev_commit_leaves(whir_ctx* ctx, const uint64_t* d_matrix, size_t rows, size_t row_u64, uint8_t* d_leaves);
int whir_dev_blake3_hash_many(whir_ctx* ctx, size_t size, const uint8_t* d_input, size_t count, uint8_t* d_output);
/* Number of nodes of the tree over num_leaves leaves (merkle_tree.hpp:58-60). */
size_t whir_merkle_num_nodes(size_t num_leaves);
/* Build the tree in d_nodes (whir_merkle_num_nodes * 32 bytes) whose first
 * num_leaves entries already hold the leaves; root = last node. */
int whir_dev_merkle_build(whir_ctx* ctx, uint8_t* d_nodes, size_t num_leaves);
/* d_out[i] = d_nodes[d_node_idx[i]] (32 bytes each). */
int whir_dev_gather_hashes(whir_ctx* ctx, const uint8_t* d_nodes, const uint64_t* d_node_idx, size_t count,
                           uint8_t* d_out);
/* The commit hot path of irs_commit::Config::commit (irs_commit.hpp:344-460)
 * for base-field vectors: RS encode into d_matrix, row leaves into the first
 * codeword_length nodes of d_nodes, the tree, and the root copied to out_root
 * (host, 32 bytes; the call synchronizes the stream). */
int whir_dev_commit_base(whir_ctx* ctx, const uint64_t* d_coeffs, size_t num_polys, size_t poly_size,
                         size_t codeword_length, size_t interleaving_depth, uint64_t* d_matrix, uint8_t* d_nodes,
                         uint8_t* out_root);
/* Same for Ext3 AoS vectors (24-byte limbs per element in each leaf). */
int whir_dev_commit_ext3(whir_ctx* ctx, const uint64_t* d_coeffs, size_t num_polys, size_t poly_size,
                         size_t codeword_length, size_t interleaving_depth, uint64_t* d_matrix, uint8_t* d_nodes,
                         uint8_t* out_root);
/* Sumcheck round polynomial on device vectors; c0, c2 copied to the host. */
int whir_dev_sumcheck_compute(whir_ctx* ctx, const uint64_t* d_a, int a_is_base, const uint64_t* d_b, size_t n,
                              uint64_t* out_c0, uint64_t* out_c2);
/* Fold on device (d_out may equal d_v when v is Ext3). */
int whir_dev_sumcheck_fold(whir_ctx* ctx, const uint64_t* d_v, int v_is_base, size_t n, const uint64_t* r,
                           uint64_t* d_out);
/* sum_i f_i z^i, f base (f_is_ext3 = 0) or Ext3 AoS; z and out on the host. */
int whir_dev_horner(whir_ctx* ctx, const uint64_t* d_f, int f_is_ext3, size_t n, const uint64_t* z, uint64_t* out);
/* One prover round of sumcheck::Config::prove (sumcheck_protocol.hpp:108-140)
 * with the folding challenge r given by the caller instead of squeezed:
 * (c0, c2) of (a, b), then a' = fold(a, r) and b' = fold(b, r) into d_a_out,
 * d_b_out (n/2 Ext3 each).  c0, c2 are copied to the host. */
int whir_dev_sumcheck_round(whir_ctx* ctx, const uint64_t* d_a, int a_is_base, const uint64_t* d_b, size_t n,
                            const uint64_t* r, uint64_t* d_a_out, uint64_t* d_b_out, uint64_t* out_c0,
                            uint64_t* out_c2);

/* whir_dev_sumcheck_round that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-workgroup (c0, c2) partials go to a file-scope device
 * array of WHIR_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it
 * runs on until the process ends (its device address is looked up when a
 * context is created on that GPU), and c0, c2 are left on the device in
 * d_c0c2[6] (c0 then c2, Ext3 AoS) instead of being copied to the host.  The
 * call is stream-ordered on ctx's stream and does not synchronize; the folds
 * are as in whir_dev_sumcheck_round.  n is at most WHIR_NOALLOC_SUMCHECK_MAX_N
 * (larger n is refused with status 3).  Every call on one GPU shares that
 * GPU's scratch, which keeps the last call's partials (each call overwrites
 * all it reads): calls that overlap on one GPU, from different streams or
 * threads, race. */
#define WHIR_NOALLOC_SCRATCH_BYTES 98304u
#define WHIR_NOALLOC_SUMCHECK_MAX_N (1ull << 32)
int whir_dev_sumcheck_round_noalloc(whir_ctx* ctx, const uint64_t* d_a, int a_is_base, const uint64_t* d_b, size_t n,
                                    const uint64_t* r, uint64_t* d_a_out, uint64_t* d_b_out, uint64_t* d_c0c2);

/* ------------------------------------------------------------ prove loop */

/* ProtocolParameters (parameters.hpp:25-34).  hash_id selects the Merkle,
 * row-leaf and PoW hash engine: WHIR_HASH_BLAKE3 (hash/blake3_engine.hpp:50,
 * the CLI's default) or WHIR_HASH_SHA2 (hash/sha2_engine.hpp:40, the CLI's
 * --hash Sha2). */
#define WHIR_HASH_BLAKE3 1u
#define WHIR_HASH_SHA2 2u
typedef struct whir_params {
    int unique_decoding;
    uint32_t starting_log_inv_rate;
    uint32_t initial_folding_factor;
    uint32_t folding_factor;
    uint32_t security_level;
    uint32_t pow_bits;
    uint32_t batch_size;
    uint32_t hash_id;
} whir_params;

/* The CLI's defaults (main.cpp Args: rate 2^-1, folding factor 4, 128-bit
 * security, 20 PoW bits, one vector, BLAKE3, list decoding). */
void whir_params_default(whir_params* params);

/* whir::Config::from_params (protocols/whir/whir.hpp:118-224) for vectors of
 * `size` coefficients over the Goldilocks cubic extension, flattened:
 *   [n_rounds | initial irs (7) | initial sumcheck size, rounds, PoW threshold |
 *    skip PoW threshold | per round: irs (7), sumcheck size, rounds, PoW
 *    threshold, round PoW threshold | final sumcheck size, rounds, PoW
 *    threshold | final PoW threshold]
 * irs = num_vectors, vector_size, codeword_length, interleaving_depth,
 * in_domain_samples, out_domain_samples, johnson_slack (IEEE-754 bits).
 * Thresholds are UINT64_MAX for no PoW.  Returns the number of words (0 on
 * invalid parameters); writes min(cap, words) words when out is non-null. */
size_t whir_config_flat(uint64_t size, const whir_params* params, uint64_t* out, size_t cap);

/* The CLI's domain separator (main.cpp:176-215): protocol_id[64] =
 * SHA3-512(CBOR map of the parameters), session_id[32] = SHA3-256(CBOR text
 * `session`, or the CLI's own text when session is NULL). */
int whir_domain_separator_cli(const whir_params* params, const char* session, uint8_t* protocol_id,
                              uint8_t* session_id);

/* A prover for one proof: ProverState::from_ds (transcript.hpp:606-611) with
 * the domain separator and instance bytes, plus the derived configuration.
 * Work runs on ctx's stream; calls on one prover (or one ctx) are serialised. */
typedef struct whir_prover whir_prover;
int whir_prover_create(whir_ctx* ctx, const whir_params* params, uint64_t size, const uint8_t* protocol_id,
                       const uint8_t* session_id, const uint8_t* instance, size_t instance_len,
                       whir_prover** out);
/* The same with all of the proof's request-sized device memory (codewords,
 * trees, folded vector, covector, LDE scratch) carved from a caller-owned,
 * 256-byte aligned region of at least whir_prover_workspace_bytes bytes
 * (the serving loop's HBM-budgeted admission uses this). */
int whir_prover_create_in(whir_ctx* ctx, const whir_params* params, uint64_t size, const uint8_t* protocol_id,
                          const uint8_t* session_id, const uint8_t* instance, size_t instance_len,
                          void* d_workspace, size_t workspace_bytes, whir_prover** out);
size_t whir_prover_workspace_bytes(const whir_params* params, uint64_t size);
int whir_prover_destroy(whir_prover* prover);
/* whir::Config::commit (protocols/whir/whir.hpp:304-309 ->
 * irs_commit.hpp:328-553): d_vectors holds batch_size base-field vectors of
 * `size` coefficients back to back on the device; it must stay valid until
 * whir_prover_prove returns. */
int whir_prover_commit(whir_prover* prover, const uint64_t* d_vectors, size_t num_vectors);
/* whir::Config::prove (protocols/whir/whir_prover.hpp:42-400) with
 * num_forms multilinear-evaluation constraints (MultilinearExtension,
 * algebra/linear_form.hpp): points = num_forms x log2(size) Ext3 coordinates
 * (host, AoS), evaluations = num_forms x batch_size Ext3 claimed values (host).
 * Optional outputs: the final claim's evaluation point (log2(size) Ext3) and
 * the constraint RLC coefficients (num_forms Ext3). */
int whir_prover_prove(whir_prover* prover, size_t num_forms, const uint64_t* points, const uint64_t* evaluations,
                      uint64_t* out_eval_point, uint64_t* out_forms_rlc);
/* ProverState::proof: the narg string and the hints.  Lengths are always
 * written; the bytes when the buffers are non-null and large enough. */
int whir_prover_proof(whir_prover* prover, uint8_t* narg, size_t narg_cap, size_t* narg_len, uint8_t* hints,
                      size_t hints_cap, size_t* hints_len);
/* PoW work of the proof: the sum over its searches of (nonce + 1), the
 * candidates a serial scan hashes. */
size_t whir_prover_pow_hashes(const whir_prover* prover);
/* Host wall time (ms, waits on the device included) per prover phase since
 * creation: 0 initial commit (encode, leaves, tree), 1 OOD evaluation,
 * 2 weight vectors (constraints, STIR), 3 sumcheck rounds, 4 PoW searches,
 * 5 round commits, 6 openings, 7 everything else. */
#define WHIR_PROVER_PHASES 8
int whir_prover_phase_times(whir_prover* prover, double* out_ms);

/* The prover transcript on its own (ProverState, transcript.hpp:595-662),
 * host only: SHAKE-128 duplex sponge, narg string and hints.  With null ids
 * the sponge starts empty (Shake128DuplexSponge, transcript.hpp:80-139);
 * otherwise protocol_id[64], session_id[32] and the instance bytes are
 * absorbed (DomainSeparator::absorb_into, transcript.hpp:531-538). */
typedef struct whir_transcript whir_transcript;
int whir_transcript_create(const uint8_t* protocol_id, const uint8_t* session_id, const uint8_t* instance,
                           size_t instance_len, whir_transcript** out);
int whir_transcript_destroy(whir_transcript* t);
/* public_message: absorb only */
int whir_transcript_absorb(whir_transcript* t, const uint8_t* data, size_t len);
/* prover_message of a fixed-size value: absorb + append to narg */
int whir_transcript_message(whir_transcript* t, const uint8_t* data, size_t len);
/* prover_hint of a fixed-size value (raw bytes) or of a vector (8-byte LE
 * element count, then the elements) */
int whir_transcript_hint(whir_transcript* t, const uint8_t* data, size_t len);
int whir_transcript_hint_vector(whir_transcript* t, const uint8_t* data, size_t count, size_t elem_bytes);
/* raw squeeze (verifier_message of a byte array) */
int whir_transcript_squeeze(whir_transcript* t, uint8_t* out, size_t len);
/* verifier_message<Goldilocks>: 40 bytes, LE integer mod p */
int whir_transcript_challenge_gl(whir_transcript* t, uint64_t* out);
/* challenge_indices (protocols/challenge_indices.hpp:96-119); returns the
 * number of indices written (num_leaves must be a power of two). */
size_t whir_transcript_indices(whir_transcript* t, size_t num_leaves, size_t count, int dedup, uint64_t* out);
int whir_transcript_proof(whir_transcript* t, uint8_t* narg, size_t narg_cap, size_t* narg_len, uint8_t* hints,
                          size_t hints_cap, size_t* hints_len);

/* find_nonce (proof_of_work.hpp:95-193): the smallest nonce whose
 * BLAKE3(challenge[32] || LE64(nonce) || 0^24), read as LE64, is <= threshold
 * (0 when threshold is UINT64_MAX). */
int whir_dev_pow_search(whir_ctx* ctx, const uint8_t* challenge, uint64_t threshold, uint64_t* out_nonce);
/* eval_eq (algebra/multilinear.hpp:120-133): d_acc[i] (+)= scalar * eq(point, i)
 * for i < 2^num_variables, point[0] pairing with the top index bit; Ext3 AoS.
 * accumulate = 0 overwrites. */
int whir_dev_eq_accumulate(whir_ctx* ctx, uint64_t* d_acc, size_t num_variables, const uint64_t* point,
                           const uint64_t* scalar, int accumulate);
/* geometric_accumulate (algebra/utilities.hpp:260-307): d_acc[i] +=
 * sum_j scalars[j] * points[j]^i, i < n (a power of two); Ext3 AoS. */
int whir_dev_geometric_accumulate(whir_ctx* ctx, uint64_t* d_acc, size_t n, const uint64_t* points,
                                  const uint64_t* scalars, size_t count);
/* mixed_multilinear_extend (algebra/multilinear.hpp:41-103): the multilinear
 * extension of d_f (2^num_variables entries, base or Ext3) at point. */
int whir_dev_mle_evaluate(whir_ctx* ctx, const uint64_t* d_f, int f_is_base, size_t num_variables,
                          const uint64_t* point, uint64_t* out);
/* No reference counterpart: whir_dev_mle_evaluate at `count` points (points:
 * count x num_variables Ext3, out: count Ext3), for callers that evaluate
 * several constraints of one vector (the prover's MLE constraints, a serving
 * loop's requests): up to four points share one pass over d_f and one wait. */
int whir_dev_mle_evaluate_many(whir_ctx* ctx, const uint64_t* d_f, int f_is_base, size_t num_variables,
                               const uint64_t* points, size_t count, uint64_t* out);

/* --------------------------------------------------------- batched commit */

/* LDE autotuner: times every valid radix split m = M1 * M2 (M2 = the second
 * pass's DFT size, which is also the number of rows, and so of row leaves,
 * each second-pass workgroup produces) under both forms of the fused-leaves
 * second pass (one workgroup per CU holding the whole 16 x 1024 tile, or two
 * per CU each staging half of it: 1024 or 512 leaves per workgroup round)
 * for one commit shape (message length, codeword rows, interleaved columns,
 * base or Ext3), and records the fastest pair for the process; later encodes
 * of that shape use it.  Every choice gives the same codeword.  out_lg_m2
 * receives log2 M2. */
int whir_autotune_lde(whir_ctx* ctx, size_t message_length, size_t codeword_length, size_t columns, int ext3,
                      double* out_best_ms, uint32_t* out_lg_m2);
/* The tuning table as 9-word records (8 shape-key words, then log2 M2 in bits
 * 0-7 and the second-pass form in bits 8-15: 0 the library default, 1 one
 * workgroup per CU, 2 two), for saving and loading across processes.  export
 * returns the word count. */
size_t whir_tuning_export(uint64_t* out, size_t cap);
int whir_tuning_import(const uint64_t* records, size_t words);

/* Many independent base-field polynomials of (possibly different) power-of-two
 * sizes, each committed as irs_commit::Config::commit does (RS at rate
 * 2^-log_inv_rate with interleaving depth 2^folding_factor, row leaves and
 * Merkle tree with the context's hash engine at create, BLAKE3 by default;
 * irs_commit.hpp:344-460).  create plans one device
 * region for all of them, autotunes each shape (autotune != 0), and captures
 * every encode, leaf and tree kernel, spread over num_branches streams, plus a
 * final root gather, into one HIP graph.  Fill the coefficients
 * (whir_batch_coeffs), then each whir_batch_run is one graph launch on the
 * context's stream; whir_batch_roots waits and copies the count x 32 roots. */
typedef struct whir_batch whir_batch;
int whir_batch_create(whir_ctx* ctx, size_t count, const uint64_t* sizes, uint32_t log_inv_rate,
                      uint32_t folding_factor, uint32_t num_branches, int autotune, whir_batch** out);
/* whir_batch_create with the batch form chosen by the caller: 1 the
 * level-synchronous form (BLAKE3 batches, folding factor 4, rate <= 1/2: every
 * polynomial with 2^10 to 2^18 coefficients per column encoded, leaves
 * included, by one first-pass and one second-pass launch together, then the
 * trees level by level across all polynomials), 0 one commit per polynomial
 * on the branch streams, -1 the library's choice (with autotune, both forms
 * are captured and the one whose graph replays faster is kept; without it,
 * the level-synchronous form where it applies).  All forms give the same
 * roots.  No reference counterpart: the reference commits one polynomial per
 * call (cuda_integration.hpp:555-621). */
int whir_batch_create_form(whir_ctx* ctx, size_t count, const uint64_t* sizes, uint32_t log_inv_rate,
                           uint32_t folding_factor, uint32_t num_branches, int autotune, int form,
                           whir_batch** out);
/* The form a batch runs (1 level-synchronous, 0 per polynomial) and, when
 * create timed both, their replay times in ms (else -1). */
int whir_batch_form(whir_batch* batch, double* per_poly_ms, double* level_ms);
int whir_batch_destroy(whir_batch* batch);
uint64_t* whir_batch_coeffs(whir_batch* batch, size_t k);
uint64_t* whir_batch_matrix(whir_batch* batch, size_t k);
uint8_t* whir_batch_nodes(whir_batch* batch, size_t k);
size_t whir_batch_rows(whir_batch* batch, size_t k);
size_t whir_batch_device_bytes(whir_batch* batch);
int whir_batch_run(whir_batch* batch);
/* The same work enqueued directly, without the graph. */
int whir_batch_run_eager(whir_batch* batch);
int whir_batch_roots(whir_batch* batch, uint8_t* out_roots);

/* --------------------------------------------------------------- serving */

/* A serving loop for commit+prove requests of any size under one parameter
 * set and the CLI domain separator: num_slots worker threads, each with its
 * own context and HIP stream, and one device arena of hbm_budget_bytes from
 * which every admitted request takes its block (prover workspace +
 * coefficients) until it completes.  A request waits until a slot is idle and
 * its block fits; smaller requests may pass a waiting one a bounded number of
 * times.  Each result is byte-identical to whir_prover_* on the same inputs. */
typedef struct whir_server whir_server;
int whir_server_create(int device, const whir_params* params, uint32_t num_slots, size_t hbm_budget_bytes,
                       whir_server** out);
/* Finishes every submitted request, then stops the workers. */
int whir_server_destroy(whir_server* server);
/* Device bytes a request of `size` coefficients holds while it runs. */
size_t whir_server_request_bytes(whir_server* server, uint64_t size);
/* Queue a request: batch_size vectors of `size` base-field coefficients
 * (host; copied at submit, or with borrow != 0 read in place, so they must
 * stay valid until the request is waited for), num_forms MLE points (host,
 * num_forms x log2(size) Ext3, copied) and their claimed values (num_forms x
 * batch_size Ext3), or evaluations = NULL to have them computed on the
 * device.  Returns its id. */
int whir_server_submit(whir_server* server, uint64_t size, const uint64_t* coeffs, int borrow, size_t num_forms,
                       const uint64_t* points, const uint64_t* evaluations, uint64_t* out_id);
/* Pinned host memory (hipHostMalloc) for request inputs that are copied to
 * the device by DMA. */
int whir_host_malloc(size_t bytes, void** out);
int whir_host_free(void* ptr);
/* Block until request `id` is done.  Lengths are always written; with
 * buffers large enough the proof is copied and the request released (with
 * null buffers it is kept for a later call).  out_ms (optional, 3 doubles):
 * queue wait, service time and end-to-end latency in ms. */
int whir_server_wait(whir_server* server, uint64_t id, uint8_t* narg, size_t narg_cap, size_t* narg_len,
                     uint8_t* hints, size_t hints_cap, size_t* hints_len, double* out_ms);
/* out[11]: completed, running, queued, arena bytes in use, peak bytes in use,
 * budget, slots, batched proof-of-work launches, searches they ran, batched
 * encode launches (initial commits of BLAKE3 requests with 2^14 to 2^22
 * coefficients at folding factor 4, several requests' encodes in one
 * level-synchronous launch pair), encodes they ran. */
int whir_server_stats(whir_server* server, uint64_t* out);
/* Test hook for the serving loop's error path (no counterpart in the
 * reference, whose GPU layer has no serving loop): request `id` (one that
 * whir_server_submit has returned or will return) fails at its at_wait-th
 * device wait (1 = the first) with status -1 and the message "injected
 * failure at a device wait", leaving the copies and kernels it enqueued since
 * its previous wait in flight.  The loop then drains the slot's stream before
 * it releases the request's HBM block.  Takes effect only for a request not
 * yet admitted; at_wait = 0 clears it.  Host-side only: no kernel faults. */
int whir_server_inject_failure(whir_server* server, uint64_t id, uint32_t at_wait);

/* --------------------------------------------------------- stage timing */

/* Stage timing with HIP events on the context's stream (the stream the
 * kernels run on).  Stages: 0 RS encode (LDE), 1 row leaves, 2 Merkle tree,
 * 3 sumcheck round.  Totals accumulate until reset. */
int whir_ctx_enable_timing(whir_ctx* ctx, int on);
int whir_ctx_reset_timing(whir_ctx* ctx);
/* out_ms[4]: summed milliseconds per stage; out_counts[4]: timed launches. */
int whir_ctx_stage_times(whir_ctx* ctx, double* out_ms, uint64_t* out_counts);

#ifdef __cplusplus
}
#endif
#endif
