// ref_driver.cpp — a thin extern "C" driver over the reference C++ prover.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file against the
// reference's own headers and vendored C sources where they lie (the
// reference repository's cpp/include and cpp/third_party), producing
// oracle/_ref/libwhirref.so.  It is used to cross-check the C restatement in
// whir_oracle.c and as the "reference" CPU baseline in bench.py.  Nothing in
// the product links it.  Each function calls one reference entry point:
//
//   ref_rs_encode          algebra/ntt/mod_ntt.hpp:133-141 interleaved_rs_encode
//   ref_ntt_batch          algebra/ntt/cooley_tukey.hpp:200-259 NttEngine::ntt_batch
//   ref_commit_leaves_*    protocols/matrix_commit.hpp:169-195 commit_leaves
//   ref_merkle_build       protocols/merkle_tree.hpp:104-147 build_tree
//   ref_merkle_open        protocols/merkle_tree.hpp:165-211 open_path
//   ref_sumcheck_*         algebra/sumcheck.hpp:44-116
//   ref_horner_base_ext3   algebra/utilities.hpp:195-246 mixed_univariate_evaluate
//   ref_horner_ext3_ext3   the same over Ext3 coefficients (Identity embedding)
//   ref_pow_find           protocols/proof_of_work.hpp:95-193 find_nonce, BLAKE3 or SHA-256
//   ref_e3_pow             GoldilocksExt3::pow (algebra/goldilocks_ext3.hpp)
//   ref_commit_base        the irs_commit.hpp:344-460 sequence (encode, leaves, tree)
//                          without the transcript
//   ref_blake3_many        hash/blake3_engine.hpp:90-199 Blake3::hash_many (both branches)
//   ref_sha256_many        hash/sha2_engine.hpp:62-79 Sha2::hash_many
//   ref_*_sha2             the SHA-256 tree and row leaves (ENGINE_ID_SHA2)
//   ref_whir_prove_hash, ref_cli_ds_hash  the PCS flow and DS with --hash Sha2
//   ref_whir_verify_hash   the main.cpp:434-446 verify flow (receive_commitment, verify,
//                          FinalClaim::verify, check_eof) on a given proof
//   ref_whir_config        protocols/whir/whir.hpp:118-224 Config::from_params, flattened
//   ref_whir_prove(_params) the main.cpp:292-424 PCS flow: from_params, the CLI's domain
//                          separator (main.cpp:176-215), commit, MLE constraints, prove
#include "whir/algebra/embedding.hpp"
#include "whir/algebra/goldilocks.hpp"
#include "whir/algebra/goldilocks_ext3.hpp"
#include "whir/algebra/ntt/cooley_tukey_goldilocks.hpp"
#include "whir/algebra/ntt/mod_ntt.hpp"
#include "whir/algebra/sumcheck.hpp"
#include "whir/algebra/utilities.hpp"
#include "whir/hash/blake3_engine.hpp"
#include "whir/hash/sha2_engine.hpp"
#include "whir/protocols/matrix_commit.hpp"
#include "whir/protocols/merkle_tree.hpp"
#include "whir/protocols/proof_of_work.hpp"
#include "whir/protocols/whir/whir.hpp"
#include "whir/transcript/transcript.hpp"

#include <cstdint>
#include <cstring>
#include <span>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {
using whir::algebra::Goldilocks;
using whir::algebra::GoldilocksExt3;
using Hash = whir::hash::Hash;

static_assert(sizeof(Goldilocks) == 8, "Goldilocks must be one u64");
static_assert(sizeof(GoldilocksExt3) == 24, "Ext3 must be three u64");

std::vector<Goldilocks> to_gl(const uint64_t* p, size_t n) {
    std::vector<Goldilocks> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = Goldilocks::from_u64(p[i]);
    return v;
}

std::vector<GoldilocksExt3> to_e3(const uint64_t* p, size_t n) {
    std::vector<GoldilocksExt3> v(n);
    for (size_t i = 0; i < n; ++i)
        v[i] = GoldilocksExt3(Goldilocks::from_u64(p[3 * i]), Goldilocks::from_u64(p[3 * i + 1]),
                              Goldilocks::from_u64(p[3 * i + 2]));
    return v;
}

void from_e3(const GoldilocksExt3& x, uint64_t* o) {
    o[0] = x.c0().as_canonical_u64();
    o[1] = x.c1().as_canonical_u64();
    o[2] = x.c2().as_canonical_u64();
}

const whir::hash::Blake3& blake3() {
    static const whir::hash::Blake3 b;
    return b;
}
const whir::hash::Sha2& sha2() {
    static const whir::hash::Sha2 s;
    return s;
}
}  // namespace

extern "C" {

int ref_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_ntt_batch(uint64_t* values, size_t total, size_t n) {
    auto& eng = whir::algebra::ntt::goldilocks_engine();
    std::span<Goldilocks> s{reinterpret_cast<Goldilocks*>(values), total};
    eng.ntt_batch(s, n);
}

void ref_rs_encode(const uint64_t* coeffs, size_t num_polys, size_t poly_size, size_t cw,
                   size_t depth, uint64_t* out) {
    std::vector<std::span<const Goldilocks>> polys;
    for (size_t k = 0; k < num_polys; ++k)
        polys.emplace_back(reinterpret_cast<const Goldilocks*>(coeffs) + k * poly_size, poly_size);
    auto m = whir::algebra::ntt::interleaved_rs_encode<Goldilocks>(
        whir::algebra::ntt::goldilocks_engine(),
        std::span<const std::span<const Goldilocks>>{polys}, cw, depth);
    std::memcpy(out, m.data(), m.size() * sizeof(uint64_t));
}

void ref_rs_encode_ext3(const uint64_t* coeffs, size_t num_polys, size_t poly_size, size_t cw,
                        size_t depth, uint64_t* out) {
    std::vector<std::span<const GoldilocksExt3>> polys;
    for (size_t k = 0; k < num_polys; ++k)
        polys.emplace_back(reinterpret_cast<const GoldilocksExt3*>(coeffs) + k * poly_size,
                           poly_size);
    auto m = whir::algebra::ntt::interleaved_rs_encode<GoldilocksExt3>(
        whir::algebra::ntt::goldilocks_ext3_engine(),
        std::span<const std::span<const GoldilocksExt3>>{polys}, cw, depth);
    std::memcpy(out, m.data(), m.size() * sizeof(GoldilocksExt3));
}

void ref_commit_leaves_base(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves) {
    std::vector<Hash> out(rows);
    whir::protocols::matrix_commit::commit_leaves<Goldilocks>(
        blake3(), std::span<const Goldilocks>{reinterpret_cast<const Goldilocks*>(matrix), rows * cols},
        cols, std::span<Hash>{out});
    std::memcpy(leaves, out.data(), rows * 32);
}

void ref_commit_leaves_ext3(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves) {
    std::vector<Hash> out(rows);
    whir::protocols::matrix_commit::commit_leaves<GoldilocksExt3>(
        blake3(),
        std::span<const GoldilocksExt3>{reinterpret_cast<const GoldilocksExt3*>(matrix), rows * cols},
        cols, std::span<Hash>{out});
    std::memcpy(leaves, out.data(), rows * 32);
}

void ref_merkle_build(const uint8_t* leaves, size_t n, uint8_t* nodes) {
    namespace mt = whir::protocols::merkle_tree;
    auto cfg = mt::make_config(whir::hash::ENGINE_ID_BLAKE3, n);
    std::vector<Hash> lv(n);
    if (n) std::memcpy(lv.data(), leaves, n * 32);
    auto w = mt::build_tree(cfg, std::move(lv),
                            [](whir::EngineId) -> const whir::hash::HashEngine& { return blake3(); });
    std::memcpy(nodes, w.nodes.data(), w.nodes.size() * 32);
}

// Sha2::hash_many (hash/sha2_engine.hpp:62-79)
void ref_sha256_many(size_t size, const uint8_t* in, size_t count, uint8_t* out) {
    std::vector<Hash> o(count);
    sha2().hash_many(size, std::span<const uint8_t>{in, size * count}, std::span<Hash>{o});
    if (count) std::memcpy(out, o.data(), count * 32);
}

// the SHA-256 tree (merkle_tree.hpp:104-147 with ENGINE_ID_SHA2)
void ref_merkle_build_sha2(const uint8_t* leaves, size_t n, uint8_t* nodes) {
    namespace mt = whir::protocols::merkle_tree;
    auto cfg = mt::make_config(whir::hash::ENGINE_ID_SHA2, n);
    std::vector<Hash> lv(n);
    if (n) std::memcpy(lv.data(), leaves, n * 32);
    auto w = mt::build_tree(cfg, std::move(lv),
                            [](whir::EngineId) -> const whir::hash::HashEngine& { return sha2(); });
    std::memcpy(nodes, w.nodes.data(), w.nodes.size() * 32);
}

// SHA-256 row leaves of a base-field matrix (matrix_commit.hpp:169-195)
void ref_commit_leaves_base_sha2(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves) {
    std::vector<Hash> out(rows);
    whir::protocols::matrix_commit::commit_leaves<Goldilocks>(
        sha2(), std::span<const Goldilocks>{reinterpret_cast<const Goldilocks*>(matrix), rows * cols}, cols,
        std::span<Hash>{out});
    std::memcpy(leaves, out.data(), rows * 32);
}

size_t ref_merkle_open(const uint8_t* nodes, size_t n_leaves, const uint64_t* idx, size_t n_idx,
                       uint8_t* hints) {
    namespace mt = whir::protocols::merkle_tree;
    auto cfg = mt::make_config(whir::hash::ENGINE_ID_BLAKE3, n_leaves);
    mt::Witness w;
    w.nodes.resize(cfg.num_nodes());
    std::memcpy(w.nodes.data(), nodes, w.nodes.size() * 32);
    std::vector<size_t> ix(idx, idx + n_idx);
    auto h = mt::open_path(cfg, w, std::span<const size_t>{ix});
    if (hints && !h.empty()) std::memcpy(hints, h.data(), h.size() * 32);
    return h.size();
}

void ref_sumcheck_compute_ext3(const uint64_t* a, const uint64_t* b, size_t n, uint64_t c0[3],
                               uint64_t c2[3]) {
    auto va = to_e3(a, n), vb = to_e3(b, n);
    auto r = whir::algebra::compute_sumcheck_polynomial<GoldilocksExt3>(va, vb);
    from_e3(r.first, c0);
    from_e3(r.second, c2);
}

void ref_sumcheck_fold_ext3(uint64_t* v, size_t n, const uint64_t r[3]) {
    auto vv = to_e3(v, n);
    GoldilocksExt3 w(Goldilocks::from_u64(r[0]), Goldilocks::from_u64(r[1]), Goldilocks::from_u64(r[2]));
    whir::algebra::fold<GoldilocksExt3>(vv, w);
    for (size_t i = 0; i < vv.size(); ++i) from_e3(vv[i], v + 3 * i);
}

void ref_horner_base_ext3(const uint64_t* coeffs, size_t n, const uint64_t z[3], uint64_t out[3]) {
    whir::algebra::Basefield<GoldilocksExt3> emb;
    GoldilocksExt3 p(Goldilocks::from_u64(z[0]), Goldilocks::from_u64(z[1]), Goldilocks::from_u64(z[2]));
    auto r = whir::algebra::mixed_univariate_evaluate<whir::algebra::Basefield<GoldilocksExt3>>(
        emb, std::span<const Goldilocks>{reinterpret_cast<const Goldilocks*>(coeffs), n}, p);
    from_e3(r, out);
}

void ref_horner_ext3_ext3(const uint64_t* coeffs, size_t n, const uint64_t z[3], uint64_t out[3]) {
    whir::algebra::Identity<GoldilocksExt3> emb;
    GoldilocksExt3 p(Goldilocks::from_u64(z[0]), Goldilocks::from_u64(z[1]), Goldilocks::from_u64(z[2]));
    auto c = to_e3(coeffs, n);
    auto r = whir::algebra::mixed_univariate_evaluate<whir::algebra::Identity<GoldilocksExt3>>(
        emb, std::span<const GoldilocksExt3>{c}, p);
    from_e3(r, out);
}

uint64_t ref_pow_find(uint32_t hash_id, const uint8_t* challenge, uint64_t threshold) {
    std::array<std::uint8_t, 32> ch;
    std::memcpy(ch.data(), challenge, 32);
    if (hash_id == 2) return whir::protocols::pow::find_nonce(sha2(), ch, threshold);
    return whir::protocols::pow::find_nonce(blake3(), ch, threshold);
}

void ref_e3_pow(const uint64_t a[3], uint64_t e, uint64_t out[3]) {
    GoldilocksExt3 x(Goldilocks::from_u64(a[0]), Goldilocks::from_u64(a[1]), Goldilocks::from_u64(a[2]));
    from_e3(x.pow(e), out);
}

void ref_blake3_many(size_t size, const uint8_t* input, size_t count, uint8_t* out) {
    std::vector<Hash> h(count);
    blake3().hash_many(size, std::span<const uint8_t>{input, size * count}, std::span<Hash>{h});
    if (count) std::memcpy(out, h.data(), count * 32);
}

// One plain commit of a single base-field polynomial: RS encode, row leaves,
// Merkle tree (irs_commit.hpp:382-460 CPU branch).  The root is written to
// out_root; the matrix to out_matrix when non-null.
void ref_commit_base(const uint64_t* coeffs, size_t poly_size, size_t cw, size_t depth,
                     uint64_t* out_matrix, uint8_t* out_root) {
    namespace mt = whir::protocols::merkle_tree;
    std::span<const Goldilocks> poly{reinterpret_cast<const Goldilocks*>(coeffs), poly_size};
    std::vector<std::span<const Goldilocks>> polys{poly};
    auto matrix = whir::algebra::ntt::interleaved_rs_encode<Goldilocks>(
        whir::algebra::ntt::goldilocks_engine(),
        std::span<const std::span<const Goldilocks>>{polys}, cw, depth);
    std::vector<Hash> leaves(cw);
    whir::protocols::matrix_commit::commit_leaves<Goldilocks>(
        blake3(), std::span<const Goldilocks>{matrix}, depth, std::span<Hash>{leaves});
    auto cfg = mt::make_config(whir::hash::ENGINE_ID_BLAKE3, cw);
    auto w = mt::build_tree(cfg, std::move(leaves),
                            [](whir::EngineId) -> const whir::hash::HashEngine& { return blake3(); });
    std::memcpy(out_root, w.nodes.back().data(), 32);
    if (out_matrix) std::memcpy(out_matrix, matrix.data(), matrix.size() * 8);
}

}  // extern "C"

namespace {
namespace whirp = whir::protocols::whir;
using Emb = whir::algebra::Basefield<GoldilocksExt3>;

whir::ProtocolParameters cli_params(uint32_t log_inv_rate, uint32_t folding_factor, uint32_t security_level,
                                    uint32_t pow_bits) {
    // main.cpp:315-323 (PCS, ConjectureList, batch 1, BLAKE3)
    whir::ProtocolParameters p;
    p.security_level = security_level;
    p.pow_bits = pow_bits;
    p.initial_folding_factor = folding_factor;
    p.folding_factor = folding_factor;
    p.unique_decoding = false;
    p.starting_log_inv_rate = log_inv_rate;
    p.batch_size = 1;
    p.hash_id = whir::hash::ENGINE_ID_BLAKE3;
    return p;
}

uint64_t dbits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}

template <typename C>
void put_irs(std::vector<uint64_t>& o, const C& c) {
    o.push_back(c.num_vectors);
    o.push_back(c.vector_size);
    o.push_back(c.codeword_length);
    o.push_back(c.interleaving_depth);
    o.push_back(c.in_domain_samples);
    o.push_back(c.out_domain_samples);
    o.push_back(dbits(c.johnson_slack));
}

// The CLI's domain separator: SHA3-512 of the CBOR map of the parameters and
// SHA3-256 of the CBOR session text (restated from main.cpp:176-215).
whir::transcript::DomainSeparator cli_ds(const whir::ProtocolParameters& p) {
    namespace tx = whir::transcript;
    std::vector<uint8_t> m;
    m.push_back(0xA0 + 8);
    tx::cbor_write_text(m, "unique_decoding");
    m.push_back(p.unique_decoding ? 0xF5 : 0xF4);
    tx::cbor_write_text(m, "starting_log_inv_rate");
    tx::cbor_write_uint(m, p.starting_log_inv_rate);
    tx::cbor_write_text(m, "initial_folding_factor");
    tx::cbor_write_uint(m, p.initial_folding_factor);
    tx::cbor_write_text(m, "folding_factor");
    tx::cbor_write_uint(m, p.folding_factor);
    tx::cbor_write_text(m, "security_level");
    tx::cbor_write_uint(m, p.security_level);
    tx::cbor_write_text(m, "pow_bits");
    tx::cbor_write_uint(m, p.pow_bits);
    tx::cbor_write_text(m, "batch_size");
    tx::cbor_write_uint(m, p.batch_size);
    tx::cbor_write_text(m, "hash_id");
    tx::cbor_array_header(m, p.hash_id.bytes().size());
    for (auto b : p.hash_id.bytes()) tx::cbor_write_uint(m, b);
    tx::DomainSeparator ds;
    sha3_512_hash(m.data(), m.size(), ds.protocol_id.data());
    std::vector<uint8_t> sess;
    tx::cbor_write_text(sess, "Example at src/bin/main.rs:133");
    sha3_256_hash(sess.data(), sess.size(), ds.session_id.data());
    return ds;
}
}  // namespace

extern "C" {

// Flattened Config::from_params: [n_rounds | initial irs (7) | initial sumcheck size, rounds,
// pow | skip pow | per round: irs (7), sumcheck size, rounds, pow, round pow | final sumcheck
// size, rounds, pow | final pow].  Returns the number of words written (out may be null).
size_t ref_whir_config(uint32_t num_variables, uint32_t log_inv_rate, uint32_t folding_factor,
                       uint32_t security_level, uint32_t pow_bits, uint64_t* out, size_t cap) {
    auto p = cli_params(log_inv_rate, folding_factor, security_level, pow_bits);
    auto c = whirp::Config<Emb>::from_params(std::size_t{1} << num_variables, p);
    std::vector<uint64_t> o;
    o.push_back(c.round_configs.size());
    put_irs(o, c.initial_committer);
    o.push_back(c.initial_sumcheck.initial_size);
    o.push_back(c.initial_sumcheck.num_rounds);
    o.push_back(c.initial_sumcheck.round_pow.threshold_val);
    o.push_back(c.initial_skip_pow.threshold_val);
    for (const auto& r : c.round_configs) {
        put_irs(o, r.irs_committer);
        o.push_back(r.sumcheck.initial_size);
        o.push_back(r.sumcheck.num_rounds);
        o.push_back(r.sumcheck.round_pow.threshold_val);
        o.push_back(r.pow.threshold_val);
    }
    o.push_back(c.final_sumcheck.initial_size);
    o.push_back(c.final_sumcheck.num_rounds);
    o.push_back(c.final_sumcheck.round_pow.threshold_val);
    o.push_back(c.final_pow.threshold_val);
    if (out) std::memcpy(out, o.data(), std::min(cap, o.size()) * 8);
    return o.size();
}

// One PCS run as the CLI does it (main.cpp:292-424) with caller-given base-field
// coefficients and Ext3 MLE constraint points (num_constraints x num_variables x 3).
// Writes the constraint evaluations (num_constraints x 3) and the proof bytes.
// Returns 0, or 1 when a buffer is too small (the lengths are set either way).
// General form: prm = {unique_decoding, starting_log_inv_rate, initial_folding_factor,
// folding_factor, security_level, pow_bits, batch_size}; coeffs holds batch_size vectors
// of 2^num_variables base-field coefficients; out_evals gets num_constraints x batch_size
// Ext3 values (constraint-major).  The domain separator is the CLI's for these parameters.
// hash_id: 1 BLAKE3, 2 SHA-256 (the CLI's --hash Blake3 / Sha2)
int ref_whir_prove_hash(const uint32_t* prm, uint32_t hash_id, uint32_t num_variables, const uint64_t* coeffs,
                        size_t num_constraints, const uint64_t* points, uint64_t* out_evals, uint8_t* narg,
                        size_t narg_cap, size_t* narg_len, uint8_t* hints, size_t hints_cap, size_t* hints_len) {
    namespace tx = whir::transcript;
    namespace alg = whir::algebra;
    whir::ProtocolParameters p;
    p.unique_decoding = prm[0] != 0;
    p.starting_log_inv_rate = prm[1];
    p.initial_folding_factor = prm[2];
    p.folding_factor = prm[3];
    p.security_level = prm[4];
    p.pow_bits = prm[5];
    p.batch_size = prm[6];
    p.hash_id = hash_id == 2 ? whir::hash::ENGINE_ID_SHA2 : whir::hash::ENGINE_ID_BLAKE3;
    const std::size_t size = std::size_t{1} << num_variables;
    const std::size_t nvec = p.batch_size;
    auto config = whirp::Config<Emb>::from_params(size, p);
    std::vector<std::vector<Goldilocks>> vecs(nvec, std::vector<Goldilocks>(size));
    std::vector<std::span<const Goldilocks>> spans;
    for (std::size_t v = 0; v < nvec; ++v) {
        for (std::size_t i = 0; i < size; ++i) vecs[v][i] = Goldilocks::from_u64(coeffs[v * size + i]);
        spans.push_back(std::span<const Goldilocks>{vecs[v]});
    }
    Emb emb{};
    std::vector<std::unique_ptr<alg::LinearForm<GoldilocksExt3>>> forms;
    std::vector<GoldilocksExt3> evaluations;
    for (std::size_t e = 0; e < num_constraints; ++e) {
        std::vector<GoldilocksExt3> pt(num_variables);
        for (std::size_t k = 0; k < num_variables; ++k) {
            const uint64_t* q = points + 3 * (e * num_variables + k);
            pt[k] = GoldilocksExt3(Goldilocks::from_u64(q[0]), Goldilocks::from_u64(q[1]),
                                   Goldilocks::from_u64(q[2]));
        }
        auto lf = std::make_unique<alg::MultilinearExtension<GoldilocksExt3>>(pt);
        for (std::size_t v = 0; v < nvec; ++v) {
            evaluations.push_back(lf->template evaluate<Emb>(emb, spans[v]));
            from_e3(evaluations.back(), out_evals + 3 * (e * nvec + v));
        }
        forms.push_back(std::move(lf));
    }
    auto ds = cli_ds(p);
    tx::Empty instance;
    auto ps = tx::ProverState::from_ds(ds, instance);
    auto witness = config.commit(ps, spans);
    std::vector<whir::protocols::irs_commit::Witness<Goldilocks, GoldilocksExt3>> wits;
    wits.push_back(std::move(witness));
    config.prove(ps, spans, std::span<const whir::protocols::irs_commit::Witness<Goldilocks, GoldilocksExt3>>{wits},
                 std::move(forms), std::span<const GoldilocksExt3>{evaluations});
    auto proof = std::move(ps).proof();
    *narg_len = proof.narg_string.size();
    *hints_len = proof.hints.size();
    if (proof.narg_string.size() > narg_cap || proof.hints.size() > hints_cap) return 1;
    std::memcpy(narg, proof.narg_string.data(), proof.narg_string.size());
    std::memcpy(hints, proof.hints.data(), proof.hints.size());
    return 0;
}

// 1 when the reference verifier accepts (narg, hints) for the claims
// f(points[e]) = evals[e * batch + v], 0 when it rejects.  Layouts as in
// ref_whir_prove_hash; the domain separator is the CLI's for these parameters.
int ref_whir_verify_hash(const uint32_t* prm, uint32_t hash_id, uint32_t num_variables, size_t num_constraints,
                         const uint64_t* points, const uint64_t* evals, const uint8_t* narg, size_t narg_len,
                         const uint8_t* hints, size_t hints_len) {
    namespace tx = whir::transcript;
    namespace alg = whir::algebra;
    whir::ProtocolParameters p;
    p.unique_decoding = prm[0] != 0;
    p.starting_log_inv_rate = prm[1];
    p.initial_folding_factor = prm[2];
    p.folding_factor = prm[3];
    p.security_level = prm[4];
    p.pow_bits = prm[5];
    p.batch_size = prm[6];
    p.hash_id = hash_id == 2 ? whir::hash::ENGINE_ID_SHA2 : whir::hash::ENGINE_ID_BLAKE3;
    const std::size_t size = std::size_t{1} << num_variables;
    const std::size_t nvec = p.batch_size;
    auto config = whirp::Config<Emb>::from_params(size, p);
    std::vector<std::unique_ptr<alg::LinearForm<GoldilocksExt3>>> forms;
    std::vector<const alg::LinearForm<GoldilocksExt3>*> refs;
    std::vector<GoldilocksExt3> evaluations;
    for (std::size_t e = 0; e < num_constraints; ++e) {
        std::vector<GoldilocksExt3> pt(num_variables);
        for (std::size_t k = 0; k < num_variables; ++k) {
            const uint64_t* q = points + 3 * (e * num_variables + k);
            pt[k] = GoldilocksExt3(Goldilocks::from_u64(q[0]), Goldilocks::from_u64(q[1]),
                                   Goldilocks::from_u64(q[2]));
        }
        forms.push_back(std::make_unique<alg::MultilinearExtension<GoldilocksExt3>>(std::move(pt)));
        refs.push_back(forms.back().get());
        for (std::size_t v = 0; v < nvec; ++v) {
            const uint64_t* q = evals + 3 * (e * nvec + v);
            evaluations.push_back(GoldilocksExt3(Goldilocks::from_u64(q[0]), Goldilocks::from_u64(q[1]),
                                                 Goldilocks::from_u64(q[2])));
        }
    }
    tx::ProverState::Proof proof;
    proof.narg_string.assign(narg, narg + narg_len);
    proof.hints.assign(hints, hints + hints_len);
    auto ds = cli_ds(p);
    tx::Empty instance;
    try {
        auto vs = tx::VerifierState::from_ds(ds, instance, proof);
        auto comm = config.receive_commitment(vs);
        std::vector<const whir::protocols::irs_commit::Commitment<GoldilocksExt3>*> cp{&comm};
        auto claim = config.verify(vs, cp, std::span<const GoldilocksExt3>{evaluations});
        return claim.verify(refs) && vs.check_eof() ? 1 : 0;
    } catch (...) {
        return 0;
    }
}

int ref_whir_prove_params(const uint32_t* prm, uint32_t num_variables, const uint64_t* coeffs,
                          size_t num_constraints, const uint64_t* points, uint64_t* out_evals, uint8_t* narg,
                          size_t narg_cap, size_t* narg_len, uint8_t* hints, size_t hints_cap,
                          size_t* hints_len) {
    return ref_whir_prove_hash(prm, 1, num_variables, coeffs, num_constraints, points, out_evals, narg, narg_cap,
                               narg_len, hints, hints_cap, hints_len);
}

// The CLI domain separator for prm (layout as above) and hash_id: protocol_id[64], session_id[32].
void ref_cli_ds_hash(const uint32_t* prm, uint32_t hash_id, uint8_t* protocol_id, uint8_t* session_id) {
    whir::ProtocolParameters p;
    p.unique_decoding = prm[0] != 0;
    p.starting_log_inv_rate = prm[1];
    p.initial_folding_factor = prm[2];
    p.folding_factor = prm[3];
    p.security_level = prm[4];
    p.pow_bits = prm[5];
    p.batch_size = prm[6];
    p.hash_id = hash_id == 2 ? whir::hash::ENGINE_ID_SHA2 : whir::hash::ENGINE_ID_BLAKE3;
    auto ds = cli_ds(p);
    std::memcpy(protocol_id, ds.protocol_id.data(), 64);
    std::memcpy(session_id, ds.session_id.data(), 32);
}

void ref_cli_ds(const uint32_t* prm, uint8_t* protocol_id, uint8_t* session_id) {
    ref_cli_ds_hash(prm, 1, protocol_id, session_id);
}

int ref_whir_prove(uint32_t num_variables, uint32_t log_inv_rate, uint32_t folding_factor,
                   uint32_t security_level, uint32_t pow_bits, const uint64_t* coeffs,
                   size_t num_constraints, const uint64_t* points, uint64_t* out_evals, uint8_t* narg,
                   size_t narg_cap, size_t* narg_len, uint8_t* hints, size_t hints_cap, size_t* hints_len) {
    uint32_t prm[7] = {0, log_inv_rate, folding_factor, folding_factor, security_level, pow_bits, 1};
    return ref_whir_prove_params(prm, num_variables, coeffs, num_constraints, points, out_evals, narg, narg_cap,
                                 narg_len, hints, hints_cap, hints_len);
}

}  // extern "C"
