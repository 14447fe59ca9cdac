// integration_check.cpp — the call-site edits of the integration notes,
// compiled against the reference's own headers and types and run beside the
// reference's CPU functions.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `integration`) builds it
// with the reference's include directory and vendored hash sources where they
// lie, linked against whirplonk_amd/_build/libwhir_gpu.so, into
// oracle/_ref/integration_check.  tests/test_gpu_dropin.py runs it on a GPU
// box.  Each site_* function is one edit a maintainer makes inside the
// reference (the replacement call written with the reference's own
// std::span / std::vector / Goldilocks / Hash values, exactly as the
// integration notes show it), followed by the reference's CPU function on the
// same inputs; the program prints one line per site and exits non-zero when
// any result differs or the library refuses a call it should serve.
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

#include "whir_gpu.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <span>
#include <vector>

namespace {
using whir::algebra::Goldilocks;
using whir::algebra::GoldilocksExt3;
using Hash = whir::hash::Hash;
namespace mt = whir::protocols::merkle_tree;
namespace mc = whir::protocols::matrix_commit;
namespace ntt = whir::algebra::ntt;

int g_failures = 0;

void report(const char* site, bool ok, int rc) {
    std::printf("%s %s (status %d)%s%s\n", ok ? "ok  " : "FAIL", site, rc, rc ? ": " : "",
                rc ? whir_last_error() : "");
    if (!ok) ++g_failures;
}

// the reference tests' LCG (cpp/tests/dump_whir.cpp:43-50), reduced by from_u64
struct Lcg {
    uint64_t s;
    uint64_t next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return s;
    }
    Goldilocks gl() { return Goldilocks::from_u64(next()); }
    GoldilocksExt3 e3() {
        Goldilocks a = gl(), b = gl(), c = gl();
        return GoldilocksExt3(a, b, c);
    }
};

template <typename T>
bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

const whir::hash::Blake3& blake3() {
    static const whir::hash::Blake3 b;
    return b;
}
const whir::hash::Sha2& sha2() {
    static const whir::hash::Sha2 s;
    return s;
}

// NttEngine::ntt_batch (cooley_tukey.hpp:200-232): the GPU branch, then the CPU path
void site_ntt_batch() {
    Lcg g{1};
    const size_t size = 1 << 14, total = 16 * size;
    std::vector<Goldilocks> values(total);
    for (auto& v : values) v = g.gl();
    std::vector<Goldilocks> cpu = values;
    ntt::goldilocks_engine().ntt_batch(std::span<Goldilocks>{cpu}, size);
    // --- the edit ---
    int rc = whir_gpu_ntt_batch(reinterpret_cast<uint64_t*>(values.data()), values.size(), size);
    report("ntt_batch (cooley_tukey.hpp:204-232)", rc == 0 && same(values, cpu), rc);
}

// NttEngine::intt_batch (cooley_tukey.hpp:767-776)
void site_intt_batch() {
    Lcg g{2};
    const size_t size = 1 << 12, total = 32 * size;
    std::vector<Goldilocks> values(total);
    for (auto& v : values) v = g.gl();
    std::vector<Goldilocks> cpu = values;
    ntt::goldilocks_engine().intt_batch(std::span<Goldilocks>{cpu}, size);
    int rc = whir_gpu_intt_batch(reinterpret_cast<uint64_t*>(values.data()), values.size(), size);
    report("intt_batch (cooley_tukey.hpp:767-776)", rc == 0 && same(values, cpu), rc);
}

std::vector<std::vector<Goldilocks>> polys_of(uint64_t seed, size_t num_polys, size_t poly_size) {
    Lcg g{seed};
    std::vector<std::vector<Goldilocks>> p(num_polys, std::vector<Goldilocks>(poly_size));
    for (auto& v : p)
        for (auto& x : v) x = g.gl();
    return p;
}

std::vector<std::span<const Goldilocks>> spans_of(const std::vector<std::vector<Goldilocks>>& p) {
    std::vector<std::span<const Goldilocks>> s;
    for (const auto& v : p) s.emplace_back(v);
    return s;
}

// ark_ntt's try_gpu_ntt_batch_transpose (mod_ntt.hpp:98) on the zero-padded matrix
void site_ntt_batch_transpose() {
    const size_t num_polys = 2, poly_size = 1 << 14, depth = 16, cw = 1 << 12;
    auto polys = polys_of(3, num_polys, poly_size);
    auto sp = spans_of(polys);
    auto cpu = ntt::interleaved_rs_encode<Goldilocks>(
        ntt::goldilocks_engine(), std::span<const std::span<const Goldilocks>>{sp}, cw, depth);
    // mod_ntt.hpp:77-93: the padded matrix ark_ntt builds before its step 2
    const size_t m = poly_size / depth;
    std::vector<Goldilocks> result(num_polys * depth * cw, Goldilocks::zero());
    for (size_t k = 0; k < num_polys; ++k)
        for (size_t b = 0; b < depth; ++b)
            for (size_t i = 0; i < m; ++i) result[k * cw * depth + b * cw + i] = polys[k][b * m + i];
    // --- the edit (rows_before_transpose = num_polys * depth, cols = codeword_length) ---
    const size_t rows_before_transpose = num_polys * depth;
    int rc = whir_gpu_ntt_batch_transpose(reinterpret_cast<uint64_t*>(result.data()), result.size(), cw,
                                          rows_before_transpose, cw);
    report("ntt_batch_transpose (mod_ntt.hpp:98)", rc == 0 && same(result, cpu), rc);
}

// try_gpu_interleaved_rs_encode (cooley_tukey.hpp:310-557)
void site_rs_encode() {
    const size_t num_polys = 3, poly_size = 1 << 13, depth = 16, cw = 1 << 11;
    auto polys = polys_of(4, num_polys, poly_size);
    auto sp = spans_of(polys);
    auto cpu = ntt::interleaved_rs_encode<Goldilocks>(
        ntt::goldilocks_engine(), std::span<const std::span<const Goldilocks>>{sp}, cw, depth);
    std::vector<uint64_t> coeffs;
    for (const auto& p : polys)
        for (const auto& x : p) coeffs.push_back(x.as_canonical_u64());
    std::vector<Goldilocks> gpu(cw * num_polys * depth);
    int rc = whir_gpu_interleaved_rs_encode(coeffs.data(), reinterpret_cast<uint64_t*>(gpu.data()), num_polys,
                                            poly_size, cw, depth);
    report("interleaved_rs_encode (cooley_tukey.hpp:310-557)", rc == 0 && same(gpu, cpu), rc);
}

// irs_commit::Config::commit's encode-and-leaves step (irs_commit.hpp:382-415)
void site_rs_encode_leaves() {
    const size_t num_vectors = 1, vector_size = 1 << 16, depth = 16, codeword_length = 1 << 14;
    auto polys = polys_of(5, num_vectors, vector_size);
    auto sp = spans_of(polys);
    auto cpu_matrix = ntt::interleaved_rs_encode<Goldilocks>(
        ntt::goldilocks_engine(), std::span<const std::span<const Goldilocks>>{sp}, codeword_length, depth);
    std::vector<Hash> cpu_leaves(codeword_length);
    mc::commit_leaves<Goldilocks>(blake3(), std::span<const Goldilocks>{cpu_matrix}, num_vectors * depth,
                                  std::span<Hash>{cpu_leaves});
    // --- the edit ---
    std::vector<uint64_t> coeffs;
    for (const auto& p : polys)
        for (const auto& x : p) coeffs.push_back(x.as_canonical_u64());
    std::vector<Goldilocks> matrix;
    std::vector<Hash> leaves;
    bool leaves_ready = false;
    matrix.resize(codeword_length * num_vectors * depth);
    leaves.resize(codeword_length);
    int rc = whir_gpu_interleaved_rs_encode_blake3_matrix_leaves(
        coeffs.data(), reinterpret_cast<uint64_t*>(matrix.data()), reinterpret_cast<uint8_t*>(leaves.data()),
        num_vectors, vector_size, codeword_length, depth);
    if (rc == 0) leaves_ready = true;
    report("rs_encode + blake3 leaves (irs_commit.hpp:382-415)",
           leaves_ready && same(matrix, cpu_matrix) && same(leaves, cpu_leaves), rc);
}

// the Ext3 SoA variant (cooley_tukey.hpp:621-737 -> cuda_integration.hpp:702)
void site_rs_encode_leaves_ext3() {
    const size_t num_polys = 1, poly_size = 1 << 12, depth = 16, cw = 1 << 11;
    Lcg g{6};
    std::vector<GoldilocksExt3> poly(poly_size);
    for (auto& x : poly) x = g.e3();
    std::vector<std::span<const GoldilocksExt3>> sp{std::span<const GoldilocksExt3>{poly}};
    auto cpu_matrix = ntt::interleaved_rs_encode<GoldilocksExt3>(
        ntt::goldilocks_ext3_engine(), std::span<const std::span<const GoldilocksExt3>>{sp}, cw, depth);
    std::vector<Hash> cpu_leaves(cw);
    mc::commit_leaves<GoldilocksExt3>(blake3(), std::span<const GoldilocksExt3>{cpu_matrix}, num_polys * depth,
                                      std::span<Hash>{cpu_leaves});
    // the caller splits the limbs into SoA arrays (cooley_tukey.hpp:436-452)
    std::vector<uint64_t> c0(poly_size), c1(poly_size), c2(poly_size);
    for (size_t i = 0; i < poly_size; ++i) {
        c0[i] = poly[i].c0().as_canonical_u64();
        c1[i] = poly[i].c1().as_canonical_u64();
        c2[i] = poly[i].c2().as_canonical_u64();
    }
    const size_t nout = cw * num_polys * depth;
    std::vector<uint64_t> o0(nout), o1(nout), o2(nout);
    std::vector<Hash> leaves(cw);
    int rc = whir_gpu_interleaved_rs_encode_blake3_ext3_matrix_leaves(
        c0.data(), c1.data(), c2.data(), o0.data(), o1.data(), o2.data(), reinterpret_cast<uint8_t*>(leaves.data()),
        num_polys, poly_size, cw, depth);
    std::vector<GoldilocksExt3> matrix(nout);  // re-interleaved (cooley_tukey.hpp:478-481)
    for (size_t i = 0; i < nout; ++i)
        matrix[i] = GoldilocksExt3(Goldilocks::from_u64(o0[i]), Goldilocks::from_u64(o1[i]), Goldilocks::from_u64(o2[i]));
    report("rs_encode + blake3 leaves, Ext3 SoA (cuda_integration.hpp:702)",
           rc == 0 && same(matrix, cpu_matrix) && same(leaves, cpu_leaves), rc);
}

std::vector<Hash> leaves_of(uint64_t seed, size_t n) {
    Lcg g{seed};
    std::vector<Hash> h(n);
    for (auto& x : h)
        for (size_t i = 0; i < 32; i += 8) {
            uint64_t v = g.next();
            std::memcpy(x.data() + i, &v, 8);
        }
    return h;
}

// irs_commit::commit's tree (irs_commit.hpp:442-459): gpu tree, else merkle_tree::commit
void site_merkle_tree(uint32_t hash_id) {
    const size_t n = (1 << 13) + 5;  // ragged: zero-padded to 2^14
    auto leaves = leaves_of(7 + hash_id, n);
    auto cfg = mt::make_config(hash_id == 2 ? whir::hash::ENGINE_ID_SHA2 : whir::hash::ENGINE_ID_BLAKE3, n);
    auto w = mt::build_tree(cfg, leaves, [hash_id](whir::EngineId) -> const whir::hash::HashEngine& {
        if (hash_id == 2) return sha2();
        return blake3();
    });
    Hash root;
    std::vector<Hash> nodes(cfg.num_nodes());
    int rc = hash_id == 2 ? whir_gpu_sha256_merkle_tree(reinterpret_cast<const uint8_t*>(leaves.data()), n,
                                                        root.data(), reinterpret_cast<uint8_t*>(nodes.data()))
                          : whir_gpu_blake3_merkle_tree(reinterpret_cast<const uint8_t*>(leaves.data()), n,
                                                        root.data(), reinterpret_cast<uint8_t*>(nodes.data()));
    report(hash_id == 2 ? "sha256 merkle tree (irs_commit.hpp:442-459)" : "blake3 merkle tree (irs_commit.hpp:442-459)",
           rc == 0 && root == w.nodes.back() && same(nodes, w.nodes), rc);
}

// irs_commit::open's path hints (irs_commit.hpp:680-703): gpu open path, else open_path
void site_open_path(uint32_t hash_id) {
    const size_t n = 1 << 12;
    auto leaves = leaves_of(9 + hash_id, n);
    auto cfg = mt::make_config(hash_id == 2 ? whir::hash::ENGINE_ID_SHA2 : whir::hash::ENGINE_ID_BLAKE3, n);
    auto w = mt::build_tree(cfg, leaves, [hash_id](whir::EngineId) -> const whir::hash::HashEngine& {
        if (hash_id == 2) return sha2();
        return blake3();
    });
    std::vector<size_t> indices = {5, 4000, 17, 5, 18, 4095, 0, 2048, 17};  // unsorted, repeated
    auto cpu = mt::open_path(cfg, w, std::span<const size_t>{indices});
    std::vector<uint64_t> idx(indices.begin(), indices.end());
    size_t nh = whir_merkle_hint_node_indices(n, idx.data(), idx.size(), nullptr);
    std::vector<Hash> hints(nh);
    Hash root;
    int rc = hash_id == 2
                 ? whir_gpu_sha256_merkle_open_path(reinterpret_cast<const uint8_t*>(leaves.data()), n, idx.data(),
                                                    idx.size(), root.data(), reinterpret_cast<uint8_t*>(hints.data()))
                 : whir_gpu_blake3_merkle_open_path(reinterpret_cast<const uint8_t*>(leaves.data()), n, idx.data(),
                                                    idx.size(), root.data(), reinterpret_cast<uint8_t*>(hints.data()));
    report(hash_id == 2 ? "sha256 open path (irs_commit.hpp:680-703)" : "blake3 open path (irs_commit.hpp:680-703)",
           rc == 0 && root == w.nodes.back() && same(hints, cpu), rc);
}

// irs_commit::commit's OOD evaluation (irs_commit.hpp:484-537)
void site_ood() {
    const size_t num_vectors = 2, vector_size = 1 << 15, num_points = 3;
    Lcg g{11};
    std::vector<Goldilocks> vecs(num_vectors * vector_size);
    for (auto& x : vecs) x = g.gl();
    std::vector<GoldilocksExt3> points(num_points);
    for (auto& p : points) p = g.e3();
    whir::algebra::Basefield<GoldilocksExt3> emb;
    std::vector<GoldilocksExt3> cpu;
    for (size_t p = 0; p < num_points; ++p)
        for (size_t v = 0; v < num_vectors; ++v)
            cpu.push_back(whir::algebra::mixed_univariate_evaluate<whir::algebra::Basefield<GoldilocksExt3>>(
                emb, std::span<const Goldilocks>{vecs.data() + v * vector_size, vector_size}, points[p]));
    std::vector<GoldilocksExt3> ood(num_points * num_vectors);
    bool ood_values_ready = false;
    int rc = whir_gpu_ood_evaluate_ext3_from_base(
        reinterpret_cast<const uint64_t*>(vecs.data()), reinterpret_cast<const uint64_t*>(points.data()),
        reinterpret_cast<uint64_t*>(ood.data()), num_vectors, vector_size, num_points);
    if (rc == 0) ood_values_ready = true;
    report("ood_evaluate_ext3_from_base (irs_commit.hpp:484-537)", ood_values_ready && same(ood, cpu), rc);

    // Ext3 vectors as SoA components
    std::vector<GoldilocksExt3> ev(num_vectors * vector_size);
    for (auto& x : ev) x = g.e3();
    whir::algebra::Identity<GoldilocksExt3> id;
    std::vector<GoldilocksExt3> cpu3;
    for (size_t p = 0; p < num_points; ++p)
        for (size_t v = 0; v < num_vectors; ++v)
            cpu3.push_back(whir::algebra::mixed_univariate_evaluate<whir::algebra::Identity<GoldilocksExt3>>(
                id, std::span<const GoldilocksExt3>{ev.data() + v * vector_size, vector_size}, points[p]));
    std::vector<uint64_t> s0, s1, s2;
    for (const auto& x : ev) {
        s0.push_back(x.c0().as_canonical_u64());
        s1.push_back(x.c1().as_canonical_u64());
        s2.push_back(x.c2().as_canonical_u64());
    }
    std::vector<GoldilocksExt3> ood3(num_points * num_vectors);
    rc = whir_gpu_ood_evaluate_ext3_from_ext3(s0.data(), s1.data(), s2.data(),
                                              reinterpret_cast<const uint64_t*>(points.data()),
                                              reinterpret_cast<uint64_t*>(ood3.data()), num_vectors, vector_size,
                                              num_points);
    report("ood_evaluate_ext3_from_ext3 (cuda_integration.hpp:1151)", rc == 0 && same(ood3, cpu3), rc);
}

// matrix_commit::encode_into (matrix_commit.hpp:84-92)
void site_encode_to_bytes() {
    Lcg g{12};
    std::vector<Goldilocks> values(1 << 17);
    for (auto& x : values) x = g.gl();
    std::vector<uint8_t> cpu(values.size() * 8), out(values.size() * 8);
    mc::encode_into<Goldilocks>(std::span<const Goldilocks>{values}, std::span<uint8_t>{cpu});
    int rc = whir_gpu_encode_to_bytes(reinterpret_cast<const uint64_t*>(values.data()), out.data(), values.size());
    report("encode_to_bytes (matrix_commit.hpp:84-92)", rc == 0 && same(out, cpu), rc);
}

// Blake3::hash_many / Sha2::hash_many (blake3_engine.hpp:90-199, sha2_engine.hpp:62-79)
void site_hash_many() {
    Lcg g{13};
    for (size_t size : {0, 64, 100, 1024, 1500}) {
        const size_t count = 300;
        std::vector<uint8_t> in(size * count);
        for (auto& b : in) b = (uint8_t)(g.next() >> 56);
        std::vector<Hash> cpu_b(count), cpu_s(count), out_b(count), out_s(count);
        blake3().hash_many(size, std::span<const uint8_t>{in}, std::span<Hash>{cpu_b});
        sha2().hash_many(size, std::span<const uint8_t>{in}, std::span<Hash>{cpu_s});
        int rc1 = whir_gpu_blake3_hash_many(size, in.data(), count, reinterpret_cast<uint8_t*>(out_b.data()));
        int rc2 = whir_gpu_sha256_hash_many(size, in.data(), count, reinterpret_cast<uint8_t*>(out_s.data()));
        char name[96];
        std::snprintf(name, sizeof name, "hash_many blake3 + sha256, %zu-byte messages", size);
        report(name, rc1 == 0 && rc2 == 0 && same(out_b, cpu_b) && same(out_s, cpu_s), rc1 | rc2);
    }
}

// compute_sumcheck_polynomial / fold (algebra/sumcheck.hpp:43-116)
void site_sumcheck() {
    Lcg g{14};
    const size_t n = 1 << 14;
    std::vector<GoldilocksExt3> a(n), b(n);
    for (auto& x : a) x = g.e3();
    for (auto& x : b) x = g.e3();
    auto cpu = whir::algebra::compute_sumcheck_polynomial<GoldilocksExt3>(a, b);
    GoldilocksExt3 c0, c2;
    int rc = whir_gpu_sumcheck_compute(reinterpret_cast<const uint64_t*>(a.data()), 0,
                                       reinterpret_cast<const uint64_t*>(b.data()), n,
                                       reinterpret_cast<uint64_t*>(&c0), reinterpret_cast<uint64_t*>(&c2));
    report("compute_sumcheck_polynomial (sumcheck.hpp:43-98)", rc == 0 && c0 == cpu.first && c2 == cpu.second, rc);
    GoldilocksExt3 r = g.e3();
    std::vector<GoldilocksExt3> cpu_a = a, out(n / 2);
    whir::algebra::fold<GoldilocksExt3>(cpu_a, r);
    rc = whir_gpu_sumcheck_fold(reinterpret_cast<const uint64_t*>(a.data()), 0, n,
                                reinterpret_cast<const uint64_t*>(&r), reinterpret_cast<uint64_t*>(out.data()));
    report("fold (sumcheck.hpp:103-116)", rc == 0 && same(out, cpu_a), rc);
}

}  // namespace

int main() {
    static_assert(sizeof(Goldilocks) == 8 && sizeof(GoldilocksExt3) == 24 && sizeof(Hash) == 32,
                  "the edits reinterpret the reference's values as raw limbs and bytes");
    if (!whir_device_available()) {
        std::printf("FAIL no gfx950 device\n");
        return 2;
    }
    site_ntt_batch();
    site_intt_batch();
    site_ntt_batch_transpose();
    site_rs_encode();
    site_rs_encode_leaves();
    site_rs_encode_leaves_ext3();
    site_merkle_tree(1);
    site_merkle_tree(2);
    site_open_path(1);
    site_open_path(2);
    site_ood();
    site_encode_to_bytes();
    site_hash_many();
    site_sumcheck();
    std::printf("%s: %d failures\n", g_failures ? "FAILED" : "passed", g_failures);
    return g_failures ? 1 : 0;
}
