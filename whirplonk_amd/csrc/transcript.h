// transcript.h — host-side Fiat–Shamir transcript of the prover.
//
// Keccak-f[1600] / SHAKE-128 / SHA3 (FIPS 202) and the duplex-sponge prover
// state of transcript/transcript.hpp:
//   * absorb appends bytes to one running SHAKE-128 input and invalidates the
//     output reader (Shake128DuplexSponge::absorb, transcript.hpp:80-86);
//   * a squeeze after an absorb starts reading SHAKE-128(all absorbed bytes)
//     from its first byte; further squeezes continue the same output stream
//     (transcript.hpp:91-99);
//   * a prover message is absorbed as its raw little-endian bytes and the same
//     bytes are appended to the proof string; a vector message or hint carries
//     an 8-byte little-endian element count first (narg_serialize,
//     transcript.hpp:263-330); hints are appended without absorbing
//     (ProverState::prover_hint, transcript.hpp:622-625);
//   * a Goldilocks challenge is 40 squeezed bytes read as a little-endian
//     integer mod p; an Ext3 challenge is three of those (decode_from_bytes,
//     transcript.hpp:401-453); byte challenges are raw bytes.
// The domain separator absorbs protocol_id (64 B), session_id (32 B) and the
// instance bytes (DomainSeparator::absorb_into, transcript.hpp:531-538).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gl.h"

namespace whir_host {

// ------------------------------------------------------------- Keccak-f[1600]
struct KeccakTables {
    uint64_t rc[24];
    int rot[25];   // rotation of lane x + 5y
    int pi[25];    // destination lane of lane x + 5y under pi
    KeccakTables() {
        // round constants from the degree-8 LFSR x^8 + x^6 + x^5 + x^4 + 1
        uint8_t r = 1;
        auto lfsr_bit = [&r]() {
            int out = r & 1;
            r = (uint8_t)((r & 0x80) ? ((r << 1) ^ 0x71) : (r << 1));
            return out;
        };
        for (int i = 0; i < 24; ++i) {
            uint64_t c = 0;
            for (int j = 0; j < 7; ++j)
                if (lfsr_bit()) c |= 1ull << ((1 << j) - 1);
            rc[i] = c;
        }
        // rho offsets: walk (x, y) -> (y, 2x + 3y) from (1, 0)
        rot[0] = 0;
        int x = 1, y = 0;
        for (int t = 0; t < 24; ++t) {
            rot[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
            int nx = y, ny = (2 * x + 3 * y) % 5;
            x = nx;
            y = ny;
        }
        // pi: lane (x, y) moves to (y, 2x + 3y)
        for (int yy = 0; yy < 5; ++yy)
            for (int xx = 0; xx < 5; ++xx) pi[xx + 5 * yy] = yy + 5 * ((2 * xx + 3 * yy) % 5);
    }
};

inline const KeccakTables& keccak_tables() {
    static const KeccakTables t;
    return t;
}

inline uint64_t rotl64(uint64_t v, int s) { return s ? (v << s) | (v >> (64 - s)) : v; }

inline void keccak_f1600(uint64_t a[25]) {
    const KeccakTables& T = keccak_tables();
    uint64_t b[25], c[5], d[5];
    for (int round = 0; round < 24; ++round) {
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; ++i) b[T.pi[i]] = rotl64(a[i] ^ d[i % 5], T.rot[i]);
        for (int y = 0; y < 25; y += 5)
            for (int x = 0; x < 5; ++x) a[y + x] = b[y + x] ^ (~b[y + (x + 1) % 5] & b[y + (x + 2) % 5]);
        a[0] ^= T.rc[round];
    }
}

// A Keccak sponge with byte-granular absorb and squeeze (little-endian lanes).
struct Sponge {
    uint64_t st[25];
    size_t rate;  // bytes
    size_t pos = 0;
    uint8_t dsbyte;

    Sponge(size_t rate_bytes, uint8_t ds) : rate(rate_bytes), dsbyte(ds) { std::memset(st, 0, sizeof st); }

    void xor_byte(size_t i, uint8_t v) { st[i >> 3] ^= (uint64_t)v << (8 * (i & 7)); }
    uint8_t get_byte(size_t i) const { return (uint8_t)(st[i >> 3] >> (8 * (i & 7))); }

    void absorb(const uint8_t* p, size_t n) {
        for (size_t k = 0; k < n; ++k) {
            xor_byte(pos++, p[k]);
            if (pos == rate) {
                keccak_f1600(st);
                pos = 0;
            }
        }
    }
    // pad10*1 with the domain byte, then switch to squeezing at offset 0
    void finalize() {
        xor_byte(pos, dsbyte);
        xor_byte(rate - 1, 0x80);
        keccak_f1600(st);
        pos = 0;
    }
    void squeeze(uint8_t* out, size_t n) {
        for (size_t k = 0; k < n; ++k) {
            if (pos == rate) {
                keccak_f1600(st);
                pos = 0;
            }
            out[k] = get_byte(pos++);
        }
    }
};

inline void sha3_256(const uint8_t* p, size_t n, uint8_t out[32]) {
    Sponge s(136, 0x06);
    s.absorb(p, n);
    s.finalize();
    s.squeeze(out, 32);
}
inline void sha3_512(const uint8_t* p, size_t n, uint8_t out[64]) {
    Sponge s(72, 0x06);
    s.absorb(p, n);
    s.finalize();
    s.squeeze(out, 64);
}
inline void shake128(const uint8_t* p, size_t n, uint8_t* out, size_t m) {
    Sponge s(168, 0x1F);
    s.absorb(p, n);
    s.finalize();
    s.squeeze(out, m);
}

// ---------------------------------------------------------------- CBOR bits
inline void cbor_head(std::vector<uint8_t>& o, uint8_t major, uint64_t v) {
    major <<= 5;
    if (v < 24) {
        o.push_back((uint8_t)(major | v));
    } else if (v < 0x100) {
        o.push_back(major | 24);
        o.push_back((uint8_t)v);
    } else if (v < 0x10000) {
        o.push_back(major | 25);
        for (int s = 8; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
    } else if (v < 0x100000000ull) {
        o.push_back(major | 26);
        for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
    } else {
        o.push_back(major | 27);
        for (int s = 56; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
    }
}
inline void cbor_uint(std::vector<uint8_t>& o, uint64_t v) { cbor_head(o, 0, v); }
inline void cbor_text(std::vector<uint8_t>& o, const char* s) {
    size_t n = std::strlen(s);
    cbor_head(o, 3, n);
    o.insert(o.end(), (const uint8_t*)s, (const uint8_t*)s + n);
}

// ------------------------------------------------------------- prover state
class Transcript {
public:
    Transcript() : absorbed_(168, 0x1F), reader_(168, 0x1F) {}

    // DomainSeparator::absorb_into + ProverState::from_ds
    void start(const uint8_t protocol_id[64], const uint8_t session_id[32], const uint8_t* instance,
               size_t instance_len) {
        absorb(protocol_id, 64);
        absorb(session_id, 32);
        absorb(instance, instance_len);
    }

    void absorb(const uint8_t* p, size_t n) {
        has_reader_ = false;
        absorbed_.absorb(p, n);
    }
    void squeeze(uint8_t* out, size_t n) {
        if (!n) return;
        if (!has_reader_) {
            reader_ = absorbed_;
            reader_.finalize();
            has_reader_ = true;
        }
        reader_.squeeze(out, n);
    }

    // prover_message of a fixed-size value: absorbed and appended as is
    void message(const void* p, size_t n) {
        absorb((const uint8_t*)p, n);
        narg_.insert(narg_.end(), (const uint8_t*)p, (const uint8_t*)p + n);
    }
    void message_u64(uint64_t v) {
        uint8_t b[8];
        for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(v >> (8 * i));
        message(b, 8);
    }
    void message_e3(const gl::e3& v) {
        uint64_t w[3] = {v.c0, v.c1, v.c2};
        message(w, 24);
    }
    // prover_hint of a vector: element count, then the raw elements
    void hint_vector(const void* p, size_t count, size_t elem_bytes) {
        uint64_t len = count;
        append_le(hints_, len);
        hints_.insert(hints_.end(), (const uint8_t*)p, (const uint8_t*)p + count * elem_bytes);
    }
    // prover_hint of a fixed-size value (a 32-byte hash)
    void hint(const void* p, size_t n) { hints_.insert(hints_.end(), (const uint8_t*)p, (const uint8_t*)p + n); }

    // verifier_message<Goldilocks>: 40 bytes, little-endian, reduced mod p
    uint64_t challenge_gl() {
        uint8_t b[40];
        squeeze(b, 40);
        return bytes_mod_p(b, 40);
    }
    gl::e3 challenge_e3() {
        gl::e3 r;
        r.c0 = challenge_gl();
        r.c1 = challenge_gl();
        r.c2 = challenge_gl();
        return r;
    }

    // challenge_indices (protocols/challenge_indices.hpp:96-119): count indices,
    // each the big-endian value of ceil(log2(n)/8) squeezed bytes mod n.
    std::vector<size_t> indices(size_t num_leaves, size_t count, bool dedup) {
        std::vector<size_t> out;
        if (count == 0) return out;
        if (num_leaves == 1) {
            if (dedup) return {0};
            return std::vector<size_t>(count, 0);
        }
        size_t lg = 0;
        while ((size_t{1} << lg) < num_leaves) ++lg;
        const size_t nb = (lg + 7) / 8;
        std::vector<uint8_t> ent(count * nb);
        squeeze(ent.data(), ent.size());
        out.reserve(count);
        for (size_t i = 0; i < count; ++i) {
            size_t acc = 0;
            for (size_t b = 0; b < nb; ++b) acc = (acc << 8) | ent[i * nb + b];
            out.push_back(acc % num_leaves);
        }
        if (dedup) {
            std::sort(out.begin(), out.end());
            out.erase(std::unique(out.begin(), out.end()), out.end());
        }
        return out;
    }

    // geometric_challenge (protocols/geometric_challenge.hpp): {} / {1} /
    // {1, x, x^2, ...} with one Ext3 challenge x
    std::vector<gl::e3> geometric(size_t count) {
        std::vector<gl::e3> out;
        if (count == 0) return out;
        if (count == 1) return {gl::e3{1, 0, 0}};
        gl::e3 x = challenge_e3(), cur{1, 0, 0};
        for (size_t i = 0; i < count; ++i) {
            out.push_back(cur);
            cur = gl::e3_mul(cur, x);
        }
        return out;
    }

    static uint64_t bytes_mod_p(const uint8_t* b, size_t n) {
        // Horner in base 256 from the most significant byte, one reduction per
        // step: acc < p, so acc * 256 + byte < 2^72 splits as hi * 2^64 + lo.
        uint64_t acc = 0;
        for (size_t i = n; i-- > 0;) {
            uint64_t hi = acc >> 56, lo = (acc << 8) | b[i];
            acc = gl::reduce128(lo, hi);
        }
        return acc;
    }

    const std::vector<uint8_t>& narg() const { return narg_; }
    const std::vector<uint8_t>& hints() const { return hints_; }

private:
    template <typename T>
    static void append_le(std::vector<uint8_t>& o, T v) {
        for (size_t i = 0; i < sizeof(T); ++i) o.push_back((uint8_t)(v >> (8 * i)));
    }

    Sponge absorbed_;
    Sponge reader_;
    bool has_reader_ = false;
    std::vector<uint8_t> narg_, hints_;
};

}  // namespace whir_host
