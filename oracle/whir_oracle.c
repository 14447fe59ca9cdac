/*
 * whir_oracle.c — CPU restatement of the WHIR commit-and-fold path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked into the product.
 * Citations are relative to the reference repository's cpp/include/whir/.
 * Written for clarity, not speed: scalar loops, no threads, no SIMD.
 */
#include "whir_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ field */

uint64_t or_gl_from_u64(uint64_t v) { return v >= OR_P ? v - OR_P : v; }

uint64_t or_gl_add(uint64_t a, uint64_t b) {
    /* goldilocks.hpp:95-101: exact sum reduced once (a, b < p) */
    u128 s = (u128)a + b;
    if (s >= OR_P) s -= OR_P;
    return (uint64_t)s;
}

uint64_t or_gl_sub(uint64_t a, uint64_t b) {
    /* goldilocks.hpp:108-113 */
    return a >= b ? a - b : (uint64_t)((u128)a + OR_P - b);
}

uint64_t or_gl_neg(uint64_t a) { return a == 0 ? 0 : OR_P - a; }

uint64_t or_gl_mul(uint64_t a, uint64_t b) {
    /* goldilocks.hpp:116-124: the full 128-bit product reduced mod p, here by
     * splitting x = x3*2^96 + x2*2^64 + lo and using 2^64 = 2^32 - 1 and
     * 2^96 = -1, with exact 128-bit intermediate sums (no lazy wrap-around):
     * x = lo + x2*(2^32 - 1) - x3 (mod p). */
    u128 x = (u128)a * b;
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    u128 t = (u128)lo + (u128)(hi & 0xFFFFFFFFULL) * 0xFFFFFFFFULL + OR_P; /* < 2^66 */
    t -= (hi >> 32);
    while (t >= OR_P) t -= OR_P; /* t < 4p: at most three subtractions */
    return (uint64_t)t;
}

uint64_t or_gl_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1, x = a;
    while (e) {
        if (e & 1) r = or_gl_mul(r, x);
        x = or_gl_mul(x, x);
        e >>= 1;
    }
    return r;
}

uint64_t or_gl_inv(uint64_t a) { return or_gl_pow(a, OR_P - 2); }

uint64_t or_gl_root(uint64_t n) {
    /* goldilocks.hpp:155-159 gives w32 = 7^((2^0)*eps) for k = 32; the
     * engine (cooley_tukey.hpp:166-177) returns w32^(2^32/n). */
    uint64_t w32 = or_gl_pow(7, 0xFFFFFFFFULL);
    uint64_t e = (1ULL << 32) / n;
    return or_gl_pow(w32, e);
}

/* ------------------------------------------------------------------- ext3 */

void or_e3_add(const uint64_t* a, const uint64_t* b, uint64_t* o) {
    for (int i = 0; i < 3; ++i) o[i] = or_gl_add(a[i], b[i]);
}

void or_e3_sub(const uint64_t* a, const uint64_t* b, uint64_t* o) {
    for (int i = 0; i < 3; ++i) o[i] = or_gl_sub(a[i], b[i]);
}

void or_e3_mul(const uint64_t* a, const uint64_t* b, uint64_t* o) {
    /* goldilocks_ext3.hpp:83-89, X^3 = 2:
     * c0 = a0b0 + 2(a1b2 + a2b1), c1 = a0b1 + a1b0 + 2 a2b2, c2 = a0b2 + a1b1 + a2b0 */
    uint64_t t0 = or_gl_add(or_gl_mul(a[1], b[2]), or_gl_mul(a[2], b[1]));
    uint64_t c0 = or_gl_add(or_gl_mul(a[0], b[0]), or_gl_mul(2, t0));
    uint64_t c1 = or_gl_add(or_gl_add(or_gl_mul(a[0], b[1]), or_gl_mul(a[1], b[0])),
                            or_gl_mul(2, or_gl_mul(a[2], b[2])));
    uint64_t c2 = or_gl_add(or_gl_add(or_gl_mul(a[0], b[2]), or_gl_mul(a[1], b[1])),
                            or_gl_mul(a[2], b[0]));
    o[0] = c0; o[1] = c1; o[2] = c2;
}

void or_e3_pow(const uint64_t* a, uint64_t e, uint64_t* o) {
    uint64_t r[3] = {1, 0, 0}, x[3] = {a[0], a[1], a[2]};
    while (e) {
        if (e & 1) or_e3_mul(r, x, r);
        or_e3_mul(x, x, x);
        e >>= 1;
    }
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}

void or_e3_inv(const uint64_t* a, uint64_t* o) {
    /* goldilocks_ext3.hpp:111-119: adjugate / norm */
    uint64_t A = or_gl_sub(or_gl_mul(a[0], a[0]), or_gl_mul(2, or_gl_mul(a[1], a[2])));
    uint64_t B = or_gl_sub(or_gl_mul(2, or_gl_mul(a[2], a[2])), or_gl_mul(a[0], a[1]));
    uint64_t C = or_gl_sub(or_gl_mul(a[1], a[1]), or_gl_mul(a[0], a[2]));
    uint64_t N = or_gl_add(or_gl_mul(a[0], A),
                           or_gl_mul(2, or_gl_add(or_gl_mul(a[2], B), or_gl_mul(a[1], C))));
    uint64_t Ni = or_gl_inv(N);
    o[0] = or_gl_mul(A, Ni); o[1] = or_gl_mul(B, Ni); o[2] = or_gl_mul(C, Ni);
}

/* -------------------------------------------------------------------- NTT */

static void ntt_block(uint64_t* v, size_t n, const uint64_t* tw /* w_n^k, k < n/2 */) {
    /* cooley_tukey.hpp:840-872: bit reversal, then radix-2 DIT passes */
    if (n <= 1) return;
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { uint64_t t = v[i]; v[i] = v[j]; v[j] = t; }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1, step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t j = 0; j < half; ++j) {
                uint64_t u = v[i + j];
                uint64_t t = or_gl_mul(v[i + j + half], tw[j * step]);
                v[i + j] = or_gl_add(u, t);
                v[i + j + half] = or_gl_sub(u, t);
            }
        }
    }
}

static uint64_t* twiddles(size_t n) {
    size_t h = n / 2 ? n / 2 : 1;
    uint64_t* tw = (uint64_t*)malloc(h * sizeof(uint64_t));
    uint64_t w = or_gl_root(n), acc = 1;
    for (size_t k = 0; k < h; ++k) { tw[k] = acc; acc = or_gl_mul(acc, w); }
    return tw;
}

void or_ntt_batch(uint64_t* values, size_t total, size_t n) {
    if (n <= 1) return;
    uint64_t* tw = twiddles(n);
    for (size_t off = 0; off + n <= total; off += n) ntt_block(values + off, n, tw);
    free(tw);
}

void or_intt_batch(uint64_t* values, size_t total, size_t n) {
    for (size_t off = 0; off + n <= total; off += n) {
        for (size_t i = 1, j = n - 1; i < j; ++i, --j) {
            uint64_t t = values[off + i]; values[off + i] = values[off + j]; values[off + j] = t;
        }
    }
    or_ntt_batch(values, total, n);
}

void or_dft_naive(const uint64_t* in, uint64_t* out, size_t n) {
    uint64_t w = n > 1 ? or_gl_root(n) : 1;
    for (size_t k = 0; k < n; ++k) {
        uint64_t wk = or_gl_pow(w, k), x = 1, acc = 0;
        for (size_t j = 0; j < n; ++j) {
            acc = or_gl_add(acc, or_gl_mul(in[j], x));
            x = or_gl_mul(x, wk);
        }
        out[k] = acc;
    }
}

/* -------------------------------------------------------------- RS encode */

void or_rs_encode(const uint64_t* coeffs, size_t num_polys, size_t poly_size,
                  size_t cw, size_t depth, uint64_t* out) {
    /* mod_ntt.hpp:49-130: zero-pad each block of m = poly_size/depth to cw,
     * NTT of size cw, transpose to row-major cw x (num_polys*depth). */
    size_t m = poly_size / depth, cols = num_polys * depth;
    uint64_t* buf = (uint64_t*)calloc(cw ? cw : 1, sizeof(uint64_t));
    uint64_t* tw = twiddles(cw > 1 ? cw : 2);
    for (size_t k = 0; k < num_polys; ++k) {
        for (size_t b = 0; b < depth; ++b) {
            memset(buf, 0, cw * sizeof(uint64_t));
            memcpy(buf, coeffs + k * poly_size + b * m, m * sizeof(uint64_t));
            ntt_block(buf, cw, tw);
            for (size_t i = 0; i < cw; ++i) out[i * cols + k * depth + b] = buf[i];
        }
    }
    free(tw);
    free(buf);
}

void or_rs_encode_ext3(const uint64_t* coeffs, size_t num_polys, size_t poly_size,
                       size_t cw, size_t depth, uint64_t* out) {
    /* the Ext3 engine embeds the base roots (cooley_tukey_goldilocks.hpp:41-46),
     * so each component transforms independently */
    size_t n_in = num_polys * poly_size, cols = num_polys * depth;
    uint64_t* comp = (uint64_t*)malloc((n_in ? n_in : 1) * sizeof(uint64_t));
    uint64_t* enc = (uint64_t*)malloc((cw * cols > 0 ? cw * cols : 1) * sizeof(uint64_t));
    for (int c = 0; c < 3; ++c) {
        for (size_t i = 0; i < n_in; ++i) comp[i] = coeffs[3 * i + c];
        or_rs_encode(comp, num_polys, poly_size, cw, depth, enc);
        for (size_t i = 0; i < cw * cols; ++i) out[3 * i + c] = enc[i];
    }
    free(comp);
    free(enc);
}

/* ----------------------------------------------------------------- BLAKE3 */
/* Standard BLAKE3 from its specification (hash mode, 32-byte output). */

static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

static uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void b3_g(uint32_t* s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx; s[d] = rotr32(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];      s[b] = rotr32(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my; s[d] = rotr32(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];      s[b] = rotr32(s[b] ^ s[c], 7);
}

static void b3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[8]) {
    uint32_t m[16], s[16], t[16];
    for (int i = 0; i < 16; ++i)
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) |
               ((uint32_t)block[4 * i + 2] << 16) | ((uint32_t)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; ++i) s[i] = cv[i];
    for (int i = 0; i < 4; ++i) s[8 + i] = B3_IV[i];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len; s[15] = flags;
    for (int r = 0; r < 7; ++r) {
        b3_g(s, 0, 4, 8, 12, m[0], m[1]);   b3_g(s, 1, 5, 9, 13, m[2], m[3]);
        b3_g(s, 2, 6, 10, 14, m[4], m[5]);  b3_g(s, 3, 7, 11, 15, m[6], m[7]);
        b3_g(s, 0, 5, 10, 15, m[8], m[9]);  b3_g(s, 1, 6, 11, 12, m[10], m[11]);
        b3_g(s, 2, 7, 8, 13, m[12], m[13]); b3_g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            for (int i = 0; i < 16; ++i) t[i] = m[B3_PERM[i]];
            memcpy(m, t, sizeof m);
        }
    }
    for (int i = 0; i < 8; ++i) out[i] = s[i] ^ s[i + 8];
}

/* chaining value (or root output when is_root) of one chunk of <= 1024 bytes */
static void b3_chunk(const uint8_t* in, size_t len, uint64_t chunk_idx, int is_root, uint32_t out[8]) {
    uint32_t cv[8];
    memcpy(cv, B3_IV, sizeof cv);
    size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (size_t b = 0; b < nblocks; ++b) {
        uint8_t block[64] = {0};
        size_t bl = len - b * 64 < 64 ? len - b * 64 : 64;
        if (len == 0) bl = 0;
        memcpy(block, in + b * 64, bl);
        uint32_t flags = 0;
        if (b == 0) flags |= B3_CHUNK_START;
        if (b + 1 == nblocks) { flags |= B3_CHUNK_END; if (is_root) flags |= B3_ROOT; }
        b3_compress(cv, block, chunk_idx, (uint32_t)bl, flags, cv);
    }
    memcpy(out, cv, sizeof cv);
}

static void b3_parent(const uint32_t l[8], const uint32_t r[8], int is_root, uint32_t out[8]) {
    uint8_t block[64];
    for (int i = 0; i < 8; ++i) {
        for (int k = 0; k < 4; ++k) {
            block[4 * i + k] = (uint8_t)(l[i] >> (8 * k));
            block[32 + 4 * i + k] = (uint8_t)(r[i] >> (8 * k));
        }
    }
    b3_compress(B3_IV, block, 0, 64, B3_PARENT | (is_root ? B3_ROOT : 0), out);
}

/* subtree over chunks [first, first+n): left part = largest power of two < n */
static void b3_subtree(const uint8_t* in, size_t len, uint64_t first, uint64_t n, int is_root, uint32_t out[8]) {
    if (n == 1) {
        size_t off = (size_t)first * 1024;
        size_t cl = len - off < 1024 ? len - off : 1024;
        b3_chunk(in + off, cl, first, is_root, out);
        return;
    }
    uint64_t left = 1;
    while (left * 2 < n) left *= 2;
    uint32_t l[8], r[8];
    b3_subtree(in, len, first, left, 0, l);
    b3_subtree(in, len, first + left, n - left, 0, r);
    b3_parent(l, r, is_root, out);
}

void or_blake3(const uint8_t* in, size_t len, uint8_t out[32]) {
    uint64_t nchunks = len == 0 ? 1 : (len + 1023) / 1024;
    uint32_t h[8];
    b3_subtree(in, len, 0, nchunks, 1, h);
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 4; ++k) out[4 * i + k] = (uint8_t)(h[i] >> (8 * k));
}

void or_blake3_many(size_t size, const uint8_t* in, size_t count, uint8_t* out) {
    for (size_t i = 0; i < count; ++i) or_blake3(in + i * size, size, out + 32 * i);
}

/* ---------------------------------------------------------- matrix commit */

static void put_le64(uint8_t* p, uint64_t v) {
    for (int b = 0; b < 8; ++b) p[b] = (uint8_t)(v >> (8 * b));
}

void or_commit_leaves_base(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves) {
    /* matrix_commit.hpp:169-195 with encode_into<Goldilocks> :77-101 */
    uint8_t* buf = (uint8_t*)malloc(cols * 8 + 1);
    for (size_t i = 0; i < rows; ++i) {
        for (size_t c = 0; c < cols; ++c) put_le64(buf + 8 * c, matrix[i * cols + c]);
        or_blake3(buf, cols * 8, leaves + 32 * i);
    }
    free(buf);
}

void or_commit_leaves_ext3(const uint64_t* matrix, size_t rows, size_t cols, uint8_t* leaves) {
    /* encode_into<GoldilocksExt3> :120-136: c0, c1, c2 LE per element */
    uint8_t* buf = (uint8_t*)malloc(cols * 24 + 1);
    for (size_t i = 0; i < rows; ++i) {
        for (size_t c = 0; c < 3 * cols; ++c) put_le64(buf + 8 * c, matrix[i * 3 * cols + c]);
        or_blake3(buf, cols * 24, leaves + 32 * i);
    }
    free(buf);
}

/* ------------------------------------------------------------ Merkle tree */

size_t or_merkle_layers(size_t n) {
    size_t pow = 1, k = 0;
    if (n <= 1) return 0;
    while (pow < n) { pow <<= 1; ++k; }
    return k;
}

size_t or_merkle_num_nodes(size_t n) { return ((size_t)1 << (or_merkle_layers(n) + 1)) - 1; }

void or_merkle_build(const uint8_t* leaves, size_t n, uint8_t* nodes) {
    size_t L = or_merkle_layers(n), width = (size_t)1 << L;
    memset(nodes, 0, or_merkle_num_nodes(n) * 32);
    memcpy(nodes, leaves, n * 32);
    size_t prev = 0, cur = width;
    while (width > 1) {
        for (size_t i = 0; i < width / 2; ++i) or_blake3(nodes + 32 * (prev + 2 * i), 64, nodes + 32 * (cur + i));
        prev = cur;
        cur += width / 2;
        width /= 2;
    }
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

size_t or_merkle_open(const uint8_t* nodes, size_t n_leaves, const uint64_t* idx, size_t n_idx, uint8_t* hints) {
    /* merkle_tree.hpp:165-211 */
    uint64_t* cur = (uint64_t*)malloc((n_idx + 1) * sizeof(uint64_t));
    memcpy(cur, idx, n_idx * sizeof(uint64_t));
    qsort(cur, n_idx, sizeof(uint64_t), cmp_u64);
    size_t m = 0;
    for (size_t i = 0; i < n_idx; ++i) if (m == 0 || cur[m - 1] != cur[i]) cur[m++] = cur[i];
    size_t nh = 0, off = 0, len = (size_t)1 << or_merkle_layers(n_leaves);
    while (len > 1) {
        size_t nm = 0;
        for (size_t k = 0; k < m;) {
            uint64_t a = cur[k];
            if (k + 1 < m && cur[k + 1] == (a ^ 1)) { cur[nm++] = a >> 1; k += 2; }
            else {
                if (hints) memcpy(hints + 32 * nh, nodes + 32 * (off + (a ^ 1)), 32);
                ++nh;
                cur[nm++] = a >> 1;
                k += 1;
            }
        }
        m = nm;
        off += len;
        len /= 2;
    }
    free(cur);
    return nh;
}

int or_merkle_verify(const uint8_t root[32], size_t n_leaves, const uint64_t* idx,
                     const uint8_t* leaf_hashes, size_t n_idx, const uint8_t* hints, size_t n_hints) {
    /* merkle_tree.hpp:220-316, restated: sort (index, hash) pairs, reject
     * inconsistent duplicates, then rebuild layer by layer */
    if (n_idx == 0) return 1;
    for (size_t i = 0; i < n_idx; ++i) if (idx[i] >= n_leaves) return 0;
    uint64_t* ix = (uint64_t*)malloc(n_idx * sizeof(uint64_t));
    uint8_t* hs = (uint8_t*)malloc(n_idx * 32);
    uint8_t* pair = (uint8_t*)malloc(64);
    size_t* ord = (size_t*)malloc(n_idx * sizeof(size_t));
    for (size_t i = 0; i < n_idx; ++i) ord[i] = i;
    for (size_t i = 1; i < n_idx; ++i) {  /* stable insertion sort by index */
        size_t t = ord[i], j = i;
        while (j > 0 && idx[ord[j - 1]] > idx[t]) { ord[j] = ord[j - 1]; --j; }
        ord[j] = t;
    }
    size_t m = 0;
    int ok = 1;
    for (size_t i = 0; i < n_idx; ++i) {
        const uint8_t* h = leaf_hashes + 32 * ord[i];
        if (m > 0 && ix[m - 1] == idx[ord[i]]) {
            if (memcmp(hs + 32 * (m - 1), h, 32) != 0) ok = 0;
            continue;
        }
        ix[m] = idx[ord[i]];
        memcpy(hs + 32 * m, h, 32);
        ++m;
    }
    size_t hc = 0, layers = or_merkle_layers(n_leaves);
    for (size_t l = 0; ok && l < layers; ++l) {
        size_t nm = 0;
        for (size_t k = 0; k < m;) {
            uint64_t a = ix[k];
            if (k + 1 < m && ix[k + 1] == (a ^ 1)) {
                memcpy(pair, hs + 32 * k, 32);
                memcpy(pair + 32, hs + 32 * (k + 1), 32);
                k += 2;
            } else {
                if (hc >= n_hints) { ok = 0; break; }
                const uint8_t* h = hints + 32 * hc++;
                if ((a & 1) == 0) { memcpy(pair, hs + 32 * k, 32); memcpy(pair + 32, h, 32); }
                else { memcpy(pair, h, 32); memcpy(pair + 32, hs + 32 * k, 32); }
                k += 1;
            }
            or_blake3(pair, 64, hs + 32 * nm);
            ix[nm++] = a >> 1;
        }
        m = nm;
    }
    if (ok) ok = (m == 1 && hc == n_hints && memcmp(hs, root, 32) == 0);
    free(ix); free(hs); free(pair); free(ord);
    return ok;
}

/* --------------------------------------------------------------- sumcheck */

void or_sumcheck_compute_ext3(const uint64_t* a, const uint64_t* b, size_t n, uint64_t c0[3], uint64_t c2[3]) {
    /* sumcheck.hpp:43-98: c0 = sum a0*b0, c2 = sum (a1-a0)(b1-b0) over the half split */
    size_t h = n / 2;
    uint64_t s0[3] = {0, 0, 0}, s2[3] = {0, 0, 0}, t[3], da[3], db[3];
    for (size_t i = 0; i < h; ++i) {
        or_e3_mul(a + 3 * i, b + 3 * i, t);
        or_e3_add(s0, t, s0);
        or_e3_sub(a + 3 * (h + i), a + 3 * i, da);
        or_e3_sub(b + 3 * (h + i), b + 3 * i, db);
        or_e3_mul(da, db, t);
        or_e3_add(s2, t, s2);
    }
    memcpy(c0, s0, sizeof s0);
    memcpy(c2, s2, sizeof s2);
}

void or_sumcheck_compute_base_ext3(const uint64_t* a, const uint64_t* b, size_t n, uint64_t c0[3], uint64_t c2[3]) {
    uint64_t* la = (uint64_t*)calloc(3 * (n ? n : 1), sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) la[3 * i] = a[i];
    or_sumcheck_compute_ext3(la, b, n, c0, c2);
    free(la);
}

void or_sumcheck_fold_ext3(uint64_t* v, size_t n, const uint64_t r[3]) {
    /* sumcheck.hpp:103-116 */
    size_t h = n / 2;
    uint64_t d[3], t[3];
    for (size_t i = 0; i < h; ++i) {
        or_e3_sub(v + 3 * (h + i), v + 3 * i, d);
        or_e3_mul(d, r, t);
        or_e3_add(v + 3 * i, t, v + 3 * i);
    }
}

void or_sumcheck_fold_base_ext3(const uint64_t* v, size_t n, const uint64_t r[3], uint64_t* out) {
    uint64_t* l = (uint64_t*)calloc(3 * (n ? n : 1), sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) l[3 * i] = v[i];
    or_sumcheck_fold_ext3(l, n, r);
    memcpy(out, l, 3 * (n / 2) * sizeof(uint64_t));
    free(l);
}

/* ----------------------------------------------------------------- Horner */

void or_horner_base_ext3(const uint64_t* c, size_t n, const uint64_t z[3], uint64_t out[3]) {
    /* utilities.hpp:195-246, serial branch :238-245 */
    uint64_t acc[3] = {0, 0, 0};
    if (n) { acc[0] = c[n - 1]; }
    for (size_t i = n ? n - 1 : 0; i > 0; --i) {
        or_e3_mul(acc, z, acc);
        acc[0] = or_gl_add(acc[0], c[i - 1]);
    }
    memcpy(out, acc, sizeof acc);
}

void or_horner_ext3_ext3(const uint64_t* c, size_t n, const uint64_t z[3], uint64_t out[3]) {
    uint64_t acc[3] = {0, 0, 0};
    if (n) memcpy(acc, c + 3 * (n - 1), sizeof acc);
    for (size_t i = n ? n - 1 : 0; i > 0; --i) {
        or_e3_mul(acc, z, acc);
        or_e3_add(acc, c + 3 * (i - 1), acc);
    }
    memcpy(out, acc, sizeof acc);
}

uint64_t or_horner_base(const uint64_t* c, size_t n, uint64_t z) {
    uint64_t acc = 0;
    for (size_t i = n; i > 0; --i) acc = or_gl_add(or_gl_mul(acc, z), c[i - 1]);
    return acc;
}

/* ------------------------------------------------------ prove-path algebra */

void or_eval_eq_ext3(uint64_t* acc, const uint64_t* point, size_t k, const uint64_t scalar[3]) {
    /* multilinear.hpp:120-133 */
    if (k == 0) {
        or_e3_add(acc, scalar, acc);
        return;
    }
    uint64_t s1[3], s0[3];
    or_e3_mul(scalar, point, s1);
    or_e3_sub(scalar, s1, s0);
    size_t half = (size_t)1 << (k - 1);
    or_eval_eq_ext3(acc, point + 3, k - 1, s0);
    or_eval_eq_ext3(acc + 3 * half, point + 3, k - 1, s1);
}

void or_geometric_accumulate_ext3(uint64_t* acc, size_t n, const uint64_t* points, const uint64_t* scalars,
                                  size_t count) {
    /* utilities.hpp:300-306: entry += s_j, then s_j *= p_j, for each entry in order */
    uint64_t* s = (uint64_t*)malloc(3 * (count ? count : 1) * sizeof(uint64_t));
    memcpy(s, scalars, 3 * count * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < count; ++j) {
            or_e3_add(acc + 3 * i, s + 3 * j, acc + 3 * i);
            or_e3_mul(s + 3 * j, points + 3 * j, s + 3 * j);
        }
    free(s);
}

void or_mle_evaluate(const uint64_t* evals, int is_base, const uint64_t* point, size_t k, uint64_t out[3]) {
    /* multilinear.hpp:93-101: f0 + (f1 - f0) * point[0] over the two halves */
    if (k == 0) {
        if (is_base) {
            out[0] = evals[0];
            out[1] = out[2] = 0;
        } else {
            memcpy(out, evals, 3 * sizeof(uint64_t));
        }
        return;
    }
    size_t half = (size_t)1 << (k - 1);
    uint64_t f0[3], f1[3], d[3];
    or_mle_evaluate(evals, is_base, point + 3, k - 1, f0);
    or_mle_evaluate(evals + (is_base ? 1 : 3) * half, is_base, point + 3, k - 1, f1);
    or_e3_sub(f1, f0, d);
    or_e3_mul(d, point, d);
    or_e3_add(f0, d, out);
}

uint64_t or_pow_find(const uint8_t ch[32], uint64_t threshold) {
    /* proof_of_work.hpp:95-193 */
    if (threshold == UINT64_MAX) return 0;
    uint8_t in[64], h[32];
    memset(in, 0, sizeof in);
    memcpy(in, ch, 32);
    for (uint64_t n = 0;; ++n) {
        for (int b = 0; b < 8; ++b) in[32 + b] = (uint8_t)(n >> (8 * b));
        or_blake3(in, 64, h);
        uint64_t v = 0;
        for (int b = 7; b >= 0; --b) v = (v << 8) | h[b];
        if (v <= threshold) return n;
    }
}

/* ------------------------------------------------------------- seeded LCG */

void or_lcg_fill(uint64_t* state, uint64_t* out, size_t n, int mode) {
    uint64_t s = *state;
    for (size_t i = 0; i < n;) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        if (mode == 1) {
            if (s >= OR_P) continue;
            out[i++] = s;
        } else {
            out[i++] = or_gl_from_u64(s);
        }
    }
    *state = s;
}
