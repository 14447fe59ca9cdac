/*
 * keccak_oracle.c — FIPS-202 Keccak-f[1600], SHA3-256/512 and SHAKE-128.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Restates the standard that the
 * reference's vendored third_party/keccak implements; the transcript
 * (transcript/transcript.hpp:80-139) is a SHAKE-128 duplex sponge over it.
 * Round constants and rotation offsets are generated from the FIPS-202
 * definitions (LFSR rc(t) and the (t+1)(t+2)/2 schedule), not tabulated.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

static uint64_t KRC[24];
static int KROT[25];
static int k_init = 0;

static int rc_bit(int t) {
    /* FIPS-202 Algorithm 5 */
    if (t % 255 == 0) return 1;
    uint32_t R = 0x80; /* bits R[0..7], R[0] is the MSB here */
    for (int i = 1; i <= t % 255; ++i) {
        uint32_t r8 = R & 1;       /* R[8] after prepending 0 is old R[7] */
        R >>= 1;                   /* prepend 0 */
        if (r8) R ^= 0x80 | 0x08 | 0x04 | 0x02; /* R[0],R[4],R[5],R[6] ^= R[8] */
    }
    return (R >> 7) & 1;
}

static void keccak_init_tables(void) {
    for (int ir = 0; ir < 24; ++ir) {
        uint64_t rc = 0;
        for (int j = 0; j <= 6; ++j)
            if (rc_bit(j + 7 * ir)) rc |= 1ULL << ((1 << j) - 1);
        KRC[ir] = rc;
    }
    KROT[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; ++t) {
        KROT[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx; y = ny;
    }
    k_init = 1;
}

static uint64_t rotl64(uint64_t v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

void or_keccak_f1600(uint64_t A[25]) {
    if (!k_init) keccak_init_tables();
    for (int ir = 0; ir < 24; ++ir) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; ++x) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; ++x) D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; ++i) A[i] ^= D[i % 5];
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y)
                B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[x + 5 * y], KROT[x + 5 * y]);
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y)
                A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= KRC[ir];
    }
}

/* generic sponge: rate in bytes, domain-separation suffix byte (0x06 SHA3, 0x1F SHAKE) */
static void sponge(size_t rate, uint8_t suffix, const uint8_t* in, size_t len, uint8_t* out, size_t olen) {
    uint64_t A[25];
    uint8_t* S = (uint8_t*)A;
    memset(A, 0, sizeof A);
    while (len >= rate) {
        for (size_t i = 0; i < rate; ++i) S[i] ^= in[i];
        or_keccak_f1600(A);
        in += rate; len -= rate;
    }
    for (size_t i = 0; i < len; ++i) S[i] ^= in[i];
    S[len] ^= suffix;
    S[rate - 1] ^= 0x80;
    or_keccak_f1600(A);
    while (olen > 0) {
        size_t n = olen < rate ? olen : rate;
        memcpy(out, S, n);
        out += n; olen -= n;
        if (olen) or_keccak_f1600(A);
    }
}

void or_sha3_256(const uint8_t* in, size_t len, uint8_t out[32]) { sponge(136, 0x06, in, len, out, 32); }
void or_sha3_512(const uint8_t* in, size_t len, uint8_t out[64]) { sponge(72, 0x06, in, len, out, 64); }
void or_shake128(const uint8_t* in, size_t len, uint8_t* out, size_t olen) { sponge(168, 0x1F, in, len, out, olen); }
