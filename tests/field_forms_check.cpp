// Host build of the device field forms (whirplonk_amd/csrc/gl_forms.h) checked
// against unsigned __int128 arithmetic.  The reduction's borrow branches need
// R = lo + hl*eps (mod 2^64) below hh < 2^32, which random inputs reach with
// probability about 2^-32, so the inputs are built per branch: each of the four
// (carry, borrow) cases is hit many times and the run fails if one is not.
// Prints "ok <cases> <c0b0> <c0b1> <c1b0> <c1b1>" and exits 0 on success.
#include <stdint.h>
#include <stdio.h>

#define GL_P 0xFFFFFFFF00000001ull
#define GL_EPS 0xFFFFFFFFull
#define GL_FORM static inline
namespace gl {
#include "gl_forms.h"
}

typedef unsigned __int128 u128;
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {  // splitmix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static long cases = 0, bad = 0, hit[4] = {0, 0, 0, 0};

static void check_reduce(uint64_t lo, uint64_t hi) {
    const u128 x = ((u128)hi << 64) | lo;
    const uint64_t want = (uint64_t)(x % GL_P), got = gl::reduce128(lo, hi);
    const uint32_t hl = (uint32_t)hi, hh = (uint32_t)(hi >> 32);
    const u128 r = (u128)lo + (u128)hl * GL_EPS;
    const int c = r >> 64 ? 1 : 0, b = (uint64_t)r < hh ? 1 : 0;
    hit[2 * c + b]++;
    ++cases;
    if (got != want && bad++ < 5)
        printf("reduce128(%016llx, %016llx) = %016llx, want %016llx\n", (unsigned long long)lo, (unsigned long long)hi,
               (unsigned long long)got, (unsigned long long)want);
}

static void check_addsub(uint64_t a, uint64_t b) {
    const uint64_t wa = (uint64_t)(((u128)a + b) % GL_P), ws = (uint64_t)(((u128)a + GL_P - b) % GL_P);
    const uint64_t wn = a ? GL_P - a : 0;
    ++cases;
    if ((gl::add(a, b) != wa || gl::sub(a, b) != ws || gl::neg(a) != wn) && bad++ < 5)
        printf("add/sub/neg(%016llx, %016llx) wrong\n", (unsigned long long)a, (unsigned long long)b);
}

int main() {
    const uint64_t edge[] = {0, 1, 2, GL_EPS - 1, GL_EPS, GL_EPS + 1, 1ull << 32, 1ull << 63, GL_P - 2, GL_P - 1,
                             GL_P, GL_P + 1, ~0ull - 1, ~0ull, 0xFFFFFFFF00000000ull, 0x00000001FFFFFFFFull,
                             0xFFFFFFFE00000000ull, 0xFFFFFFFE00000001ull, 0xFFFFFFFEFFFFFFFFull, 0x7FFFFFFF80000000ull};
    const int ne = sizeof(edge) / sizeof(edge[0]);
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j)
            for (int d = -2; d <= 2; ++d) check_reduce(edge[i] + d, edge[j] - d);
    for (long i = 0; i < 2000000; ++i) {
        uint64_t lo = rnd(), hi = rnd();
        const uint32_t hl = (uint32_t)hi, hh = (uint32_t)(hi >> 32) | 1u;
        switch (i & 7) {
            case 0: break;                                           // c = 1 or 0 at random, b = 0
            case 1:                                                  // c = 0, b = 1: R = lo + hl*eps < hh
                hi = ((uint64_t)hh << 32) | 0u;
                lo = rnd() % hh;
                break;
            case 2: {                                                // c = 1, b = 1: R wraps below hh
                const uint64_t prod = (uint64_t)hl * GL_EPS;         // lo = 2^64 - prod + r, r < hh
                if (prod == 0) break;
                lo = (0 - prod) + rnd() % hh;
                hi = ((uint64_t)hh << 32) | hl;
                break;
            }
            case 3: lo = ~0ull - (rnd() & 0xFFFF); hi |= 0xFFFFFFFFull; break;   // c = 1, R near its maximum
            case 4: hi &= 0xFFFFFFFFull; break;                                   // hh = 0
            case 5: hi &= 0xFFFFFFFF00000000ull; break;                           // hl = 0
            case 6: {                                                             // products of canonical values
                const uint64_t a = rnd() % GL_P, bb = (i & 8) ? GL_P - 1 - (rnd() & 0xFF) : rnd() % GL_P;
                const u128 x = (u128)a * bb;
                lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
                break;
            }
            case 7: lo = GL_P - 1 + (rnd() & 3); hi &= 0xFFFFFFFFull; break;      // W around p
        }
        check_reduce(lo, hi);
    }
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j) check_addsub(edge[i] % GL_P, edge[j] % GL_P);
    for (long i = 0; i < 1000000; ++i) {
        const uint64_t a = (i & 3) == 0 ? GL_P - 1 - (rnd() & 0xFFF) : rnd() % GL_P;
        const uint64_t b = (i & 3) == 1 ? GL_P - 1 - (rnd() & 0xFFF) : rnd() % GL_P;
        check_addsub(a, b);
    }
    const long need = 1000;
    if (bad || hit[0] < need || hit[1] < need || hit[2] < need || hit[3] < need) {
        printf("FAIL bad=%ld hits %ld %ld %ld %ld\n", bad, hit[0], hit[1], hit[2], hit[3]);
        return 1;
    }
    printf("ok %ld %ld %ld %ld %ld\n", cases, hit[0], hit[1], hit[2], hit[3]);
    return 0;
}
