// gl_forms.h — the device forms of Goldilocks add, sub, neg and the 128-bit
// reduction, included by gl.h inside namespace gl with GL_FORM defined as
// `__device__ __forceinline__`.  The file has no HIP dependency so that the
// CPU tests compile the same source with GL_FORM = `static inline` (clang,
// for __builtin_subc) and check it against 128-bit arithmetic on the inputs
// that reach every branch.  Needs GL_P, GL_EPS and <stdint.h>.
//
// The forms are shaped for gfx950's VALU: 64-bit adds (v_lshl_add_u64) and
// 64-bit compares where no carry has to chain, 32-bit borrow chains only where
// the borrow itself is the condition.  A carry or compare result read by the
// very next VALU instruction costs wait states, so each function keeps such
// flag chains to one or two.
GL_FORM uint32_t lo32(uint64_t x) { return (uint32_t)x; }
GL_FORM uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
GL_FORM uint64_t mk64(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

// a, b canonical.  t = a + b + eps (mod 2^64), and b + eps < 2^64.  If a + b
// carried, t = a + b - p; if a + b in [p, 2^64), t wraps to a + b - p; both
// are below a (b < p).  If a + b < p, t = a + b + eps >= a.  So t < a exactly
// when the answer is a + b - p = t, and otherwise the answer is a + b.
GL_FORM uint64_t add(uint64_t a, uint64_t b) {
    uint64_t t = a + (b + GL_EPS);
    return t < a ? t : a + b;
}

// a, b canonical: d = a - b; on borrow the answer is d + p (mod 2^64).
GL_FORM uint64_t sub(uint64_t a, uint64_t b) {
    uint32_t b1, b2;
    const uint32_t d0 = __builtin_subc(lo32(a), lo32(b), 0u, &b1);
    const uint32_t d1 = __builtin_subc(hi32(a), hi32(b), b1, &b2);
    const uint64_t d = mk64(d0, d1);
    return b2 ? d + GL_P : d;
}

GL_FORM uint64_t neg(uint64_t a) { return sub(0, a); }

// x = hi*2^64 + lo (any 128-bit value) -> x mod p, canonical.  With
// hi = hh*2^32 + hl, x = lo + hl*eps - hh (mod p).
//   R = lo + hl*eps (mod 2^64), carry c weighing 2^64 = eps; c = 1 means
//       R <= (2^32-1)^2 - 1 = 2^64 - 2^33.
//   W = R - hh (mod 2^64), borrow b weighing -2^64 = -eps.
//   x = W + (c - b)*eps (mod p), and per case:
//     c=1 b=0: W + eps <= 2^64 - 2^32 - 1 < p       -> W + eps
//     c=1 b=1: W in [p, 2^64), x = W                -> W - p = W + eps (mod 2^64)
//     c=0 b=1: W in [p, 2^64), x = W - eps < p      -> W - eps = W + p (mod 2^64)
//     c=0 b=0: x = W                                -> W >= p ? W + eps (mod 2^64) : W
// Every case with W >= p or c = 1 adds k = (b && !c) ? p : eps; the rest keep W.
GL_FORM uint64_t reduce128(uint64_t lo, uint64_t hi) {
    const uint32_t hl = lo32(hi), hh = hi32(hi);
    const uint64_t R = (uint64_t)hl * GL_EPS + lo;
    const bool c = R < lo;
    uint32_t b1, b;
    const uint32_t w0 = __builtin_subc(lo32(R), hh, 0u, &b1);
    const uint32_t w1 = __builtin_subc(hi32(R), 0u, b1, &b);
    const uint64_t W = mk64(w0, w1);
    const uint64_t k = (b && !c) ? GL_P : GL_EPS;
    return (c || W >= GL_P) ? W + k : W;
}
