/* CPU check of the DDA's division by reciprocal and one FMA correction
 * (q0 = a*y, q = fma(fma(-q0, d, a), y, q0) with y = 1/d) against IEEE a/d,
 * over random signed pairs with exponents in [-60, 60] and pairs whose
 * quotient lies next to a rounding midpoint.  Built by __graft_entry__.build(). */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float f_of(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static uint32_t u_of(float x) { uint32_t v; memcpy(&v, &x, 4); return v; }

static uint64_t next(uint64_t* s) {
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return *s;
}

static float fast_div(float a, float d) {
    float y = 1.0f / d;
    float q0 = a * y;
    return fmaf(fmaf(-q0, d, a), y, q0);
}

/* mode 0: random signed and scaled pairs; mode 1: quotients next to midpoints. */
long divcheck(long n, uint64_t seed, int mode) {
    uint64_t s = seed | 1;
    long bad = 0;
    for (long i = 0; i < n; i++) {
        uint64_t r = next(&s);
        float a, d;
        if (mode == 0) {
            a = f_of(0x3f800000u | (uint32_t)(r & 0x7fffff));
            d = f_of(0x3f800000u | (uint32_t)((r >> 23) & 0x7fffff));
            a = ldexpf((r >> 63) ? -a : a, (int)((r >> 46) % 121) - 60);
            d = ldexpf(((r >> 62) & 1) ? -d : d, -(int)((r >> 53) % 61));
        } else {
            d = f_of(0x3f800000u | (uint32_t)(r & 0x7fffff));
            float q = f_of(0x3f800000u | (uint32_t)((r >> 23) & 0x7fffff));
            double mid = (double)q + ldexp(1.0, -24);
            a = f_of(u_of((float)(mid * (double)d)) + (int)((r >> 50) % 3) - 1);
        }
        if (u_of(fast_div(a, d)) != u_of(a / d)) bad++;
    }
    return bad;
}
