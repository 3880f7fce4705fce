/*
 * Test infrastructure: the engine's math functions (pumas_amd/csrc/libm.h)
 * compiled for the host, compared bit for bit with the host C library.
 * Returns the number of mismatching arguments; NaN results compare equal.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "libm.h"

using namespace pg;

static inline uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline uint32_t fbits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline bool same(double a, double b) { return bits(a) == bits(b) || (a != a && b != b); }
static inline bool fsame(float a, float b) { return fbits(a) == fbits(b) || (a != a && b != b); }

/* Every float bit pattern i with i % stride == offset. */
extern "C" long check_float(int which, long stride, long offset)
{
        long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
        for (long i = offset; i < (1L << 32); i += stride) {
                float x; uint32_t u = (uint32_t)i; memcpy(&x, &u, 4);
                const float r = which == 0 ? lm::logf(x) : lm::expf(x);
                const float g = which == 0 ? ::logf(x) : ::expf(x);
                if (!fsame(r, g)) bad++;
        }
        return bad;
}

static inline uint64_t mix(uint64_t z)
{
        z += 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
}
static inline double unit(uint64_t h) { return ((h >> 11) + 0.5) * 0x1p-53; }

/* n random arguments per function over its whole range and the ranges the
 * transport uses. which: 0 exp, 1 log, 2 pow, 3 sin, 4 cos, 5 sincos,
 * 6 log10. */
extern "C" long check_double(int which, long n, uint64_t seed)
{
        long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
        for (long i = 0; i < n; i++) {
                const uint64_t h1 = mix(seed ^ (uint64_t)i * 2654435761ull);
                const uint64_t h2 = mix(h1);
                const int kind = (int)(h2 & 3);
                double x, y = 0;
                switch (which) {
                case 0:
                        x = kind == 0 ? (unit(h1) * 2 - 1) * 745 :
                            kind == 1 ? (unit(h1) * 2 - 1) * 20 :
                            kind == 2 ? (unit(h1) * 2 - 1) : [&] { double v; uint64_t b = h1; memcpy(&v, &b, 8); return v; }();
                        if (!same(lm::exp(x), ::exp(x))) bad++;
                        break;
                case 1:
                case 6: {
                        uint64_t b = h1 >> 1;
                        x = kind == 0 ? [&] { double v; memcpy(&v, &b, 8); return v; }() :
                            kind == 1 ? unit(h1) : kind == 2 ? 0.9 + 0.2 * unit(h1) :
                            exp((unit(h1) * 2 - 1) * 30);
                        if (which == 1 ? !same(lm::log(x), ::log(x)) : !same(lm::log10(x), ::log10(x))) bad++;
                        break;
                }
                case 2:
                        x = kind == 0 ? unit(h1) * 10 : kind == 1 ? exp((unit(h1) * 2 - 1) * 50) :
                            kind == 2 ? unit(h1) : unit(h1) * 1e4;
                        y = kind == 0 ? (unit(h2) * 2 - 1) * 3 : kind == 1 ? (unit(h2) * 2 - 1) * 10 :
                            kind == 2 ? 1. / (1. - unit(h2) * 3) : (unit(h2) * 2 - 1) * 2;
                        if (!same(lm::pow(x, y), ::pow(x, y))) bad++;
                        break;
                default: {
                        const double lim[4] = { 0.9, 2.5, 7., 1e8 };
                        x = (unit(h1) * 2 - 1) * lim[kind];
                        if (which == 3 && !same(lm::sin(x), ::sin(x))) bad++;
                        if (which == 4 && !same(lm::cos(x), ::cos(x))) bad++;
                        if (which == 5) {
                                double s, c, gs, gc;
                                lm::sincos(x, &s, &c);
                                ::sincos(x, &gs, &gc);
                                if (!same(s, gs) || !same(c, gc)) bad++;
                        }
                }
                }
        }
        return bad;
}
