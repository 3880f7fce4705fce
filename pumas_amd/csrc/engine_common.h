/*
 * Device transport engine: common definitions, random stream and the physics
 * table view with its interpolation routines.
 *
 * The engine is written once for the device (hipcc, gfx950). Every routine
 * that reproduces a reference routine cites it (src/pumas.c:line) and keeps
 * its floating point operation order, because per-history parity with the
 * reference is the acceptance test. Host pointers and callbacks have no
 * counterpart here. Table lookups return the unique bracketing row, which is
 * what the reference's cached and uncached paths both return for the non
 * decreasing tables checked at export time.
 */
#ifndef PUMAS_ENGINE_COMMON_H
#define PUMAS_ENGINE_COMMON_H

#include <stdint.h>

#include "libm.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PG_FN __device__ __forceinline__
#define PG_HD __host__ __device__ __forceinline__
#define PG_NOINLINE __device__ __noinline__
#else
#include <math.h>
#define PG_FN inline
#define PG_HD inline
#define PG_NOINLINE __attribute__((noinline))
#endif

namespace pg {

/* Physical constants and tuning parameters, src/pumas.c:53-200. */
constexpr double ALPHA_EM = 7.2973525693E-03;
constexpr double HBAR_C = 1.973269804E-16;
constexpr double BOHR_RADIUS = 0.529177210903E-10;
constexpr double ELECTRON_MASS = 0.510998910E-03;
constexpr double ELECTRON_RADIUS = 2.817940285E-15;
constexpr double PROTON_MASS = 0.938272;
constexpr double NEUTRON_MASS = 0.939565;
constexpr double PION_MASS = 0.13957018;
constexpr double AVOGADRO_NUMBER = 6.02214076E+23;
constexpr double LARMOR_FACTOR = 0.299792458;
constexpr double BMC_ALPHA = 2.0;
constexpr double EPSILON_X = 3E-03;
constexpr double STEP_MIN = 1E-07;
constexpr double PI = 3.14159265358979323846;
constexpr double FLT_EPS = 1.19209290e-07; /* FLT_EPSILON as a double */
constexpr double DBL_EPS = 2.2204460492503131e-16;
constexpr double DBL_MAXV = 1.7976931348623157e+308;
constexpr int N_DEL_PROCESSES = 4;

/* Enum values of include/pumas.h. */
enum {
        MODE_DISABLED = -1,
        MODE_CSDA = 0,
        MODE_MIXED = 1,
        MODE_STRAGGLED = 2,
        MODE_WEIGHTED = 0,
        MODE_RANDOMISED = 1,
        MODE_FORWARD = 0,
        MODE_BACKWARD = 1
};

enum {
        EVENT_NONE = 0,
        EVENT_LIMIT_ENERGY = 1,
        EVENT_LIMIT_DISTANCE = 2,
        EVENT_LIMIT_GRAMMAGE = 4,
        EVENT_LIMIT_TIME = 8,
        EVENT_LIMIT = 15,
        EVENT_MEDIUM = 16,
        EVENT_VERTEX_BREMSSTRAHLUNG = 32,
        EVENT_VERTEX_PAIR_CREATION = 64,
        EVENT_VERTEX_PHOTONUCLEAR = 128,
        EVENT_VERTEX_DELTA_RAY = 256,
        EVENT_VERTEX_DEL = 480,
        EVENT_VERTEX_COULOMB = 512,
        EVENT_VERTEX_DECAY = 1024,
        EVENT_VERTEX = 2016,
        EVENT_WEIGHT = 2048,
        EVENT_START = 4096,
        EVENT_STOP = 8192
};

enum {
        RETURN_SUCCESS = 0,
        RETURN_ACCURACY_ERROR = 1,
        RETURN_DECAY_ERROR = 3,
        RETURN_DENSITY_ERROR = 4,
        RETURN_DIRECTION_ERROR = 5,
        RETURN_MEDIUM_ERROR = 12,
        RETURN_MISSING_LIMIT = 15,
        RETURN_VALUE_ERROR = 24
};

/* Per-history status of a history whose end state is not the reference's
 * (enum pumas_gpu_history_status of include/pumas_gpu.h). Not errors of
 * the call: they are counted, and the history's rc carries them. */
enum {
        HISTORY_STEP_LIMIT = 100, /* the step cap ended it */
        HISTORY_LOOP_LIMIT = 101, /* a sampling or bisection loop gave up */
        HISTORY_NO_TARGET = 102   /* backward DEL target selection found no
                                     target: the reference aborts */
};

enum { STEP_CHECK = 0, STEP_RAW = 1 };

/* ------------------------------------------------------------------------
 * Counter based random stream: Philox4x32-10 (Salmon et al., SC'11).
 *
 * key     = (seed low 32 bits, seed high 32 bits)
 * counter = (draw index / 2, stream, history id low, history id high)
 * draw 2k uses output words (x0, x1), draw 2k+1 uses (x2, x3), as the 64-bit
 * word (x1 << 32 | x0); the uniform is ((w >> 11) + 0.5) 2^-53 in (0, 1).
 * ------------------------------------------------------------------------ */
struct Philox4 {
        uint32_t x[4];
};

PG_FN Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
    uint32_t c3, uint32_t k0, uint32_t k1)
{
        const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
        const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
        for (int r = 0; r < 10; r++) {
                const uint64_t p0 = (uint64_t)M0 * c0;
                const uint64_t p1 = (uint64_t)M1 * c2;
                const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
                const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
                c0 = hi1 ^ c1 ^ k0;
                c1 = lo1;
                c2 = hi0 ^ c3 ^ k1;
                c3 = lo0;
                k0 += W0;
                k1 += W1;
        }
        Philox4 out;
        out.x[0] = c0;
        out.x[1] = c1;
        out.x[2] = c2;
        out.x[3] = c3;
        return out;
}

/* Stream identifiers. */
enum { STREAM_SAMPLING = 0, STREAM_TRANSPORT = 1 };

struct Rng {
        uint32_t k0, k1;
        uint32_t h0, h1;
        uint32_t stream;
        uint32_t count;
        /* Second word of the last block, for the following odd draw. */
        uint32_t next_lo, next_hi;

        PG_FN void init(uint64_t seed, uint64_t history, uint32_t stream_,
            uint32_t count_)
        {
                k0 = (uint32_t)seed;
                k1 = (uint32_t)(seed >> 32);
                h0 = (uint32_t)history;
                h1 = (uint32_t)(history >> 32);
                stream = stream_;
                count = count_;
                if (count & 1u) {
                        const Philox4 r =
                            philox4x32_10(count >> 1, stream, h0, h1, k0, k1);
                        next_lo = r.x[2];
                        next_hi = r.x[3];
                }
        }

        /* 64 random bits of draw `count`; a block serves two draws. */
        PG_FN uint64_t bits()
        {
                const uint32_t i = count++;
                if (i & 1u) return ((uint64_t)next_hi << 32) | next_lo;
                const Philox4 r = philox4x32_10(i >> 1, stream, h0, h1, k0, k1);
                next_lo = r.x[2];
                next_hi = r.x[3];
                return ((uint64_t)r.x[1] << 32) | r.x[0];
        }

        PG_FN double uniform()
        {
                const uint64_t w = bits();
                return ((double)(w >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        }

        /* The same draw rounded to single precision (FP32-bulk engine),
         * so that both engines consume one random sequence. */
        PG_FN float uniformf() { return (float)uniform(); }
};

/* ------------------------------------------------------------------------
 * Physics tables in device memory (PTAB sections, see ptab.h).
 * ------------------------------------------------------------------------ */
struct PhysicsView {
        int n_energies, n_materials, n_elements, n_components, n_table_dcs;
        int max_components;
        double mass, ctau, cutoff;

        const double *K, *K_dX, *K_dNI_in, *K_dNI_el;
        const double *X, *X_dK, *X_dT, *T, *T_dK, *dE, *dE_dK;
        const double *Omega, *Omega_dK, *NI_el, *NI_el_dK, *NI_in, *NI_in_dK;
        const double *CS, *CS_dK, *CSf, *CSf_dK, *CSn, *CSn_dK;
        const double *Kt, *a_max, *b_max, *Mu0, *Mu0_dK, *Lb, *Lb_dK;
        const double *Ms1, *Ms1_dK;
        const double *material_density, *component_fraction;
        const double *element_Z, *element_A, *element_I;
        const double *element_nscr, *element_amplitude, *element_exponent;
        const double *element_rn, *element_lrad;
        const float *DCS, *DCS_x, *DCS_envelope;
        const int *elements_in, *component_element, *element_index;
        const int *material_component0; /* first component of a material */
        /* Coarse index of the kinetic energy grid (see k_index_fast), or
         * NULL. Lookups through this view use it on a far miss. */
        const int16_t *k_coarse;
        int k_bin0, k_nbins;
        static constexpr bool coarse_index = true;
};

/* Transport configuration (value part of struct pumas_context). */
struct Config {
        int energy_loss, decay, direction, scattering;
        int event;
        double limit_energy, limit_distance, limit_grammage, limit_time;
        double accuracy;
        /* Iterations after which a rejection or bisection loop gives up. */
        int loop_max;
};

/* Table accessors, src/pumas.c:4260-4752. */
PG_FN int scheme2(int scheme) { return (scheme > MODE_MIXED) ? MODE_MIXED : scheme; }

template <class V, class C>
PG_FN C tab2(const V & p, C t, int scheme, int material, int row)
{
        return t + (scheme2(scheme) * p.n_materials + material) * p.n_energies +
            row;
}

template <class V, class C>
PG_FN C tab1(const V & p, C t, int material, int row)
{
        return t + material * p.n_energies + row;
}

template <class V, class C>
PG_FN C tab_NI_el(const V & p, C t, int scheme, int material, int row)
{
        const int s = (scheme >= MODE_MIXED) ? MODE_MIXED : MODE_CSDA;
        return t + (s * p.n_materials + material) * p.n_energies + row;
}

template <class V, class C>
PG_FN C tab_Ms1(const V & p, C t, int scheme, int material, int row)
{
        return t + ((scheme2(scheme) + 1) * p.n_materials + material) *
            p.n_energies + row;
}

template <class V>
PG_FN double b_max(const V & p, int scheme, int material)
{
        return p.b_max[scheme2(scheme) * p.n_materials + material];
}

/* Hermite cubic, src/pumas.c:4118-4125. */
PG_FN double pchip(double t, double p0, double p1, double m0, double m1)
{
        const double c2 = -3 * (p0 - p1) - 2 * m0 - m1;
        const double c3 = 2 * (p0 - p1) + m0 + m1;
        return p0 + t * (m0 + t * (c2 + t * c3));
}

/*
 * Row i1 with table[i1] <= value < table[i1 + 1], -1 below the table and
 * n - 1 above it: the result of table_index / table_bracket
 * (src/pumas.c:4174-4233).
 */
template <class C>
PG_FN int table_index(C table, int n, double value)
{
        const int imax = n - 1;
        if (value < table[0]) return -1;
        if (value >= table[imax]) return imax;
        int i1 = 0, i2 = imax;
        do {
                const int i3 = (i1 + i2) / 2;
                if (value >= table[i3])
                        i1 = i3;
                else
                        i2 = i3;
        } while (i2 - i1 >= 2);
        return i1;
}

/*
 * Memory of the last bracketing rows, one pair for the kinetic energy grid
 * and one pair shared by the other grids, as the reference's per-context
 * index cache (src/pumas.c:437-442, 4174-4210). The bracketing row of a
 * value in a non decreasing table is unique, so the cache changes the cost
 * of a lookup, not its result.
 */
struct Cache {
        int k[2];
        int x[2];

        PG_FN void reset(int n_energies)
        {
                k[0] = k[1] = x[0] = x[1] = n_energies - 2;
        }
};

/*
 * Coarse index of the kinetic energy grid: the positive doubles are binned
 * by their exponent and three leading mantissa bits (bins of 1/8 octave,
 * narrower than the grid's rows), and bin b holds the bracketing row of its
 * lower edge. The bracketing row of x is then that of its bin or one of the
 * next rows, found by a few comparisons instead of a bisection; the result
 * is the same row, the grid being increasing.
 */
PG_HD int k_bin(double x)
{
        return (int)(__builtin_bit_cast(uint64_t, x) >> 49);
}

/* Bracketing row of K[0] <= x < K[n - 1] from the coarse index, or -1. */
template <class V>
PG_FN int k_index_fast(const V & p, double x)
{
        if (p.k_coarse == nullptr) return -1;
        const int b = k_bin(x) - p.k_bin0;
        if ((b < 0) || (b >= p.k_nbins)) return -1;
        int i = p.k_coarse[b];
        for (int k = 0; k < 4; k++, i++)
                if (x < p.K[i + 1]) return i;
        return -1;
}

/* Builds the coarse index of the grid K[0..n) into out[max_bins] (host
 * side, at table upload). Returns the number of bins, 0 if they do not fit
 * (the lookups then bisect). */
inline int k_coarse_build(
    const double * K, int n, int16_t * out, int max_bins, int * bin0)
{
        if ((n < 3) || !(K[1] > 0.)) return 0;
        const int b0 = k_bin(K[1]);
        const int nb = k_bin(K[n - 1]) - b0 + 1;
        if ((nb <= 0) || (nb > max_bins)) return 0;
        int i = 0;
        for (int b = 0; b < nb; b++) {
                const double edge =
                    __builtin_bit_cast(double, (uint64_t)(b0 + b) << 49);
                while ((i + 1 < n - 1) && (K[i + 1] <= edge)) i++;
                out[b] = (int16_t)i;
        }
        *bin0 = b0;
        return nb;
}

/*
 * Bracketing row of a value missed by the first cached row. The second
 * cached row and the two rows next to the first one are checked together
 * (one round trip to memory): between two lookups of a history the value
 * usually moves by less than a row. Then the coarse index when `kp` (the
 * tables, for a lookup on their kinetic energy grid) has one, and the
 * bisection last.
 */
template <class C, class V = PhysicsView>
PG_FN int cached_index_miss(
    C table, int n, double value, int * last, const V * kp = nullptr)
{
        const int i0 = last[0];
        const double below = table[(i0 > 0) ? i0 - 1 : 0];
        const double lo = table[i0];
        const double hi = table[i0 + 1];
        const double above = table[(i0 + 2 < n) ? i0 + 2 : n - 1];
        const double l0 = table[last[1]], l1 = table[last[1] + 1];
        if ((value >= l0) && (value < l1)) return last[1];
        int i1;
        if ((i0 > 0) && (value >= below) && (value < lo)) {
                i1 = i0 - 1;
        } else if ((i0 + 2 < n) && (value >= hi) && (value < above)) {
                i1 = i0 + 1;
        } else {
                i1 = -1;
                if (kp && (value >= table[0]) && (value < table[n - 1]))
                        i1 = k_index_fast(*kp, value);
                if (i1 < 0) {
                        i1 = table_index(table, n, value);
                        if ((i1 < 0) || (i1 >= n - 1)) return i1;
                }
        }
        if (i1 != last[0])
                last[0] = i1;
        else if (i1 != last[1])
                last[1] = i1;
        return i1;
}

template <class C, class V = PhysicsView>
PG_FN int cached_index(
    C table, int n, double value, int * last, const V * kp = nullptr)
{
        if ((value >= table[last[0]]) && (value < table[last[0] + 1]))
                return last[0];
        return cached_index_miss(table, n, value, last, kp);
}

/* Bracketing row in the kinetic energy grid. */
template <class V>
PG_FN int index_K(const V & p, Cache & ch, double kinetic)
{
        if constexpr (V::coarse_index)
                return cached_index(p.K, p.n_energies, kinetic, ch.k, &p);
        else
                return cached_index(p.K, p.n_energies, kinetic, ch.k);
}

/* Far miss on the kinetic energy grid of `p`. */
template <class V>
PG_FN int k_miss(const V & p, Cache & ch, double kinetic)
{
        if constexpr (V::coarse_index)
                return cached_index_miss(p.K, p.n_energies, kinetic, ch.k, &p);
        else
                return cached_index_miss(p.K, p.n_energies, kinetic, ch.k);
}

/* src/pumas.c:4146-4157. */
/*
 * The rows of the first cached bracket are loaded together with its bounds,
 * before the bracket is checked, so that a cache hit (the common case: the
 * kinetic energy changes little between lookups) costs one round trip to
 * memory instead of two. A miss reloads them.
 */
template <class V, class A, class B, class C>
PG_FN double interpolate(const V & p, int * last, A tx, B ty, C tm, double x)
{
        int i1 = last[0];
        double x0 = tx[i1], x1 = tx[i1 + 1];
        double y0 = ty[i1], y1 = ty[i1 + 1];
        double d0 = tm[i1], d1 = tm[i1 + 1];
        if (!((x >= x0) && (x < x1))) {
                if constexpr (V::coarse_index)
                        i1 = cached_index_miss(tx, p.n_energies, x, last,
                            (tx == p.K) ? &p : nullptr);
                else
                        i1 = cached_index_miss(tx, p.n_energies, x, last);
                x0 = tx[i1];
                x1 = tx[i1 + 1];
                y0 = ty[i1];
                y1 = ty[i1 + 1];
                d0 = tm[i1];
                d1 = tm[i1 + 1];
        }
        const double dX = x1 - x0;
        const double t = (x - x0) / dX;
        const double m1 = d0 * dX;
        const double m2 = (i1 + 1 > 1) ? d1 * dX : m1;
        return pchip(t, y0, y1, m1, m2);
}

/*
 * Lookups without a cached row (a history's first lookups, in the closed
 * form CSDA of the entry kernel). They return the row table_index returns
 * (the bracketing row of a non decreasing table is unique), but the loads
 * of each round are independent of one another: a lookup costs two or
 * three round trips to memory instead of the bisection's chain of
 * log2(n) dependent loads.
 */

/* A table value read through the constant address space: a scalar load
 * when the address is the same for every lane of the wavefront (the
 * tables are not written while a kernel runs). */
PG_FN double uload(const double * t)
{
#ifdef __HIP_DEVICE_COMPILE__
        typedef const __attribute__((address_space(4))) double cdouble;
        return *(cdouble *)t;
#else
        return *t;
#endif
}

/* Rows i and i + 1 of a table, as one 16 B load (8 B aligned). */
PG_FN void row_pair(const double * t, int i, double & a, double & b)
{
        double v[2];
        __builtin_memcpy(v, t + i, sizeof(v));
        a = v[0];
        b = v[1];
}

/* Row of table[0] <= value < table[n - 1]: rounds of W - 1 pivots loaded
 * together, down to at most W rows, then those rows loaded together. The
 * rows with table[j] <= value are a prefix of the table, so the count of
 * loaded rows at or below the value places it. */
template <int W = 8, class C>
PG_FN int table_row_wide(C table, int n, double value, int lo = 0,
    int hi = -1)
{
        if (hi < 0) hi = n - 1; /* table[lo] <= value < table[hi] */
        double v[W - 1];
        while (hi - lo > W) {
                /* Pivots lo + span k / W, k = 1..W-1: strictly between lo
                 * and hi, increasing (span > W). */
                const int span = hi - lo;
#pragma unroll
                for (int k = 1; k < W; k++)
                        v[k - 1] = table[lo + (span * k) / W];
                int count = 0;
#pragma unroll
                for (int k = 1; k < W; k++)
                        count += (v[k - 1] <= value) ? 1 : 0;
                const int nhi = lo + (span * (count + 1)) / W;
                lo += (span * count) / W;
                hi = nhi;
        }
        /* The rows past hi count 0: table[j] >= table[hi] > value. */
        int count = 0;
        if ((W % 2 == 0) && (lo + W <= n - 1)) {
                /* Rows lo + 1 .. lo + W in 16 B pairs. */
#pragma unroll
                for (int k = 1; k < W; k += 2) {
                        double a, b;
                        row_pair(table, lo + k, a, b);
                        count += ((a <= value) ? 1 : 0) + ((b <= value) ? 1 : 0);
                }
                return lo + count;
        }
#pragma unroll
        for (int k = 1; k < W; k++)
                v[k - 1] = table[(lo + k < hi) ? lo + k : hi];
#pragma unroll
        for (int k = 1; k < W; k++)
                count += (v[k - 1] <= value) ? 1 : 0;
        return lo + count;
}

/* table_row_wide for a table that is the same for every active lane of
 * the wavefront: its first round, 15 pivots over the whole table, reads
 * them with scalar loads (the tables are not written while a kernel runs),
 * then the lanes go on in their own sixteenth of the table. */
template <class C>
PG_FN int table_row_uniform(C table, int n, double value)
{
#ifdef __HIP_DEVICE_COMPILE__
        typedef const __attribute__((address_space(4))) double cdouble;
        cdouble * const t = (cdouble *)table;
#else
        const double * const t = table;
#endif
        const int span = n - 1;
        int lo = 0, hi = span;
        if (span > 16) {
                /* Pivots strictly increasing: a valid bracket [lo, hi]. */
                int count = 0;
#pragma unroll
                for (int k = 1; k < 16; k++)
                        count += (t[(span * k) / 16] <= value) ? 1 : 0;
                lo = (span * count) / 16;
                hi = (span * (count + 1)) / 16;
        }
        return table_row_wide(table, n, value, lo, hi);
}

/* Row of K[0] <= x < K[n - 1] on the kinetic energy grid: the coarse index
 * with its four candidate rows loaded together (k_index_fast), else the
 * wide search. */
template <class V>
PG_FN int k_row(const V & p, double x)
{
        if constexpr (V::coarse_index) {
                const int b = k_bin(x) - p.k_bin0;
                if ((p.k_coarse != nullptr) && (b >= 0) && (b < p.k_nbins)) {
                        const int i = p.k_coarse[b];
                        const int last = p.n_energies - 1;
                        int count = 0;
                        if (i + 4 <= last) {
                                /* K[i + 1 .. i + 4] in two 16 B pairs. */
                                double a, b1, c, d;
                                row_pair(p.K, i + 1, a, b1);
                                row_pair(p.K, i + 3, c, d);
                                count = ((a <= x) ? 1 : 0) +
                                    ((b1 <= x) ? 1 : 0) + ((c <= x) ? 1 : 0) +
                                    ((d <= x) ? 1 : 0);
                        } else {
#pragma unroll
                                for (int k = 1; k <= 4; k++) {
                                        const int j =
                                            (i + k < last) ? i + k : last;
                                        count += (p.K[j] <= x) ? 1 : 0;
                                }
                        }
                        if (count < 4) return i + count;
                }
        }
        return table_row_wide(p.K, p.n_energies, x);
}

/* The interpolation of `interpolate` on a known row i1: the same
 * operations, so the same result. */
PG_FN double interpolate_row(
    int i1, const double * tx, const double * ty, const double * tm, double x)
{
        double x0, x1, y0, y1, d0, d1;
        row_pair(tx, i1, x0, x1);
        row_pair(ty, i1, y0, y1);
        row_pair(tm, i1, d0, d1);
        const double dX = x1 - x0;
        const double t = (x - x0) / dX;
        const double m1 = d0 * dX;
        const double m2 = (i1 + 1 > 1) ? d1 * dX : m1;
        return pchip(t, y0, y1, m1, m2);
}

/* cel_grammage, src/pumas.c:3644-3669. */
template <class V>
PG_FN double cel_grammage(const V & p, Cache & ch, int scheme, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[0]) return 0.;
        if (kinetic >= p.K[imax]) {
                const double a = p.a_max[material];
                const double b = b_max(p, scheme, material);
                const double K0 = p.K[imax];
                const double K1 = a / b + p.mass;
                return *tab2(p, p.X, scheme, material, imax) +
                    1. / b * lm::log((kinetic + K1) / (K0 + K1));
        }
        return interpolate(p, ch.k, p.K, tab2(p, p.X, scheme, material, 0),
            tab2(p, p.X_dK, scheme, material, 0), kinetic);
}

/* cel_grammage_as_time, src/pumas.c:3680-3709. */
template <class V>
PG_FN double cel_grammage_as_time(const V & p, Cache & ch, int scheme, int material, double time)
{
        const int imax = p.n_energies - 1;
        const auto T = tab2(p, p.T, scheme, material, 0);
        if (time < T[0]) return 0.;
        if (time >= T[imax]) {
                const double a = p.a_max[material];
                const double b = b_max(p, scheme, material);
                const double E0 = p.K[imax] + p.mass;
                const double tmax = T[imax];
                const double r =
                    E0 / (a + b * E0) * lm::exp(a * (time - tmax) / p.mass);
                const double kinetic = a * r / (1. - b * r) - p.mass;
                const double K1 = a / b + p.mass;
                return *tab2(p, p.X, scheme, material, imax) +
                    1. / b * lm::log((kinetic + K1) / (E0 - p.mass + K1));
        }
        return interpolate(p, ch.x, T, tab2(p, p.X, scheme, material, 0),
            tab2(p, p.X_dT, scheme, material, 0), time);
}

/* cel_proper_time, src/pumas.c:3720-3745. */
template <class V>
PG_FN double cel_proper_time(const V & p, Cache & ch, int scheme, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[0]) return 0.;
        if (kinetic >= p.K[imax]) {
                const double a = p.a_max[material];
                const double b = b_max(p, scheme, material);
                const double E0 = p.K[imax] + p.mass;
                const double E1 = kinetic + p.mass;
                return *tab2(p, p.T, scheme, material, imax) +
                    p.mass / a * lm::log((E1 / E0) * (a + b * E0) / (a + b * E1));
        }
        return interpolate(p, ch.k, p.K, tab2(p, p.T, scheme, material, 0),
            tab2(p, p.T_dK, scheme, material, 0), kinetic);
}

/* cel_kinetic_energy, src/pumas.c:3758-3784. */
template <class V>
PG_FN double cel_kinetic_energy(const V & p, Cache & ch, int scheme, int material, double grammage)
{
        const int imax = p.n_energies - 1;
        const auto X = tab2(p, p.X, scheme, material, 0);
        if (grammage < X[0]) return 0.;
        if (grammage >= X[imax]) {
                const double a = p.a_max[material];
                const double b = b_max(p, scheme, material);
                const double K0 = p.K[imax];
                const double K1 = a / b + p.mass;
                return (K0 + K1) * lm::exp(b * (grammage - X[imax])) - K1;
        }
        return interpolate(
            p, ch.x, X, p.K, tab2(p, p.K_dX, scheme, material, 0), grammage);
}

/* cel_energy_loss, src/pumas.c:3796-3814. */
template <class V>
PG_FN double cel_energy_loss(const V & p, Cache & ch, int scheme, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[0]) return 0.;
        if (kinetic >= p.K[imax]) {
                return p.a_max[material] +
                    b_max(p, scheme, material) * (kinetic + p.mass);
        }
        return interpolate(p, ch.k, p.K, tab2(p, p.dE, scheme, material, 0),
            tab2(p, p.dE_dK, scheme, material, 0), kinetic);
}

/* cel_straggling, src/pumas.c:3826-3846. */
template <class V>
PG_FN double cel_straggling(const V & p, Cache & ch, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[0]) return 0.;
        const double Kmax = p.K[imax];
        if (kinetic >= Kmax) {
                const double r = (kinetic + p.mass) / (Kmax + p.mass);
                return *tab1(p, p.Omega, material, imax) * r * r;
        }
        return interpolate(p, ch.k, p.K, tab1(p, p.Omega, material, 0),
            tab1(p, p.Omega_dK, material, 0), kinetic);
}

/* del_cross_section, src/pumas.c:3895-3912. */
template <class V>
PG_FN double del_cross_section(const V & p, Cache & ch, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[0]) return 0.;
        if (kinetic >= p.K[imax]) return *tab1(p, p.CS, material, imax);
        return interpolate(p, ch.k, p.K, tab1(p, p.CS, material, 0),
            tab1(p, p.CS_dK, material, 0), kinetic);
}

/* del_interaction_length, src/pumas.c:3923-3947. */
template <class V>
PG_FN double del_interaction_length(const V & p, Cache & ch, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[0]) return 0.;
        if (kinetic >= p.K[imax]) {
                const double k0 = p.K[imax];
                const double a0 = p.a_max[material];
                const double b0 = b_max(p, MODE_MIXED, material);
                const double cs = *tab1(p, p.CS, material, imax);
                const double dZ = cs / b0 *
                    lm::log((a0 + b0 * (kinetic + p.mass)) /
                        (a0 + b0 * (k0 + p.mass)));
                return *tab1(p, p.NI_in, material, imax) + dZ;
        }
        return interpolate(p, ch.k, p.K, tab1(p, p.NI_in, material, 0),
            tab1(p, p.NI_in_dK, material, 0), kinetic);
}

/* del_kinetic_from_interaction_length, src/pumas.c:3959-3987. */
template <class V>
PG_FN double del_kinetic_from_interaction_length(const V & p, Cache & ch, int material, double nI)
{
        const int imax = p.n_energies - 1;
        const auto NI = tab1(p, p.NI_in, material, 0);
        if (nI < NI[0]) return 0.;
        if (nI >= NI[imax]) {
                const double k0 = p.K[imax];
                const double a0 = p.a_max[material];
                const double b0 = b_max(p, MODE_MIXED, material);
                const double cs = *tab1(p, p.CS, material, imax);
                const double nI0 = NI[imax];
                return ((a0 + b0 * (k0 + p.mass)) * lm::exp(b0 * (nI - nI0) / cs) -
                           a0) / b0 - p.mass;
        }
        return interpolate(
            p, ch.x, NI, p.K, tab1(p, p.K_dNI_in, material, 0), nI);
}

/* ehs_interaction_length, src/pumas.c:3998-4023. */
template <class V>
PG_FN double ehs_interaction_length(const V & p, Cache & ch, int scheme, int material, double kinetic)
{
        const int imax = p.n_energies - 1;
        const auto NI = tab_NI_el(p, p.NI_el, scheme, material, 0);
        if (kinetic < p.K[0]) return 0.;
        if (kinetic >= p.K[imax]) {
                const double k0 = p.K[imax - 1];
                const double k1 = p.K[imax];
                const double nI0 = NI[imax - 1];
                const double nI1 = NI[imax];
                return nI1 + (kinetic - k1) * (nI1 - nI0) / (k1 - k0);
        }
        return interpolate(p, ch.k, p.K, NI,
            tab_NI_el(p, p.NI_el_dK, scheme, material, 0), kinetic);
}

/* ehs_kinetic_from_interaction_length, src/pumas.c:4034-4058. */
template <class V>
PG_FN double ehs_kinetic_from_interaction_length(const V & p, Cache & ch, int scheme, int material, double nI)
{
        const int imax = p.n_energies - 1;
        const auto NI = tab_NI_el(p, p.NI_el, scheme, material, 0);
        if (nI < NI[0]) return 0.;
        if (nI >= NI[imax]) {
                const double k0 = p.K[imax - 1];
                const double k1 = p.K[imax];
                const double nI0 = NI[imax - 1];
                const double nI1 = NI[imax];
                return k1 + (nI - nI1) * (k1 - k0) / (nI1 - nI0);
        }
        /* K_dNI_el shares the two-scheme layout of the other CEL tables. */
        return interpolate(
            p, ch.x, NI, p.K, tab2(p, p.K_dNI_el, scheme, material, 0), nI);
}

/* table_get_msc, src/pumas.c:4077-4111. */
template <class V>
PG_FN void table_get_msc(const V & p, Cache & ch, int scheme, int material,
    double kinetic, double * mu0, double * invlb1)
{
        const int imax = p.n_energies - 1;
        if (kinetic < p.K[1]) {
                *mu0 = *tab1(p, p.Mu0, material, 1);
                *invlb1 = *tab_Ms1(p, p.Ms1, scheme, material, 1) *
                    sqrt(p.K[1] / kinetic);
        } else if (kinetic >= p.K[imax]) {
                *mu0 = *tab1(p, p.Mu0, material, imax);
                *invlb1 = *tab_Ms1(p, p.Ms1, scheme, material, imax) *
                    p.K[imax] / kinetic;
        } else {
                /* Speculative rows at the first cached bracket, as in
                 * interpolate. */
                const auto y0 = tab1(p, p.Mu0, material, 0);
                const auto m0 = tab1(p, p.Mu0_dK, material, 0);
                const auto y1 = tab_Ms1(p, p.Ms1, scheme, material, 0);
                const auto m1 = tab_Ms1(p, p.Ms1_dK, scheme, material, 0);
                int i = ch.k[0];
                double x0 = p.K[i], x1 = p.K[i + 1];
                double a0 = y0[i], a1 = y0[i + 1], b0 = m0[i], b1 = m0[i + 1];
                double c0 = y1[i], c1 = y1[i + 1], e0 = m1[i], e1 = m1[i + 1];
                if (!((kinetic >= x0) && (kinetic < x1))) {
                        i = k_miss(p, ch, kinetic);
                        x0 = p.K[i];
                        x1 = p.K[i + 1];
                        a0 = y0[i];
                        a1 = y0[i + 1];
                        b0 = m0[i];
                        b1 = m0[i + 1];
                        c0 = y1[i];
                        c1 = y1[i + 1];
                        e0 = m1[i];
                        e1 = m1[i + 1];
                }
                const double dx = x1 - x0;
                const double t = (kinetic - x0) / dx;
                *mu0 = pchip(t, a0, a1, b0 * dx, b1 * dx);
                *invlb1 = pchip(t, c0, c1, e0 * dx, e1 * dx);
        }
}

/* coulomb_ehs_length, src/pumas.c:7446-7466. */
template <class V>
PG_FN double coulomb_ehs_length(const V & p, Cache & ch, int material, double kinetic)
{
        const double p2 = kinetic * (kinetic + 2. * p.mass);
        const int imax = p.n_energies - 1;
        const auto Lb = tab1(p, p.Lb, material, 0);
        if (kinetic < p.K[1]) {
                return Lb[1] / p2;
        } else if (kinetic >= p.K[imax]) {
                return Lb[imax] / p2;
        } else {
                int i = ch.k[0];
                double x0 = p.K[i], x1 = p.K[i + 1];
                double l0 = Lb[i], l1 = Lb[i + 1];
                if (!((kinetic >= x0) && (kinetic < x1))) {
                        i = k_miss(p, ch, kinetic);
                        x0 = p.K[i];
                        x1 = p.K[i + 1];
                        l0 = Lb[i];
                        l1 = Lb[i + 1];
                }
                const double h = (kinetic - x0) / (x1 - x0);
                return (l0 + h * (l1 - l0)) / p2;
        }
}

} /* namespace pg */

#endif
