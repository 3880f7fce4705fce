/*
 * FP32-bulk table interpolation: single precision mirrors of the CEL, DEL
 * and scattering tables and the lookups the FP32-bulk engine makes on them
 * every step. Layouts are those of the FP64 tables (src/pumas.c:4260-4752);
 * the arithmetic is the same as in engine_common.h, in float.
 */
#ifndef PUMAS_ENGINE_BULK_H
#define PUMAS_ENGINE_BULK_H

#include "engine_common.h"

namespace pg {

/* The single precision tables, reached through pointers of type P: plain
 * pointers to HBM, or pointers to a copy in LDS (see the stepping kernel). */
template <class P> struct FloatTablesT {
        int n_energies, n_materials;
        float mass;
        P K, X, X_dK, dE, dE_dK, Omega, Omega_dK;
        P CS, CS_dK, Mu0, Mu0_dK, Ms1, Ms1_dK, Lb;
        P a_max, b_max;
};

using FloatTables = FloatTablesT<const float *>;

/* Bracketing rows memory for float grids (see Cache). */
struct FloatCache {
        int k[2];
        int x[2];

        PG_FN void reset(int n_energies)
        {
                k[0] = k[1] = x[0] = x[1] = n_energies - 2;
        }
};

/* Bracketing row of a value missed by the first cached row: second cached
 * row and neighbours of the first one, then bisection (as
 * cached_index_miss). */
template <class P>
PG_FN int findex_miss(P table, int n, float value, int * last)
{
        const int i0 = last[0];
        const float below = table[(i0 > 0) ? i0 - 1 : 0];
        const float lo = table[i0];
        const float hi = table[i0 + 1];
        const float above = table[(i0 + 2 < n) ? i0 + 2 : n - 1];
        const float l0 = table[last[1]], l1 = table[last[1] + 1];
        if ((value >= l0) && (value < l1)) return last[1];
        int i1;
        if ((i0 > 0) && (value >= below) && (value < lo)) {
                i1 = i0 - 1;
        } else if ((i0 + 2 < n) && (value >= hi) && (value < above)) {
                i1 = i0 + 1;
        } else {
                const int imax = n - 1;
                if (value < table[0]) return -1;
                if (value >= table[imax]) return imax;
                int j1 = 0, j2 = imax;
                do {
                        const int j3 = (j1 + j2) / 2;
                        if (value >= table[j3])
                                j1 = j3;
                        else
                                j2 = j3;
                } while (j2 - j1 >= 2);
                i1 = j1;
        }
        if (i1 != last[0])
                last[0] = i1;
        else if (i1 != last[1])
                last[1] = i1;
        return i1;
}

PG_FN float fpchip(float t, float p0, float p1, float m0, float m1)
{
        const float c2 = -3.f * (p0 - p1) - 2.f * m0 - m1;
        const float c3 = 2.f * (p0 - p1) + m0 + m1;
        return p0 + t * (m0 + t * (c2 + t * c3));
}

/* Interpolation with speculative rows at the first cached bracket (one
 * round trip to memory on a cache hit), as interpolate. */
template <class P>
PG_FN float finterpolate(const FloatTablesT<P> & f, int * last, P tx, P ty,
    P tm, float x)
{
        int i1 = last[0];
        float x0 = tx[i1], x1 = tx[i1 + 1];
        float y0 = ty[i1], y1 = ty[i1 + 1];
        float d0 = tm[i1], d1 = tm[i1 + 1];
        if (!((x >= x0) && (x < x1))) {
                i1 = findex_miss(tx, f.n_energies, x, last);
                x0 = tx[i1];
                x1 = tx[i1 + 1];
                y0 = ty[i1];
                y1 = ty[i1 + 1];
                d0 = tm[i1];
                d1 = tm[i1 + 1];
        }
        const float dX = x1 - x0;
        const float t = (x - x0) / dX;
        const float m1 = d0 * dX;
        const float m2 = (i1 + 1 > 1) ? d1 * dX : m1;
        return fpchip(t, y0, y1, m1, m2);
}

template <class P>
PG_FN P fcol2(const FloatTablesT<P> & f, P t, int scheme, int material)
{
        return t + (scheme2(scheme) * f.n_materials + material) * f.n_energies;
}

template <class P>
PG_FN P fcol1(const FloatTablesT<P> & f, P t, int material)
{
        return t + material * f.n_energies;
}

/* CSDA range, as cel_grammage. */
template <class P>
PG_FN float fcel_grammage(const FloatTablesT<P> & f, FloatCache & ch, int scheme,
    int material, float kinetic)
{
        const int imax = f.n_energies - 1;
        if (kinetic < f.K[0]) return 0.f;
        const P X = fcol2(f, f.X, scheme, material);
        if (kinetic >= f.K[imax]) {
                const float a = f.a_max[material];
                const float b = f.b_max[scheme2(scheme) * f.n_materials +
                    material];
                const float K1 = a / b + f.mass;
                return X[imax] +
                    1.f / b * logf((kinetic + K1) / (f.K[imax] + K1));
        }
        return finterpolate(f, ch.k, f.K, X, fcol2(f, f.X_dK, scheme, material),
            kinetic);
}

/* Stopping power, as cel_energy_loss. */
template <class P>
PG_FN float fcel_energy_loss(const FloatTablesT<P> & f, FloatCache & ch,
    int scheme, int material, float kinetic)
{
        const int imax = f.n_energies - 1;
        if (kinetic < f.K[0]) return 0.f;
        if (kinetic >= f.K[imax])
                return f.a_max[material] +
                    f.b_max[scheme2(scheme) * f.n_materials + material] *
                    (kinetic + f.mass);
        return finterpolate(f, ch.k, f.K, fcol2(f, f.dE, scheme, material),
            fcol2(f, f.dE_dK, scheme, material), kinetic);
}

/* Straggling, as cel_straggling. */
template <class P>
PG_FN float fcel_straggling(
    const FloatTablesT<P> & f, FloatCache & ch, int material, float kinetic)
{
        const int imax = f.n_energies - 1;
        if (kinetic < f.K[0]) return 0.f;
        const P Om = fcol1(f, f.Omega, material);
        if (kinetic >= f.K[imax]) {
                const float r = (kinetic + f.mass) / (f.K[imax] + f.mass);
                return Om[imax] * r * r;
        }
        return finterpolate(
            f, ch.k, f.K, Om, fcol1(f, f.Omega_dK, material), kinetic);
}

/* DEL cross section, as del_cross_section. */
template <class P>
PG_FN float fdel_cross_section(
    const FloatTablesT<P> & f, FloatCache & ch, int material, float kinetic)
{
        const int imax = f.n_energies - 1;
        if (kinetic < f.K[0]) return 0.f;
        const P CS = fcol1(f, f.CS, material);
        if (kinetic >= f.K[imax]) return CS[imax];
        return finterpolate(
            f, ch.k, f.K, CS, fcol1(f, f.CS_dK, material), kinetic);
}

/* Multiple scattering parameters, as table_get_msc. */
template <class P>
PG_FN void ftable_get_msc(const FloatTablesT<P> & f, FloatCache & ch, int scheme,
    int material, float kinetic, float * mu0, float * invlb1)
{
        const int imax = f.n_energies - 1;
        const int s = scheme2(scheme) + 1;
        const P Mu0 = fcol1(f, f.Mu0, material);
        const P Ms1 = f.Ms1 + (s * f.n_materials + material) *
            f.n_energies;
        if (kinetic < f.K[1]) {
                *mu0 = Mu0[1];
                *invlb1 = Ms1[1] * sqrtf(f.K[1] / kinetic);
        } else if (kinetic >= f.K[imax]) {
                *mu0 = Mu0[imax];
                *invlb1 = Ms1[imax] * f.K[imax] / kinetic;
        } else {
                const P m0 = fcol1(f, f.Mu0_dK, material);
                const P m1 = f.Ms1_dK +
                    (s * f.n_materials + material) * f.n_energies;
                int i = ch.k[0];
                float x0 = f.K[i], x1 = f.K[i + 1];
                float a0 = Mu0[i], a1 = Mu0[i + 1], b0 = m0[i], b1 = m0[i + 1];
                float c0 = Ms1[i], c1 = Ms1[i + 1], e0 = m1[i], e1 = m1[i + 1];
                if (!((kinetic >= x0) && (kinetic < x1))) {
                        i = findex_miss(f.K, f.n_energies, kinetic, ch.k);
                        x0 = f.K[i];
                        x1 = f.K[i + 1];
                        a0 = Mu0[i];
                        a1 = Mu0[i + 1];
                        b0 = m0[i];
                        b1 = m0[i + 1];
                        c0 = Ms1[i];
                        c1 = Ms1[i + 1];
                        e0 = m1[i];
                        e1 = m1[i + 1];
                }
                const float dx = x1 - x0;
                const float t = (kinetic - x0) / dx;
                *mu0 = fpchip(t, a0, a1, b0 * dx, b1 * dx);
                *invlb1 = fpchip(t, c0, c1, e0 * dx, e1 * dx);
        }
}

/* Hard elastic path, as coulomb_ehs_length. */
template <class P>
PG_FN float fcoulomb_ehs_length(
    const FloatTablesT<P> & f, FloatCache & ch, int material, float kinetic)
{
        const float p2 = kinetic * (kinetic + 2.f * f.mass);
        const int imax = f.n_energies - 1;
        const P Lb = fcol1(f, f.Lb, material);
        if (kinetic < f.K[1]) return Lb[1] / p2;
        if (kinetic >= f.K[imax]) return Lb[imax] / p2;
        int i = ch.k[0];
        float x0 = f.K[i], x1 = f.K[i + 1], l0 = Lb[i], l1 = Lb[i + 1];
        if (!((kinetic >= x0) && (kinetic < x1))) {
                i = findex_miss(f.K, f.n_energies, kinetic, ch.k);
                x0 = f.K[i];
                x1 = f.K[i + 1];
                l0 = Lb[i];
                l1 = Lb[i + 1];
        }
        const float h = (kinetic - x0) / (x1 - x0);
        return (l0 + h * (l1 - l0)) / p2;
}

} /* namespace pg */

#endif
