/*
 * Device transport engine: polar angle samplers of discrete energy losses
 * and the elastic hard scattering (Coulomb) model.
 */
#ifndef PUMAS_ENGINE_INTERACTIONS_H
#define PUMAS_ENGINE_INTERACTIONS_H

#include "engine_dcs.h"

namespace pg {

/* ------------------------------------------------------------------------
 * Minimum search used by the photonuclear polar sampler: an initial
 * bisection, src/pumas.c:12105-12183, then Brent's method,
 * src/pumas.c:12233-12325. The objective is -Q2 d2S/(dq dQ2) in ln Q2.
 * ------------------------------------------------------------------------ */
struct PhotonuclearObjective {
        double Z, A, m, K, q;

        PG_FN double operator()(double lnQ2) const
        {
                const double Q2 = lm::exp(lnQ2);
                return -dcs_photonuclear_d2_DRSS(Z, A, m, K, q, Q2) * Q2;
        }
};

PG_FN double sign_of(double a, double b) { return (b >= 0.0) ? fabs(a) : -fabs(a); }

PG_NOINLINE int find_minimum_brent(const PhotonuclearObjective & f, double xa,
    double xb, double xc, double fb, double tol, int max_iter, double * xopt,
    double * fopt)
{
        const double EPS = 1E-15;
        const double GOLD = 0.3819660;

        tol *= fabs(xc - xa);

        double x = xb, w = xb, v = xb;
        double fw = fb, fv = fb, fx = fb;
        double d = 0., e = 0.;
        int i;
        for (i = 0; i < max_iter; i++) {
                const double xm = 0.5 * (xa + xb);
                const double tol1 = tol * fabs(x) + EPS;
                const double tol2 = 2 * tol1;
                if (fabs(x - xm) <= (tol2 - 0.5 * (xb - xa))) break;
                if (fabs(e) > tol1) {
                        const double r = (x - w) * (fx - fv);
                        double q = (x - v) * (fx - fw);
                        double p = (x - v) * q - (x - w) * r;
                        q = 2. * (q - r);
                        if (q > 0.) p = -p;
                        q = fabs(q);
                        const double etemp = e;
                        e = d;
                        if (fabs(p) >= fabs(0.5 * q * etemp) ||
                            (p <= q * (xa - x)) || (p >= q * (xb - x))) {
                                e = (x >= xm) ? xa - x : xb - x;
                                d = GOLD * e;
                        } else {
                                d = p / q;
                                const double u = x + d;
                                if ((u - xa < tol2) || (xb - u < tol2)) {
                                        d = sign_of(tol1, xm - x);
                                }
                        }
                } else {
                        e = (x >= xm) ? xa - x : xb - x;
                        d = GOLD * e;
                }
                const double u =
                    (fabs(d) >= tol1) ? x + d : x + sign_of(tol1, d);
                const double fu = f(u);
                if (fu <= fx) {
                        if (u >= x) {
                                xa = x;
                        } else {
                                xb = x;
                        }
                        v = w, w = x, x = u;
                        fv = fw, fw = fx, fx = fu;
                } else {
                        if (u < x) {
                                xa = u;
                        } else {
                                xb = u;
                        }
                        if ((fu <= fw) || (w == x)) {
                                v = w;
                                w = u;
                                fv = fw;
                                fw = fu;
                        } else if ((fu <= fv) || (v == x) || (v == w)) {
                                v = u;
                                fv = fu;
                        }
                }
        }
        *xopt = x;
        *fopt = fx;
        return (i < max_iter) ? i : -1;
}

PG_NOINLINE int find_minimum(const PhotonuclearObjective & f, double xa,
    double xc, double tol, int max_iter, double * xopt, double * fopt)
{
        double fa = f(xa);
        double fc = f(xc);
        const double atol = tol * fabs(xc - xa);

        double xb = 0, fb = 0;
        int i;
        for (i = 0; i < max_iter; i++) {
                xb = 0.5 * (xa + xc);
                fb = f(xb);
                if ((fb < fa) && (fb < fc)) break;

                if (fabs(xc - xa) <= atol) {
                        *xopt = xb;
                        *fopt = fb;
                        return i;
                }
                if (fb < fc) {
                        xc = xb;
                        fc = fb;
                } else {
                        xa = xb;
                        fa = fb;
                }
        }
        if (i == max_iter) {
                if (fa < fb) {
                        fb = fa;
                        xb = xa;
                } else if (fc < fb) {
                        fb = fc;
                        xb = xc;
                }
                *xopt = xb;
                *fopt = fb;
                return -1;
        }

        const int status = find_minimum_brent(
            f, xa, xb, xc, fb, tol, max_iter - i, xopt, fopt);
        if (status < 0)
                return -2;
        else
                return status + i;
}

/* ------------------------------------------------------------------------
 * Polar angle of a DEL, mu = (1 - cos theta) / 2 (src/pumas.c:11736-11976).
 * ------------------------------------------------------------------------ */

/* Tsai's DDCS with two step rejection, polar_bremsstrahlung,
 * src/pumas.c:11759-11827 (also used for pair production, 11841). */
PG_NOINLINE double polar_bremsstrahlung(const PhysicsView & p, Rng & rng,
    int iel, double ki, double kf)
{
        const double m = p.mass;
        const double E = ki + m;
        const double nu = ki - kf;
        const double y = nu / E;

        const double tmp = 0.5 * m * nu / (E * (E - nu));
        const double mu0 = tmp * tmp;
        const double c1 = (2 * (1. - y) + y * y) * mu0;
        const double c2 = 4 * (1. - y) * mu0 * mu0;

        const double RN = p.element_rn[iel];
        const double tmp2 = m * RN / HBAR_C;
        const double muni = tmp2 * tmp2 / 6.;
        const double muc = sqrt(mu0) - mu0;
        if (muc < 0) return 0;

        double x0;
        {
                const double q = muni;
                const double r = mu0;
                const double qr = q * r;
                x0 = (1. + 2 * qr) * lm::log((1. + qr) / (r + qr)) -
                    (1. - r) * (1 + 2 * q) / (1 + q);
        }

        for (int i = 0; i < 100; i++) {
                const double r0 = rng.uniform();
                const double r1 = rng.uniform();

                const double mu = r0 * mu0 / (mu0 + 1. - r0);
                const double mus = mu0 + mu;
                const double mu2 = mus * mus;
                const double d1 = 1 / mu2;
                const double d2 = d1 * d1;
                const double score = r1 * c1 * d1;

                double pdf = c1 * d1 - mu * c2 * d2;
                if (score > pdf) continue;

                const double l2 = mu2 / (mu0 * mu0);
                const double q = muni * l2;
                const double r = mu0 * l2;
                const double qr = q * r;
                if (qr > 1E+05) continue;

                const double x = (1. + 2 * qr) * lm::log((1. + qr) / (r + qr)) -
                    (1. - r) * (1 + 2 * q) / (1 + q);

                pdf *= x / x0;
                if (score <= pdf) return mu;
        }
        return 0.;
}

/* Photonuclear polar angle, polar_photonuclear, src/pumas.c:11874-11950. */
PG_NOINLINE double polar_photonuclear(const PhysicsView & p, Rng & rng,
    int iel, double ki, double kf)
{
        const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
        const double q = ki - kf;
        const double ml = p.mass;
        const double E = ki + ml;
        const double ml2 = ml * ml;
        const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
        const double Q2max = 2.0 * M * (q - PION_MASS) - PION_MASS * PION_MASS;
        if ((Q2max < Q2min) | (Q2min < 0)) return 0.;

        const PhotonuclearObjective f = { p.element_Z[iel], p.element_A[iel],
                ml, ki, q };
        double fopt = 0., xopt = 0.;
        const double lnQ2min = lm::log(Q2min);
        const double lnQ2max = lm::log(Q2max);
        const int status =
            find_minimum(f, lnQ2min, lnQ2max, 1E-06, 100, &xopt, &fopt);
        fopt = -fopt;
        if (status == -2) return 0.;

        const double x = lm::exp(xopt) / (2 * q * M);
        const double lne = 0.1;
        if ((x <= 0.04) && (xopt + lne < lnQ2max)) {
                double f2, x2;
                const int status2 =
                    find_minimum(f, xopt + lne, lnQ2max, 1E-06, 100, &x2, &f2);
                if (status2 >= 0) {
                        f2 = -f2;
                        if (f2 > fopt) {
                                xopt = x2;
                                fopt = f2;
                        }
                }
        }

        const double rQ2 = lnQ2max - lnQ2min;
        double Q2 = 0.;
        int i;
        for (i = 0; i < 100; i++) {
                const double u = rng.uniform();
                Q2 = Q2min * lm::exp(rQ2 * u);
                const double r =
                    dcs_photonuclear_d2_DRSS(f.Z, f.A, ml, ki, q, Q2) * Q2;
                if (rng.uniform() * fopt <= r) break;
        }
        if (i == 100) return 0.;

        const double pm = sqrt(ki * (ki + 2 * ml));
        const double E1 = E - q;
        const double eps = ml / E1;
        const double p1 = E1 * sqrt((1. + eps) * (1. - eps));
        const double p2 = pm * p1;
        double tmp = p2 + ml * ml - E * E1;
        if (fabs(tmp) <= 3 * DBL_EPS * p2) tmp = 0.;
        const double a_mu = 0.5 * tmp / p2;
        const double b_mu = 0.25 / p2;
        double mu = a_mu + b_mu * Q2;
        if (mu < 0)
                mu = 0.;
        else if (mu > 1.)
                mu = 1.;
        return mu;
}

/* Ionisation kinematics, polar_ionisation, src/pumas.c:11965-11976. */
PG_FN double polar_ionisation(const PhysicsView & p, double ki, double kf)
{
        const double nu = ki - kf;
        const double p2 = ki * (ki + 2 * p.mass);
        const double E = ki + p.mass;
        const double c = (p2 - nu * (E + ELECTRON_MASS)) /
            sqrt(p2 * (p2 + nu * nu - 2 * nu * E));
        return 0.5 * (1. - c);
}

PG_FN double polar_sample(const PhysicsView & p, Rng & rng, int process,
    int iel, double ki, double kf)
{
        if (process <= 1) return polar_bremsstrahlung(p, rng, iel, ki, kf);
        if (process == 2) return polar_photonuclear(p, rng, iel, ki, kf);
        return polar_ionisation(p, ki, kf);
}

/* ------------------------------------------------------------------------
 * Coulomb scattering helpers (src/pumas.c:7203-7683). The reference
 * evaluates the pole reduction and the transport coefficients in x87
 * extended precision (`long double`); the device evaluates them in double
 * precision. They only enter the choice of the target element of a hard
 * elastic event, through cumulative cross sections.
 * ------------------------------------------------------------------------ */
struct CoulombData {
        double cs_hard;
        double normalisation;
        int n_parameters;
        double amplitude[3];
        double screening[4];
        double a[3], b[3], c[4];
        double fspin;
        double fCM[2];
};

/*
 * The element screening data hold at most 3 atomic terms and the nuclear
 * one (screening[n_parameters - 1]). Loops over the terms run to the fixed
 * bound with a guard, and the nuclear term is selected by comparisons, so
 * that the arrays are indexed by constants and stay in registers; the
 * operations of each term are unchanged.
 */
PG_FN double screening_at(const CoulombData & cd, int k)
{
        return (k == 0) ? cd.screening[0] :
            (k == 1)    ? cd.screening[1] :
            (k == 2)    ? cd.screening[2] :
                          cd.screening[3];
}

/* coulomb_frame_parameters, src/pumas.c:7492-7510. */
PG_FN void coulomb_frame_parameters(double Z, double A, double mass,
    double kinetic, double * kinetic0, double * parameters)
{
        const double Ma = Z * PROTON_MASS + (A - Z) * NEUTRON_MASS;
        double M2 = mass + Ma;
        M2 *= M2;
        const double sCM12i = 1. / sqrt(M2 + 2. * Ma * kinetic);
        parameters[0] = (kinetic + mass + Ma) * sCM12i;
        *kinetic0 = (kinetic * Ma + mass * (mass + Ma)) * sCM12i - mass;
        if (*kinetic0 < 1E-09) *kinetic0 = 1E-09;
        const double etot = kinetic + mass + Ma;
        const double betaCM2 = kinetic * (kinetic + 2. * mass) / (etot * etot);
        double rM2 = mass / Ma;
        rM2 *= rM2;
        parameters[1] = sqrt(rM2 * (1. - betaCM2) + betaCM2);
}

/* coulomb_spin_factor, src/pumas.c:7475-7479. */
PG_FN double coulomb_spin_factor(double mass, double kinetic)
{
        const double e = kinetic + mass;
        return kinetic * (e + mass) / (e * e);
}

/* coulomb_screening_parameters, src/pumas.c:7203-7387, with the per element
 * Salvat amplitudes and exponents and the nuclear radius as exported. */
PG_FN void coulomb_screening_parameters(const PhysicsView & p, int iel,
    double kinetic, double kinetic0, int * n_parameters, double * amplitude,
    double * screening)
{
        const double Z = p.element_Z[iel];
        const double mass = p.mass;
        const double aZE = ALPHA_EM * Z * (kinetic + mass);
        const double p2 = kinetic * (kinetic + 2. * mass);
        const double zeta2 = aZE * aZE / p2;
        double f;
        if (zeta2 < 1) {
                f = 1. / (1. + zeta2) + 0.20205690315959424 -
                    0.03692775514336999 * zeta2 +
                    0.008349277381922926 * zeta2 * zeta2;
                f *= zeta2;
        } else {
                const double r2i = 1. / (1. + zeta2);
                const double phi = -atan(sqrt(zeta2));
                f = -0.5 * lm::log(r2i) + 0.5 * (1. - r2i) +
                    1. / 12. * (1. - lm::cos(2 * phi) * r2i) -
                    1. / 120. * (1. - lm::cos(4 * phi) * r2i * r2i) +
                    1. / 252. * (1. - lm::cos(6 * phi) * r2i * r2i * r2i);
        }
        const double ktv = lm::exp(2 * f);

        const double p02 = kinetic0 * (kinetic0 + 2. * mass);
        const double d = 0.25 * HBAR_C * HBAR_C / p02 * ktv;
        const int n = (int)p.element_nscr[iel];
#pragma unroll
        for (int i = 0; i < 3; i++) {
                if (i >= n) break;
                amplitude[i] = p.element_amplitude[3 * iel + i];
                const double a = p.element_exponent[3 * iel + i] / BOHR_RADIUS;
                screening[i] = d * a * a;
        }
        const double RN = p.element_rn[iel];
        const double nuclear = 12. * d / (RN * RN);
#pragma unroll
        for (int i = 0; i < 4; i++)
                if (i == n) screening[i] = nuclear;
        *n_parameters = n + 1;
}

/* coulomb_nuclear_form_factor, src/pumas.c:7396-7417. */
PG_FN double coulomb_nuclear_form_factor(double mu, double N)
{
        if (mu <= 0.)
                return 1.;
        else if (mu > 1.)
                return 0.;
        const double x2 = 10. * mu / N;
        double d;
        if (x2 <= 1E-04) {
                d = 1. / (1. + 0.1 * x2);
        } else {
                const double x = sqrt(x2);
                double sx, cx;
                lm::sincos(x, &sx, &cx);
                d = 3 * (sx - x * cx) / (x2 * x);
        }
        return d * d;
}

/* coulomb_normalisation, src/pumas.c:7428-7436. */
PG_FN double coulomb_normalisation(
    double Z, double A, double mass, double kinetic, double kinetic0)
{
        const double p2 = kinetic * (kinetic + 2. * mass);
        const double p02 = kinetic0 * (kinetic0 + 2. * mass);
        double d = ALPHA_EM * Z * HBAR_C * (kinetic + mass);
        return A * 1E-03 * p2 * p02 / (d * d * PI * AVOGADRO_NUMBER);
}

/* coulomb_pole_reduction, src/pumas.c:7524-7576. */
PG_FN void coulomb_pole_reduction(CoulombData & cd)
{
        const int n = cd.n_parameters - 1;
        for (int i = 0; i < 3; i++) cd.a[i] = cd.b[i] = 0.;
        for (int i = 0; i < 4; i++) cd.c[i] = 0.;

#pragma unroll
        for (int i = 0; i < 3; i++) {
                if (i >= n) break;
                const double Ai = cd.amplitude[i];
                const double Bi = cd.screening[i];
                cd.b[i] = Ai * Ai;
#pragma unroll
                for (int j = i + 1; j < 3; j++) {
                        if (j >= n) break;
                        const double Aj = cd.amplitude[j];
                        const double Bj = cd.screening[j];
                        const double d = 2 * Ai * Aj / (Bj - Bi);
                        cd.a[i] += d;
                        cd.a[j] -= d;
                }
        }

        const double N = screening_at(cd, n);
        double Sa[4] = { 0, 0, 0, 0 }, Sb[4] = { 0, 0, 0, 0 };
#pragma unroll
        for (int i = 0; i < 3; i++) {
                if (i >= n) break;
                const double Bi = cd.screening[i];
                const double x = 1. - Bi / N;
                double r = x;
                for (int j = 0; j < 4; j++) {
                        Sa[j] += cd.a[i] / r;
                        r *= x;
                        Sb[j] += cd.b[i] / r;
                }
        }

        double tmp = 1;
        for (int i = 0; i < 4; i++) {
                cd.c[i] = ((4 - i) * Sb[3 - i] / N - Sa[3 - i]) * tmp;
                tmp *= N;
        }

#pragma unroll
        for (int i = 0; i < 3; i++) {
                if (i >= n) break;
                const double Bi = cd.screening[i];
                double d = 1. - Bi / N;
                d *= d;
                d *= d;
                cd.a[i] = (cd.a[i] - 4 * cd.b[i] / (N - Bi)) / d;
                cd.b[i] /= d;
        }
}

/* Order 0 transport coefficient, coulomb_transport_coefficients,
 * src/pumas.c:7623-7683 (only coefficient[0] is used at transport time). */
PG_FN double coulomb_cs0(double mu, const CoulombData & cd)
{
        const double S = cd.fspin;
        const int n = cd.n_parameters - 1;
        double cs0 = 0.;
#pragma unroll
        for (int i = 0; i < 3; i++) {
                if (i >= n) break;
                const double alp = cd.screening[i];
                const double r = mu / (mu + alp);
                const double L = lm::log(1. + mu / alp);
                const double I0 = r / alp;
                const double J0 = L;
                const double I1 = L - r;
                const double J1 = -alp * L;
                cs0 += cd.a[i] * (J0 - S * J1) + cd.b[i] * (I0 - S * I1);
        }

        const double N = screening_at(cd, n);
        const double r = mu / (mu + N);
        const double L = lm::log(1. + mu / N);
        const double I0 = r / N;
        const double J0 = L;
        const double I1 = L - r;
        const double J1 = -N * L;

        const double rn = 1. / (1. + mu / N);
        const double K0 = (1. - rn) * (1. + rn) / (2 * N * N);
        const double L0 = (1. - rn) * ((1. + rn) * (1. + rn) - rn) /
            (3 * N * N * N);
        const double K1 = I0 - N * K0;
        const double L1 = K0 - N * L0;

        cs0 += cd.c[0] * (J0 - S * J1) + cd.c[1] * (I0 - S * I1) +
            cd.c[2] * (K0 - S * K1) + cd.c[3] * (L0 - S * L1);
        return cs0;
}

/* coulomb_restricted_cs, src/pumas.c:7591-7610. */
PG_FN double coulomb_restricted_cs(double mu, const CoulombData & cd)
{
        if ((mu >= 1.) || (mu >= 1E+06 * screening_at(cd, cd.n_parameters - 1)))
                return 0.;
        const double cs0 = coulomb_cs0(1., cd);
        const double cs1 = coulomb_cs0(mu, cd);
        if (cs1 > cs0)
                return 0;
        else
                return cs0 - cs1;
}

} /* namespace pg */

#endif
