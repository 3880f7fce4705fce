/*
 * Device transport engine: differential cross sections of the inelastic
 * processes (default models SSR / SSR / DRSS and close ionisation) and the
 * FP32 spline evaluation of tabulated DCS.
 */
#ifndef PUMAS_ENGINE_DCS_H
#define PUMAS_ENGINE_DCS_H

#include "engine_common.h"

namespace pg {

/* Extra per physics data used by the exact DCS models. */
struct ModelView {
        const double * z13;   /* lm::pow(Z, 1/3) per element */
        const double * zm13;  /* lm::pow(Z, -1/3) */
        const double * z23;   /* lm::pow(Z, 2/3) */
        const double * dn;    /* 1.54 lm::pow(A, 0.27) */
        const double * lrad;  /* radiation logarithm */
        const double * cst;   /* e, e^0.5, e^-1/6, e^1/3, e^-1/3, (alpha re)^2 */
        const double *gq9_x, *gq9_w, *gq12_x, *gq12_w;
};

/* Units of the PROPOSAL based models, src/pumas.c:13800-13803, 14410-14412. */
constexpr double SSR_SQRTE = 1.648721270700128;
constexpr double SSR_ME = 0.5109989461;
constexpr double SSR_RE = 2.8179403227E-13;
constexpr double SSR_MMU = 105.6583745;

PG_FN double sq(double x) { return x * x; }

/*
 * Bremsstrahlung, Sandrock, Soedingrekso & Rhode (arXiv:1910.07050), as
 * dcs_bremsstrahlung_SSR, src/pumas.c:13915-14005. Returns m^2/GeV.
 */
PG_FN double dcs_bremsstrahlung_SSR(const ModelView & mv, int iel, double Z,
    double A, double m, double K, double q)
{
        if ((Z <= 0) || (A <= 0) || (m <= 0) || (K <= 0) || (q <= 0))
                return 0.;

        const double z13 = mv.z13[iel];
        if (q >= K + m * (1. - 0.75 * SSR_SQRTE * z13)) return 0.;

        /* Radiative correction coefficients of the SSR parameterisation. */
        const double a0 = -0.00349, a1 = 148.84, a2 = -987.531;
        const double b0 = 0.1642, b1 = 132.573, b2 = -585.361, b3 = 1407.77;
        const double c0 = -2.8922, c1 = -19.0156, c2 = 57.698, c3 = -63.418,
                     c4 = 14.1166, c5 = 1.84206;
        const double d0 = 2134.19, d1 = 581.823, d2 = -2708.85, d3 = 4767.05,
                     d4 = 1.52918, d5 = 0.361933;

        const double energy = (K + m) * 1E+03;
        const double v = q * 1E+03 / energy;
        m *= 1E+03;

        const double Z13 = 1. / z13;
        const double rad_log = mv.lrad[iel];
        const double rad_log_inel = (Z == 1.) ? 446. : 1429.;
        const double Dn = mv.dn[iel];

        const double mu_qc = m / (SSR_MMU * mv.cst[0] / Dn);
        const double rho = sqrt(1. + 4. * mu_qc * mu_qc);

        const double log_rho = lm::log((rho + 1.) / (rho - 1.));
        const double delta1 = lm::log(mu_qc) + 0.5 * rho * log_rho;
        const double delta2 = lm::log(mu_qc) +
            0.25 * (3. * rho - rho * rho * rho) * log_rho + 2. * mu_qc * mu_qc;

        const double delta = m * m * v / (2. * energy * (1. - v));

        double phi1 = lm::log(rad_log * Z13 * (m / SSR_ME) /
            (1. + rad_log * Z13 * mv.cst[1] * delta / SSR_ME));
        double phi2 = lm::log(rad_log * Z13 * mv.cst[2] * (m / SSR_ME) /
            (1. + rad_log * Z13 * mv.cst[3] * delta / SSR_ME));
        phi1 -= delta1 * (1. - 1. / Z);
        phi2 -= delta2 * (1. - 1. / Z);

        const double s_atomic_1 = lm::log(m / delta /
            (m * delta / (SSR_ME * SSR_ME) + SSR_SQRTE));
        const double s_atomic_2 = lm::log(1. + SSR_ME /
            (delta * rad_log_inel * Z13 * Z13 * SSR_SQRTE));
        const double s_atomic = (4. / 3. * (1. - v) + v * v) *
            (s_atomic_1 - s_atomic_2);

        double s_rad;
        if (v < .0 || v > 1.) {
                s_rad = 0.;
        } else if (v < 0.02) {
                s_rad = a0 + a1 * v + a2 * v * v;
        } else if (v >= 0.02 && v < 0.1) {
                s_rad = b0 + b1 * v + b2 * v * v + b3 * v * v * v;
        } else if (v >= 0.01 && v < 0.9) {
                s_rad = c0 + c1 * v + c2 * v * v;
                const double tmp = lm::log(1. - v);
                s_rad += c3 * v * lm::log(v) + c4 * tmp + c5 * tmp * tmp;
        } else {
                s_rad = d0 + d1 * v + d2 * v * v;
                const double tmp = lm::log(1. - v);
                s_rad += d3 * v * lm::log(v) + d4 * tmp + d5 * tmp * tmp;
        }

        const double result = ((2. - 2. * v + v * v) * phi1 - 2. / 3. *
            (1. - v) * phi2) + 1. / Z * s_atomic +
            0.25 * ALPHA_EM * phi1 * s_rad;
        if (result <= 0.) return 0.;

        const double aux = 2 * (SSR_ME / m) * SSR_RE * Z;
        return aux * aux * (ALPHA_EM / q) * result * 1E-04;
}

/* Kinematic range of pair production, src/pumas.c:14230-14239. */
PG_FN void dcs_pair_production_range(const ModelView & mv, int iel, double m,
    double K, double * qmin, double * qmax)
{
        const double sqrte = 1.6487212707;
        const double Z13 = mv.z13[iel];
        *qmax = K + m * (1. - 0.75 * sqrte * Z13);
        *qmin = 4. * ELECTRON_MASS;
}

/*
 * Real dilogarithm by the Chebyshev expansion of CERNLIB RDILOG (C332), the
 * method of math_dilog, src/pumas.c:12609-12686, with the same argument
 * transformations and coefficients so that the pair production DDCS rounds
 * identically (CERN Program Library, http://cds.cern.ch/record/2050865),
 * every branch of math_dilog, src/pumas.c:12609-12686.
 */
PG_FN double dilog(double x)
{
        const double C[20] = { 0.42996693560813697, 0.40975987533077105,
                -0.01858843665014592, 0.00145751084062268,
                -0.00014304184442340, 0.00001588415541880,
                -0.00000190784959387, 0.00000024195180854,
                -0.00000003193341274, 0.00000000434545063,
                -0.00000000060578480, 0.00000000008612098,
                -0.00000000001244332, 0.00000000000182256,
                -0.00000000000027007, 0.00000000000004042,
                -0.00000000000000610, 0.00000000000000093,
                -0.00000000000000014, 0.00000000000000002 };
        const double PI3 = PI * PI / 3.;
        const double PI6 = PI * PI / 6.;
        const double PI12 = PI * PI / 12.;
        /* The reference's special values, kept as they are (x = 10, and
         * -pi^2/12 at both x = -1 and x = 1). */
        if (x == 10) return PI6;
        if ((x == -1.) || (1. - x == 0.)) return -PI12;
        const double T = -x;
        double A, S, Y;
        /* The reference squares with pow(y, 2.), which its compiler
         * evaluates as y * y. */
        if (T <= -2) {
                Y = -1. / (1. + T);
                S = 1.;
                const double l0 = lm::log(-T), l1 = lm::log(1. + 1. / T);
                A = -PI3 + 0.5 * (l0 * l0 - l1 * l1);
        } else if (T < -1.) {
                Y = -1. - T;
                S = -1.;
                A = lm::log(-T);
                A = -PI6 + A * (A + lm::log(1. + 1. / T));
        } else if (T <= -0.5) {
                Y = -(1. + T) / T;
                S = 1.;
                A = lm::log(-T);
                A = -PI6 + A * (-0.5 * A + lm::log(1. + T));
        } else if (T < 0.) {
                Y = -T / (1. + T);
                S = -1.;
                const double l = lm::log(1 + T);
                A = 0.5 * (l * l);
        } else if (T <= 1.) {
                Y = T;
                S = 1.;
                A = 0.;
        } else {
                Y = 1. / T;
                S = -1.;
                const double l = lm::log(T);
                A = PI6 + 0.5 * (l * l);
        }
        const double H = Y + Y - 1;
        const double ALFA = H + H;
        double B0 = 0., B1 = 0., B2 = 0.;
        for (int i = 19; i >= 0; i--) {
                B0 = C[i] + ALFA * B1 - B2;
                B2 = B1;
                B1 = B0;
        }
        return -(S * (B0 - H * B2) + A);
}

/*
 * Pair production DDCS in the e+e- asymmetry, SSR model, as
 * dcs_pair_production_d2_SSR, src/pumas.c:14394-14544.
 */
PG_FN double dcs_pair_production_d2_SSR(const ModelView & mv, int iel,
    double Z, double A, double m, double K, double q, double rho)
{
        if ((Z <= 0) || (A <= 0) || (m <= 0) || (K <= 0) || (q <= 0))
                return 0.;

        double qmin, qmax;
        dcs_pair_production_range(mv, iel, m, K, &qmin, &qmax);
        if ((q < qmin) || (q > qmax)) return 0.;

        const double energy = (K + m) * 1E+03;
        const double v = q / (K + m);
        m *= 1E+03;
        const double rad_log = mv.lrad[iel];

        const double const_prefactor = 4. / (3. * PI) * Z * mv.cst[5];
        const double Z13 = mv.zm13[iel];
        const double d_n = mv.dn[iel];

        rho = 1 - rho;
        const double rho2 = rho * rho;

        double g1, g2;
        if (Z == 1.) {
                g1 = 4.4E-05;
                g2 = 4.8E-05;
        } else {
                g1 = 1.95E-05;
                g2 = 5.3E-05;
        }

        const double zeta1 = (0.073 * lm::log(energy / m /
            (1. + g1 * mv.z23[iel] * energy / m)) - 0.26);
        const double zeta2 = (0.058 * lm::log(energy / m /
            (1 + g2 / Z13 * energy / m)) - 0.14);

        double zeta;
        if ((zeta1 > 0.) && (zeta2 > 0.)) {
                zeta = zeta1 / zeta2;
        } else {
                zeta = 0.;
        }

        const double beta = v * v / (2. * (1. - v));
        const double xi = sq(m * v / (2. * SSR_ME)) * (1. - rho2) / (1. - v);

        const double e05 = mv.cst[1], e13 = mv.cst[3], e16 = mv.cst[2],
                     em13 = mv.cst[4];

        /* Diagram e */
        const double Be = ((2. + rho2) * (1. + beta) +
            xi * (3. + rho2)) * lm::log(1. + 1. / xi) +
            (1. - rho2 - beta) / (1. + xi) - (3. + rho2);

        const double Ce2 = ((1. - rho2) * (1. + beta) +
            xi * (3. - rho2)) * lm::log(1. + 1. / xi) +
            2. * (1. - beta - rho2) / (1. + xi) - (3. - rho2);
        const double Ce1 = Be - Ce2;

        const double De = ((2. + rho2) * (1. + beta) +
            xi * (3. + rho2)) * dilog(1. / (1. + xi)) -
            (2. + rho2) * xi * lm::log(1. + 1. / xi) -
            (xi + rho2 + beta) / (1. + xi);

        double Le1, Le2;
        if (De / Be > 0.) {
                const double tmp = De / Be;
                const double Xe = (tmp < 100) ? lm::exp(-tmp) : 0.;
                Le1 = lm::log(rad_log * Z13 * sqrt(1. + xi) /
                    (Xe + 2. * SSR_ME * e05 * rad_log * Z13 * (1. + xi) /
                    (energy * v * (1. - rho2)))) - tmp -
                    0.5 * lm::log(Xe + sq(SSR_ME / m * d_n) * (1. + xi));

                Le2 = lm::log(rad_log * Z13 * e16 * sqrt(1 + xi) /
                    (Xe + 2. * SSR_ME * e13 * rad_log * Z13 * (1. + xi) /
                    (energy * v * (1. - rho2)))) - tmp -
                    0.5 * lm::log(Xe + sq(SSR_ME / m * d_n) *
                    em13 * (1. + xi));
        } else {
                const double tmp = De / Be;
                const double Xe_inv = (tmp > -100) ? lm::exp(De / Be) : 0;
                Le1 = lm::log(rad_log * Z13 * sqrt(1. + xi) /
                    (1. + Xe_inv * 2. * SSR_ME * e05 * rad_log * Z13 *
                    (1. + xi) / (energy * v * (1. - rho2)))) - 0.5 * De / Be -
                    0.5 * lm::log(1. + Xe_inv * sq(SSR_ME / m * d_n) * (1. + xi));

                Le2 = lm::log(rad_log * Z13 * e16 * sqrt(1 + xi) /
                    (1. + Xe_inv * 2. * SSR_ME * e13 * rad_log * Z13 *
                    (1. + xi) / (energy * v * (1. - rho2)))) - 0.5 * De / Be -
                    0.5 * lm::log(1. + Xe_inv * sq(SSR_ME / m * d_n) *
                    em13 * (1. + xi));
        }

        double diagram_e = const_prefactor * (Z + zeta) * (1. - v) / v *
            (Ce1 * Le1 + Ce2 * Le2);
        if (diagram_e < 0.) diagram_e = 0.;

        /* Diagram mu */
        const double Bm = ((1. + rho2) * (1. + (3. * beta) / 2) - 1. / xi *
            (1. + 2. * beta) * (1. - rho2)) * lm::log(1. + xi) +
            xi * (1. - rho2 - beta) / (1. + xi) +
            (1. + 2. * beta) * (1. - rho2);

        const double Cm2 = ((1. - beta) * (1. - rho2) -
            xi * (1. + rho2)) * lm::log(1. + xi) / xi -
            2. * (1. - beta - rho2) / (1. + xi) +
            1. - beta - (1. + beta) * rho2;
        const double Cm1 = Bm - Cm2;

        const double Dm = ((1. + rho2) * (1. + (3. * beta) / 2.) -
            1. / xi * (1. + 2. * beta) * (1. - rho2)) *
            dilog(xi / (1. + xi)) + (1. + (3. * beta) / 2.) *
            (1. - rho2) / xi * lm::log(1. + xi) + (1. - rho2 - beta / 2. *
            (1. + rho2) + (1. - rho2) / (2. * xi) * beta) * xi / (1. + xi);

        double Lm1, Lm2;
        if (Dm / Bm > 0.) {
                const double tmp = Dm / Bm;
                const double Xm = (tmp < 100) ? lm::exp(-tmp) : 0.;
                Lm1 = lm::log(m / SSR_ME * rad_log * Z13 / d_n /
                    (Xm + 2. * SSR_ME * e05 * rad_log * Z13 * (1. + xi) /
                    (energy * v * (1. - rho2)))) - tmp;
                Lm2 = lm::log(m / SSR_ME * rad_log * Z13 / d_n /
                    (Xm + 2. * SSR_ME * e13 * rad_log * Z13 * (1. + xi) /
                    (energy * v * (1. - rho2)))) - tmp;
        } else {
                const double tmp = Dm / Bm;
                const double Xmv = (tmp > -100) ? lm::exp(tmp) : 0.;
                Lm1 = lm::log(m / SSR_ME * rad_log * Z13 / d_n /
                    (1. + 2. * SSR_ME * e05 * rad_log * Z13 * (1. + xi) /
                    (energy * v * (1. - rho2)) * Xmv));
                Lm2 = lm::log(m / SSR_ME * rad_log * Z13 / d_n /
                    (1. + 2. * SSR_ME * e13 * rad_log * Z13 * (1. + xi) /
                    (energy * v * (1. - rho2)) * Xmv));
        }

        double diagram_mu = const_prefactor * (Z + zeta) * (1. - v) / v *
            sq(SSR_ME / m) * (Cm1 * Lm1 + Cm2 * Lm2);
        if (diagram_mu < 0.) diagram_mu = 0.;

        return (diagram_e + diagram_mu) * 1E-01 / energy;
}

/*
 * Pair production DCS, SSR model: 12 point Gauss-Legendre integration of
 * the DDCS over t = ln(rho), as dcs_pair_production_SSR,
 * src/pumas.c:14560-14604.
 */
PG_NOINLINE double dcs_pair_production_SSR(const ModelView & mv, int iel,
    double Z, double A_, double mass, double K, double q)
{
        if ((Z <= 0) || (A_ <= 0) || (mass <= 0) || (K <= 0) || (q <= 0))
                return 0.;

        if (q <= 4. * ELECTRON_MASS) return 0.;
        const double sqrte = 1.6487212707;
        const double Z13 = mv.z13[iel];
        if (q >= K + mass * (1. - 0.75 * sqrte * Z13)) return 0.;

        const double gamma = 1. + K / mass;
        const double x0 = 1. - 4. * ELECTRON_MASS / q;
        const double x1 = 1. - 6. / (gamma * (gamma - q / mass));

        double rmax;
        if ((x0 > 0) && (x1 > 0)) {
                rmax = sqrt(x0) * x1;
        } else {
                return 0.;
        }

        const double ri = 1. - rmax;
        if ((ri <= 0.) || (ri >= 1.)) return 0.;
        const double tmin = lm::log(ri);

        double I = 0.;
        for (int i = 0; i < 12; i++) {
                const double rho = lm::exp(mv.gq12_x[i] * tmin);
                I -= dcs_pair_production_d2_SSR(mv, iel, Z, A_, mass, K, q,
                         rho) *
                    rho * mv.gq12_w[i] * tmin;
        }
        return (I < 0) ? 0. : I;
}

/*
 * Proton structure function F2, ALLM97 parameterisation (Abramowicz &
 * Levy, DESY 97-251), as dcs_photonuclear_f2p_ALLM97, src/pumas.c:14626-14673.
 */
PG_FN double f2p_ALLM97(double x, double Q2)
{
        const double m02 = 0.31985;
        const double mP2 = 49.457;
        const double mR2 = 0.15052;
        const double Q02 = 0.52544;
        const double Lambda2 = 0.06527;

        const double cP1 = 0.28067;
        const double cP2 = 0.22291;
        const double cP3 = 2.1979;
        const double aP1 = -0.0808;
        const double aP2 = -0.44812;
        const double aP3 = 1.1709;
        const double bP1 = 0.36292;
        const double bP2 = 1.8917;
        const double bP3 = 1.8439;

        const double cR1 = 0.80107;
        const double cR2 = 0.97307;
        const double cR3 = 3.4942;
        const double aR1 = 0.58400;
        const double aR2 = 0.37888;
        const double aR3 = 2.6063;
        const double bR1 = 0.01147;
        const double bR2 = 3.7582;
        const double bR3 = 0.49338;

        const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
        const double M2 = M * M;
        const double W2 = M2 + Q2 * (1.0 / x - 1.0);
        const double t = lm::log(lm::log((Q2 + Q02) / Lambda2) / lm::log(Q02 / Lambda2));
        const double xP = (Q2 + mP2) / (Q2 + mP2 + W2 - M2);
        const double xR = (Q2 + mR2) / (Q2 + mR2 + W2 - M2);
        const double cP =
            cP1 + (cP1 - cP2) * (1.0 / (1.0 + lm::pow(t, cP3)) - 1.0);
        const double aP =
            aP1 + (aP1 - aP2) * (1.0 / (1.0 + lm::pow(t, aP3)) - 1.0);
        const double bP = bP1 + bP2 * lm::pow(t, bP3);
        const double cR = cR1 + cR2 * lm::pow(t, cR3);
        const double aR = aR1 + aR2 * lm::pow(t, aR3);
        const double bR = bR1 + bR2 * lm::pow(t, bR3);

        const double F2P = cP * lm::exp(aP * lm::log(xP) + bP * lm::log(1 - x));
        const double F2R = cR * lm::exp(aR * lm::log(xR) + bR * lm::log(1 - x));

        return Q2 / (Q2 + m02) * (F2P + F2R);
}

/*
 * Photonuclear DDCS d2S/(dq dQ2), Dutta, Reno, Sarcevic & Seckel
 * (Phys. Rev. D63 (2001) 094020) with ALLM97, as dcs_photonuclear_d2_DRSS
 * (src/pumas.c:14923-14944) with its shadowing and F2A factors
 * (src/pumas.c:14689-14720).
 */
PG_FN double dcs_photonuclear_d2_DRSS(
    double Z, double A, double ml, double K, double q, double Q2)
{
        const double cf = 4 * PI * ALPHA_EM * ALPHA_EM * HBAR_C * HBAR_C;
        const double M = 0.5 * (PROTON_MASS + NEUTRON_MASS);
        const double E = K + ml;
        const double y = q / E;
        const double x = 0.5 * Q2 / (M * q);
        const double F2p = f2p_ALLM97(x, Q2);
        double shadowing;
        if (Z == 1.)
                shadowing = 1.;
        else if (x < 0.0014)
                shadowing = lm::exp(-0.1 * lm::log(A));
        else if (x < 0.04)
                shadowing = lm::exp((0.069 * lm::log10(x) + 0.097) * lm::log(A));
        else
                shadowing = 1.;
        const double F2A = F2p * shadowing *
            (Z + (A - Z) * (1.0 + x * (-1.85 + x * (2.45 + x * (-2.35 + x)))));
        const double R = 0.;

        const double dds = (1 - y +
                               0.5 * (1 - 2 * ml * ml / Q2) *
                                   (y * y + Q2 / (E * E)) / (1 + R)) /
                (Q2 * Q2) -
            0.25 / (E * E * Q2);

        return cf * F2A * dds / q;
}

/*
 * Photonuclear DCS: 9 point Gauss-Legendre integration of the DDCS over
 * ln Q2, as dcs_photonuclear_integrated (mode 0), src/pumas.c:15002-15060.
 */
PG_NOINLINE double dcs_photonuclear_DRSS(const ModelView & mv, double Z,
    double A, double ml, double K, double q)
{
        if ((Z <= 0) || (A <= 0) || (ml <= 0) || (K <= 0)) return 0.;

        const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
        const double mpi = PION_MASS;
        if ((q <= mpi + 0.5 * mpi * mpi / M) ||
            (q >= K + ml - 0.5 * (M + ml * ml / M)))
                return 0.;

        const double E = K + ml;
        const double ml2 = ml * ml;
        const double Q2min = ml2 * (q * q - 0.5 * ml2) / (E * (E - q));
        const double Q2max = 2.0 * M * (q - mpi) - mpi * mpi;
        if ((Q2max < Q2min) | (Q2min < 0)) return 0.;

        /* Logarithmic nodes, as math_gauss_quad_initialise,
         * src/pumas.c:12575-12607. */
        const double xmin = lm::log(Q2min);
        const double xmax = lm::log(Q2max);
        const double dx = xmax - xmin;
        double ds = 0.;
        for (int i = 0; i < 9; i++) {
                const double Q2 = lm::exp(xmin + dx * mv.gq9_x[i]);
                const double w = mv.gq9_w[i] * (dx * Q2);
                double tmp = dcs_photonuclear_d2_DRSS(Z, A, ml, K, q, Q2);
                if (tmp <= 0.) continue;
                ds += tmp * w;
        }
        return (ds < 0.) ? 0. : ds;
}

/* Kinematic range of photonuclear events, src/pumas.c:11127-11134. */
PG_FN void dcs_photonuclear_range(double m, double K, double * qmin,
    double * qmax)
{
        const double M = 0.5 * (NEUTRON_MASS + PROTON_MASS);
        const double mpi = PION_MASS;
        *qmin = mpi + 0.5 * mpi * mpi / M;
        *qmax = K + m - 0.5 * (M + m * m / M);
}

/* Close collisions DCS, pumas_electronic_dcs, src/pumas.c:11082-11101. */
PG_FN double electronic_dcs(double Z, double I, double m, double K, double q)
{
        const double P2 = K * (K + 2. * m);
        const double E = K + m;
        const double Wmax = 2. * ELECTRON_MASS * P2 /
            (m * m + ELECTRON_MASS * (ELECTRON_MASS + 2. * E));
        if (q > Wmax) return 0.;
        const double Wmin = (13.6 / 19.2) * I;
        if (q <= Wmin) return 0.;

        const double E2 = E * E;
        const double beta2 = P2 / E2;
        const double a0 = 0.5 / E2;
        const double a1 = beta2 / Wmax;
        const double cs = 2 * PI * ELECTRON_MASS * ELECTRON_RADIUS *
            ELECTRON_RADIUS * Z * (a0 + 1. / q * (1 / q - a1)) / beta2;
        return cs;
}

/*
 * Macroscopic DCS of process `ip` (0 Bremsstrahlung, 1 pair production,
 * 2 photonuclear, 3 ionisation) on element `iel`, in m^2/kg, as the wrappers
 * dcs_bremsstrahlung / dcs_pair_production / dcs_photonuclear /
 * dcs_ionisation, src/pumas.c:11003-11079, 11117-11121.
 */
PG_NOINLINE double dcs_exact(const PhysicsView & p, const ModelView & mv,
    int ip, int iel, double K, double q)
{
        const double Z = p.element_Z[iel];
        const double A = p.element_A[iel];
        const double mass = p.mass;
        if (ip == 0) {
                return dcs_bremsstrahlung_SSR(mv, iel, Z, A, mass, K, q) *
                    1E+03 * AVOGADRO_NUMBER * (mass + K) / A;
        } else if (ip == 1) {
                return dcs_pair_production_SSR(mv, iel, Z, A, mass, K, q) *
                    1E+03 * AVOGADRO_NUMBER * (mass + K) / A;
        } else if (ip == 2) {
                double qmin, qmax;
                dcs_photonuclear_range(mass, K, &qmin, &qmax);
                if ((q < qmin) || (q > qmax)) return 0.;
                return dcs_photonuclear_DRSS(mv, Z, A, mass, K, q) * 1E+03 *
                    AVOGADRO_NUMBER * (mass + K) / A;
        } else {
                return electronic_dcs(Z, p.element_I[iel], mass, K, q) *
                    (mass + K) * AVOGADRO_NUMBER / (A * 1E-03);
        }
}

/* Float bracket of the DCS x nodes, src/pumas.c:4236-4244. */
PG_FN int bracketf(const float * table, int n, float value)
{
        int i1 = 0, i2 = n - 1;
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
 * DCS evaluation during the Monte Carlo with the FP32 log-spline tables
 * and the exact model fallbacks, dcs_evaluate, src/pumas.c:11243-11347.
 * Includes the Jacobian factor K / (K + m).
 */
PG_FN double dcs_evaluate(const PhysicsView & p, const ModelView & mv,
    Cache & ch, int ip, int iel, double K, double q)
{
        const double wj = K / (K + p.mass);
        if (ip == 3) return dcs_exact(p, mv, ip, iel, K, q) * wj;

        double qlow = 0;
        if (ip == 2) {
                double qmin, qmax;
                dcs_photonuclear_range(p.mass, K, &qmin, &qmax);
                if ((q < qmin) || (q > qmax))
                        return 0.;
                else
                        qlow = 2 * qmin;
        } else {
                double qmin, qmax;
                dcs_pair_production_range(mv, iel, p.mass, K, &qmin, &qmax);
                if (ip == 1)
                        qlow = 4 * qmin;
                else
                        qmin = 0.;
                if ((q < qmin) || (q > qmax)) return 0.;
        }
        if (q < qlow) return dcs_exact(p, mv, ip, iel, K, q) * wj;

        const int nx = p.n_table_dcs;
        const float xmin = p.DCS_x[0];
        const float xmax = p.DCS_x[nx - 1];
        const float x = lm::logf((float)(q / K));
        if ((x < xmin) || (x >= xmax))
                return dcs_exact(p, mv, ip, iel, K, q) * wj;
        const int ix = bracketf(p.DCS_x, nx, x);
        if ((ix < 0) || (ix >= nx - 1))
                return dcs_exact(p, mv, ip, iel, K, q) * wj;

        const int eidx = p.element_index[iel];
        float p00, p01, m00, m01, p10, p11, m10, m11, h0, h1;
        const int imax = p.n_energies - 1;
        const double Kmax = p.K[imax];
        int i0 = 0;
        if ((K >= Kmax) ||
            ((i0 = index_K(p, ch, K)) >= imax)) {
                const float * data = p.DCS +
                    (size_t)nx * 2 * ((ip * p.n_elements + eidx) * p.n_energies +
                                         imax);
                p00 = p10 = data[2 * ix];
                m00 = m10 = data[2 * ix + 1];
                p01 = p11 = data[2 * ix + 2];
                m01 = m11 = data[2 * ix + 3];
                h0 = 1.f;
                h1 = 0.f;
        } else {
                const float * data = p.DCS +
                    (size_t)nx * 2 * ((ip * p.n_elements + eidx) * p.n_energies +
                                         i0);
                p00 = data[2 * ix];
                m00 = data[2 * ix + 1];
                p01 = data[2 * ix + 2];
                m01 = data[2 * ix + 3];
                data += 2 * nx;
                p10 = data[2 * ix];
                m10 = data[2 * ix + 1];
                p11 = data[2 * ix + 2];
                m11 = data[2 * ix + 3];

                const float K0 = (float)(p.K[i0]);
                const float K1 = (float)(p.K[i0 + 1]);
                h1 = lm::logf(((float)K) / K0) / lm::logf(K1 / K0);
                h0 = 1.f - h1;
        }

        if (((p00 == 0.f) && (m00 == 0.f)) || ((p01 == 0.f) && (m01 == 0.f)) ||
            ((p10 == 0.f) && (m10 == 0.f)) || ((p11 == 0.f) && (m11 == 0.f))) {
                return dcs_exact(p, mv, ip, iel, K, q) * wj;
        }

        const float x0 = p.DCS_x[ix];
        const float x1 = p.DCS_x[ix + 1];
        const float dx = x1 - x0;
        m00 *= dx;
        m01 *= dx;
        const float t = (x - x0) / dx;
        const float c02 = -3 * (p00 - p01) - 2 * m00 - m01;
        const float c03 = 2 * (p00 - p01) + m00 + m01;
        float r = lm::expf(p00 + t * (m00 + t * (c02 + t * c03))) * h0;
        if (h1 > 0.f) {
                m10 *= dx;
                m11 *= dx;
                const float c12 = -3 * (p10 - p11) - 2 * m10 - m11;
                const float c13 = 2 * (p10 - p11) + m10 + m11;
                r += lm::expf(p10 + t * (m10 + t * (c12 + t * c13))) * h1;
        }
        return r * wj;
}

/* Forward DEL envelope normalisation, src/pumas.c:5887-5894. */
PG_FN double envelope_norm(double a, double xmin, double xmax)
{
        if (a == 1) {
                return lm::log(xmax / xmin);
        } else {
                return (lm::pow(xmin, 1. - a) - lm::pow(xmax, 1. - a)) / (a - 1.);
        }
}

} /* namespace pg */

#endif
