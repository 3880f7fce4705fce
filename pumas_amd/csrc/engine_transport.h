/*
 * Device transport engine: one Monte Carlo history, from a start state to an
 * end condition, as pumas_context_transport (src/pumas.c:3114-3251) with the
 * geometry and random stream of this library.
 *
 * A history keeps the reference's per-context stepping memory
 * (struct simulation_context, src/pumas.c:426-477) in the Track below, so
 * that draws, step lengths and events follow the reference one for one.
 */
#ifndef PUMAS_ENGINE_TRANSPORT_H
#define PUMAS_ENGINE_TRANSPORT_H

#include "engine_bulk.h"
#include "engine_interactions.h"
#include "geometry.h"

namespace pg {

constexpr int MAX_COMPONENTS = 8;

struct Track {
        /* struct pumas_state */
        double charge, energy, distance, grammage, time, weight;
        double pos[3], dir[3];
        int decayed;
        /* Stepping memory of the reference simulation context. */
        int step_first, step_event, step_foreseen;
        double step_X_limit, step_invlb1, lifetime;
        int randn_done;
        double randn_next;
        /* Table row memory, random stream and step counter. */
        Cache ix;
        FloatCache fx;
        Rng rng;
        uint32_t steps;
        int max_steps;
};

enum {
        PHASE_EMPTY = 0,
        PHASE_STEP,
        PHASE_DEL,
        PHASE_EHS,
        PHASE_DONE
};

/* A history in flight: its Track and the loop carried variables of the
 * reference's stepping loop (src/pumas.c:5231-5273). */
struct Walker {
        Track t;
        int phase;
        int record; /* frames of this history are recorded */
        int event, rc;
        int media0, end_medium;
        int medium, material, straight, step_index, step_max_type;
        /* density: the local density at the start of the next step;
         * step_max_locals: the locals step times the accuracy (0: uniform
         * medium). */
        double density, step_max_medium, step_max_locals;
        double ti, wi, Xi, Xf, dei, grammage_max;
};

/* A recorded frame: layout of struct pumas_gpu_frame (include/pumas_gpu.h),
 * the device counterpart of struct pumas_frame (include/pumas.h:386-395). */
struct Frame {
        double charge, energy, distance, grammage, time, weight;
        double position[3], direction[3];
        int decayed;
        int pad;
        int medium; /* index, -1 = outside */
        int event;
        uint64_t history;
        uint64_t reserved;
};

/* Device frame recorder, as the reference's pumas_recorder
 * (include/pumas.h:432-458): frames go to a device buffer in the order they
 * are recorded (a counter is incremented per frame, frames beyond the
 * capacity are counted and dropped), for histories with ids in
 * [first, last). `period` has the reference's meaning: every period-th
 * step, or only events and medium changes when zero or less. */
struct Recorder {
        Frame * frames; /* NULL: recording disabled */
        unsigned long long * count;
        unsigned long long capacity;
        int period;
        uint64_t first, last;
};

/*
 * The rejection and bisection loops of the reference have no iteration
 * limit. For finite inputs within the tables each of them accepts with a
 * probability bounded away from zero (at most 120 iterations over 10^7
 * voxel histories). Above the tables (kinetic energies beyond 1e9 GeV,
 * which backward discrete losses can reach from a few hundred GeV) the
 * elastic hard scattering angle loop accepts ever more rarely (the nuclear
 * form factor cuts the large angles): the reference needs up to 3e7 trials
 * per event around 1e10 GeV, and at 1e10 GeV forward one of its loops never
 * ends; a NaN or infinite input, or a sampling range that rounds to
 * nothing, can also make a loop never accept. A single such lane keeps its
 * wavefront, and so the launch, from ending: an iteration takes a few
 * hundred cycles. After Config::loop_max iterations (2^16 by default) a
 * loop gives up: its last values are used once and the history is marked
 * (max_steps = MARK_LOOP_LIMIT) so that it ends at its next step boundary
 * with rc HISTORY_LOOP_LIMIT and no event. Parity with the reference on
 * histories above the tables takes a limit of 2^25.
 * Target selections that reach the reference's assert(0) mark the history
 * the same way (MARK_NO_TARGET, rc HISTORY_NO_TARGET).
 */
constexpr int LOOP_MAX_DEFAULT = 1 << 16;
constexpr int MARK_LOOP_LIMIT = -1;
constexpr int MARK_NO_TARGET = -2;

PG_FN bool loop_stalled(Track & t, int & iterations, int limit)
{
        if (++iterations < limit) return false;
        if (t.max_steps >= 0) t.max_steps = MARK_LOOP_LIMIT;
        return true;
}

/* Appends a frame of `t` to the recorder. Out of line, with its arguments
 * copied: neither the history nor the engine object is pinned to memory on
 * the paths that do not record. */
static PG_NOINLINE void record_state(const Recorder rec, const Track t,
    int medium,
    int event)
{
#ifdef __HIP_DEVICE_COMPILE__
        const unsigned long long k = atomicAdd(rec.count, 1ull);
#else
        const unsigned long long k = (*rec.count)++;
#endif
        if (k >= rec.capacity) return;
        Frame & f = rec.frames[k];
        f.charge = t.charge;
        f.energy = t.energy;
        f.distance = t.distance;
        f.grammage = t.grammage;
        f.time = t.time;
        f.weight = t.weight;
        for (int j = 0; j < 3; j++) {
                f.position[j] = t.pos[j];
                f.direction[j] = t.dir[j];
        }
        f.decayed = t.decayed;
        f.medium = medium;
        f.event = event;
        f.pad = 0;
        f.history = ((uint64_t)t.rng.h1 << 32) | t.rng.h0;
        f.reserved = 0;
}

/*
 * The transport engine. With F32 false every operation is the reference's
 * FP64 arithmetic. With F32 true (FP32-bulk) the per-step table lookups,
 * the straggling, scattering and random number sampling run in single
 * precision; the state (energy, distance, grammage, time, weight, position,
 * direction), the range to energy inversion, the event weights and the
 * geometry stay in double precision.
 */
template <bool F32, class FT = FloatTables, class ST = PhysicsView>
struct EngineT {
        const PhysicsView & p;
        const ModelView & mv;
        const Config & c;
        const pg_geometry & g;
        const FT & ft;
        /* The FP64 tables read by every step (the same tables as `p`, or a
         * copy of them in LDS). */
        const ST & st;
        Recorder rec;

        /* Records the state of `w` in `medium` with `event`, if the history
         * is recorded (record_state, src/pumas.c:7738-7780). */
        PG_FN void record_frame(const Walker & w, int medium, int event) const
        {
                if (!w.record) return;
                record_state(rec, w.t, medium, event);
        }

        /* Per-step table lookups, in the bulk precision. */
        PG_FN double bulk_grammage(Track & t, int scheme, int material,
            double kinetic) const
        {
                if constexpr (F32)
                        return fcel_grammage(
                            ft, t.fx, scheme, material, (float)kinetic);
                else
                        return cel_grammage(st, t.ix, scheme, material, kinetic);
        }

        PG_FN double bulk_energy_loss(Track & t, int scheme, int material,
            double kinetic) const
        {
                if constexpr (F32)
                        return fcel_energy_loss(
                            ft, t.fx, scheme, material, (float)kinetic);
                else
                        return cel_energy_loss(
                            st, t.ix, scheme, material, kinetic);
        }

        PG_FN double bulk_straggling(Track & t, int material,
            double kinetic) const
        {
                if constexpr (F32)
                        return fcel_straggling(
                            ft, t.fx, material, (float)kinetic);
                else
                        return cel_straggling(st, t.ix, material, kinetic);
        }

        PG_FN double bulk_del_cross_section(Track & t, int material,
            double kinetic) const
        {
                if constexpr (F32)
                        return fdel_cross_section(
                            ft, t.fx, material, (float)kinetic);
                else
                        return del_cross_section(st, t.ix, material, kinetic);
        }

        PG_FN double bulk_ehs_length(Track & t, int material,
            double kinetic) const
        {
                if constexpr (F32)
                        return fcoulomb_ehs_length(
                            ft, t.fx, material, (float)kinetic);
                else
                        return coulomb_ehs_length(st, t.ix, material, kinetic);
        }

        PG_FN void bulk_msc(Track & t, int scheme, int material,
            double kinetic, double * mu0, double * invlb1) const
        {
                if constexpr (F32) {
                        float m, l;
                        ftable_get_msc(ft, t.fx, scheme, material,
                            (float)kinetic, &m, &l);
                        *mu0 = m;
                        *invlb1 = l;
                } else {
                        table_get_msc(
                            st, t.ix, scheme, material, kinetic, mu0, invlb1);
                }
        }

        /* -ln(u) for an exponential path length or angle. */
        PG_FN double bulk_neglog_uniform(Track & t) const
        {
                if constexpr (F32)
                        return -(double)logf(t.rng.uniformf());
                else
                        return -lm::log(t.rng.uniform());
        }

        PG_FN double sgn() const
        {
                return (c.direction == MODE_FORWARD) ? 1. : -1.;
        }

        /* Medium callback with a step proposal (counts Monte Carlo steps). */
        PG_FN int medium_step(Track & t, int * medium, double * step) const
        {
                t.steps++;
                return geo_locate(&g, t.pos, t.dir, sgn(), medium, step);
        }

        /* Location only medium callback. */
        PG_FN int medium_at(const double * pos, const Track & t) const
        {
                int m;
                geo_locate(&g, pos, t.dir, sgn(), &m, 0);
                return m;
        }

        /* transport_set_locals, src/pumas.c:5528-5552: the local density
         * of `medium` at the state's position, and the locals step times
         * the context accuracy (zero for a uniform medium). The
         * exponential profile is the device form of a non uniform
         * `pumas_locals_cb`: density * exp(-(z - z_ref) / scale), proposing
         * scale / max(|u_z|, cos_min). */
        PG_FN double set_locals(const Track & t, int medium, int * material,
            double * density) const
        {
                double d;
                geo_medium_properties(&g, medium, material, &d);
                double step_max = 0.;
                const int k = geo_locals_entry(&g, medium);
                if ((k >= 0) && (g.locals[k].kind == GEO_LOCALS_EXPONENTIAL)) {
                        const pg_locals & l = g.locals[k];
                        d = d * lm::exp(-(t.pos[2] - l.z_ref) / l.scale);
                        const double uz = fabs(t.dir[2]);
                        step_max =
                            l.scale / ((uz <= l.cos_min) ? l.cos_min : uz);
                }
                if (d <= 0.) d = p.material_density[*material];
                *density = d;
                return step_max * c.accuracy;
        }

        /* step_randn, src/pumas.c:7016-7029. */
        PG_FN double randn(Track & t) const
        {
                t.randn_done = !t.randn_done;
                if (!t.randn_done) return t.randn_next;
                if constexpr (F32) {
                        const float r = sqrtf(-2.f * logf(t.rng.uniformf()));
                        const float phi =
                            2.f * (float)PI * t.rng.uniformf();
                        float s, cp;
                        sincosf(phi, &s, &cp);
                        t.randn_next = r * cp;
                        return r * s;
                }
                const double r = sqrt(-2. * lm::log(t.rng.uniform()));
                const double phi = 2. * PI * t.rng.uniform();
                double s, cp;
                lm::sincos(phi, &s, &cp);
                t.randn_next = r * cp;
                return r * s;
        }

        /* step_rotate_direction, src/pumas.c:7042-7095. */
        PG_FN void rotate(Track & t, double mu) const
        {
                double * const direction = t.dir;
                const double cos_theta = 1 - 2 * mu;
                const double stsq = 4 * mu * (1. - mu);
                if (stsq <= 0.) return;
                const double st = sqrt(stsq);

                /* The reference's choice of the axis pair (the two
                 * components other than the smallest), with the pair
                 * selected first and normalised once: lanes of a wavefront
                 * with different pairs then run the same instructions. The
                 * operations of each pair are unchanged. */
                const double d0 = direction[0], d1 = direction[1],
                             d2 = direction[2];
                const double a0 = fabs(d0), a1 = fabs(d1), a2 = fabs(d2);
                const bool pair_xz = (a0 > a1) && (a0 > a2);
                const bool pair_xy = !(a0 > a1) && (a1 > a2);
                const bool pair_yz = !pair_xz && !pair_xy;
                const double e0 = pair_yz ? d1 : d0;
                const double e1 = pair_xy ? d1 : d2;
                const double nrm = 1. / sqrt(e0 * e0 + e1 * e1);
                const double u0x =
                    pair_xz ? -d2 * nrm : (pair_xy ? d1 * nrm : 0.);
                const double u0y =
                    pair_yz ? d2 * nrm : (pair_xy ? -d0 * nrm : 0.);
                const double u0z =
                    pair_xz ? d0 * nrm : (pair_yz ? -d1 * nrm : 0.);
                const double u1x = u0y * direction[2] - u0z * direction[1];
                const double u1y = u0z * direction[0] - u0x * direction[2];
                const double u1z = u0x * direction[1] - u0y * direction[0];

                double cp, sp;
                if constexpr (F32) {
                        const float phi =
                            (float)PI * (1.f - 2.f * t.rng.uniformf());
                        float sf, cf;
                        sincosf(phi, &sf, &cf);
                        cp = cf;
                        sp = sf;
                } else {
                        const double phi = PI * (1. - 2. * t.rng.uniform());
                        lm::sincos(phi, &sp, &cp);
                }
                direction[0] =
                    cos_theta * direction[0] + st * (cp * u0x + sp * u1x);
                direction[1] =
                    cos_theta * direction[1] + st * (cp * u0y + sp * u1y);
                direction[2] =
                    cos_theta * direction[2] + st * (cp * u0z + sp * u1z);
        }

        /* cel_kinetic_energy in CSDA for a lookup without a cached row;
         * *row is the range table's row that brackets the grammage, or -1
         * outside the table. */
        PG_FN double csda_kinetic_energy(Track & t, int material,
            double grammage, int * row, double * k0, double * k1) const
        {
                const int imax = p.n_energies - 1;
                const auto X = tab2(p, p.X, MODE_CSDA, material, 0);
#ifdef __HIP_DEVICE_COMPILE__
                /* One material for the wavefront (the usual case): the
                 * bounds and the search's first round from scalar loads. */
                const int m0 = __builtin_amdgcn_readfirstlane(material);
                const bool uniform = __all(material == m0);
#else
                const int m0 = material;
                const bool uniform = false;
#endif
                const auto Xu = tab2(p, p.X, MODE_CSDA, m0, 0);
                const double x_lo = uniform ? uload(Xu) : X[0];
                const double x_hi = uniform ? uload(Xu + imax) : X[imax];
                if ((grammage >= x_lo) && (grammage < x_hi)) {
                        const int j = uniform ?
                            table_row_uniform(Xu, p.n_energies, grammage) :
                            table_row_wide(X, p.n_energies, grammage);
                        *row = j;
                        row_pair(p.K, j, *k0, *k1);
                        return interpolate_row(j, X, p.K,
                            tab2(p, p.K_dX, MODE_CSDA, material, 0), grammage);
                }
                *row = -1;
                return cel_kinetic_energy(p, t.ix, MODE_CSDA, material, grammage);
        }

        /* transport_with_csda, src/pumas.c:4777-4949 (no magnetic field).
         * Inlined into the entry kernel, its only caller. */
        PG_FN int transport_with_csda(
            Track & t, int material, double density) const
        {
                const double ki = t.energy;
                const double di = t.distance;
                const double ti = t.time;
                /* The grammage, proper time and energy loss at ki, from
                 * one row lookup (the loss is the reference's last lookup;
                 * the order of these pure lookups does not matter). */
                double xi, Ti, dEi;
                const int imax = p.n_energies - 1;
                if ((ki >= uload(p.K)) && (ki < uload(p.K + imax))) {
                        const int i = k_row(p, ki);
                        xi = interpolate_row(i, p.K,
                            tab2(p, p.X, MODE_CSDA, material, 0),
                            tab2(p, p.X_dK, MODE_CSDA, material, 0), ki);
                        Ti = interpolate_row(i, p.K,
                            tab2(p, p.T, MODE_CSDA, material, 0),
                            tab2(p, p.T_dK, MODE_CSDA, material, 0), ki);
                        dEi = interpolate_row(i, p.K,
                            tab2(p, p.dE, MODE_CSDA, material, 0),
                            tab2(p, p.dE_dK, MODE_CSDA, material, 0), ki);
                } else {
                        xi = cel_grammage(p, t.ix, MODE_CSDA, material, ki);
                        Ti = cel_proper_time(p, t.ix, MODE_CSDA, material, ki);
                        dEi = cel_energy_loss(
                            p, t.ix, MODE_CSDA, material, ki);
                }

                int event = EVENT_NONE;
                double xB;
                if (c.event & EVENT_LIMIT_GRAMMAGE) {
                        xB = c.limit_grammage - t.grammage;
                        event = EVENT_LIMIT_GRAMMAGE;
                } else
                        xB = DBL_MAXV;
                if (c.event & EVENT_LIMIT_DISTANCE) {
                        const double xD = density * (c.limit_distance - di);
                        if (xD < xB) {
                                xB = xD;
                                event = EVENT_LIMIT_DISTANCE;
                        }
                }
                int decayed = 0;
                double time_max =
                    (c.event & EVENT_LIMIT_TIME) ? c.limit_time : 0.;
                if (c.decay == MODE_RANDOMISED) {
                        if ((time_max <= 0.) || (t.lifetime < time_max)) {
                                time_max = t.lifetime;
                                decayed = 1;
                        }
                }
                if (time_max > 0.) {
                        const double dt = time_max - ti;
                        if (dt <= 0.) {
                                xB = 0.;
                                event = EVENT_LIMIT_TIME;
                        } else {
                                const double s = sgn();
                                const double Tf = Ti - s * dt * density;
                                if (Tf > 0.) {
                                        const double xT = fabs(xi -
                                            cel_grammage_as_time(p, t.ix, MODE_CSDA, material, Tf));
                                        if (xT < xB) {
                                                xB = xT;
                                                event = EVENT_LIMIT_TIME;
                                        }
                                }
                        }
                }
                if (xB <= 0.) return event;

                double xf, kf;
                /* Row of kf on the kinetic energy grid when kf comes from
                 * the range table's row (the same row brackets it there,
                 * up to rounding), else -1. */
                int kf_row = -1;
                double kf_lo = 0., kf_hi = 0.; /* K[kf_row], K[kf_row + 1] */
                if (c.direction == MODE_FORWARD) {
                        if (xi > xB) {
                                xf = xi - xB;
                                if (xB < EPSILON_X * xi) {
                                        kf = ki - dEi * xB;
                                        if (kf < 0.) kf = 0.;
                                } else {
                                        kf = csda_kinetic_energy(t, material,
                                            xf, &kf_row, &kf_lo, &kf_hi);
                                }
                        } else {
                                xf = kf = 0.;
                                event = EVENT_LIMIT_ENERGY;
                        }
                        if (c.event & EVENT_LIMIT_ENERGY) {
                                if (ki <= c.limit_energy)
                                        return EVENT_LIMIT_ENERGY;
                                if (kf < c.limit_energy) {
                                        kf = c.limit_energy;
                                        kf_row = -1;
                                        xf = cel_grammage(p, t.ix, MODE_CSDA, material, kf);
                                        event = EVENT_LIMIT_ENERGY;
                                }
                        }
                } else {
                        if (xB == DBL_MAXV) {
                                xf = xB;
                                kf = DBL_MAXV;
                        } else {
                                xf = xB + xi;
                                if (xB < EPSILON_X * xi) {
                                        kf = ki + dEi * xB;
                                } else {
                                        kf = csda_kinetic_energy(t, material,
                                            xf, &kf_row, &kf_lo, &kf_hi);
                                }
                        }
                        if (c.event & EVENT_LIMIT_ENERGY) {
                                if (ki >= c.limit_energy)
                                        return EVENT_LIMIT_ENERGY;
                                if (kf > c.limit_energy) {
                                        kf = c.limit_energy;
                                        kf_row = -1;
                                        xf = cel_grammage(p, t.ix, MODE_CSDA, material, kf);
                                        event = EVENT_LIMIT_ENERGY;
                                }
                        }
                }
                /* The row of kf: checked against the grid, found afresh
                 * if it is not the range table's row. */
                if ((kf_row >= 0) && !((kf >= kf_lo) && (kf < kf_hi)))
                        kf_row = -1;
                if ((kf_row < 0) && (kf >= uload(p.K)) &&
                    (kf < uload(p.K + imax)))
                        kf_row = k_row(p, kf);

                const double distance = fabs(xf - xi) / density;
                t.energy = kf;
                if (event == EVENT_LIMIT_DISTANCE)
                        t.distance = c.limit_distance;
                else
                        t.distance += distance;
                if (event == EVENT_LIMIT_GRAMMAGE)
                        t.grammage = c.limit_grammage;
                else
                        t.grammage += distance * density;
                if (event == EVENT_LIMIT_TIME) {
                        t.time = time_max;
                        t.decayed = decayed;
                        if (decayed) event = EVENT_VERTEX_DECAY;
                } else {
                        const double Tf = (kf_row >= 0) ?
                            interpolate_row(kf_row, p.K,
                                tab2(p, p.T, MODE_CSDA, material, 0),
                                tab2(p, p.T_dK, MODE_CSDA, material, 0), kf) :
                            cel_proper_time(p, t.ix, MODE_CSDA, material, kf);
                        t.time += fabs(Ti - Tf) / density;
                }
                if (c.direction == MODE_BACKWARD) {
                        const double dEf = (kf_row >= 0) ?
                            interpolate_row(kf_row, p.K,
                                tab2(p, p.dE, MODE_CSDA, material, 0),
                                tab2(p, p.dE_dK, MODE_CSDA, material, 0), kf) :
                            cel_energy_loss(p, t.ix, MODE_CSDA, material, kf);
                        t.weight *= dEf / dEi;
                }
                if (c.decay == MODE_WEIGHTED)
                        t.weight *= lm::exp(-fabs(ti - t.time) / p.ctau);

                const double path =
                    (c.direction == MODE_FORWARD) ? distance : -distance;
                t.pos[0] += path * t.dir[0];
                t.pos[1] += path * t.dir[1];
                t.pos[2] += path * t.dir[2];
                return event;
        }

        /* transport_limit, src/pumas.c:5569-5696. */
        PG_FN void transport_limit(Track & t, int material, double Xi,
            double Xf, double * grammage_max) const
        {
                t.step_event = t.step_foreseen = EVENT_NONE;
                if (t.weight <= 0.) {
                        t.step_event = EVENT_WEIGHT;
                        return;
                }
                if (c.direction == MODE_FORWARD) {
                        const double kinetic_min =
                            (c.event & EVENT_LIMIT_ENERGY) ? c.limit_energy :
                                                             0.;
                        if (t.energy <= kinetic_min) {
                                t.step_event = EVENT_LIMIT_ENERGY;
                                return;
                        }
                } else if ((c.event & EVENT_LIMIT_ENERGY) &&
                    (t.energy >= c.limit_energy)) {
                        t.step_event = EVENT_LIMIT_ENERGY;
                        return;
                }

                if (c.event & EVENT_LIMIT_GRAMMAGE) {
                        *grammage_max = c.limit_grammage;
                        t.step_foreseen = EVENT_LIMIT_GRAMMAGE;
                } else {
                        *grammage_max = 0.;
                }

                const int scheme = c.energy_loss;
                if (scheme == MODE_DISABLED) {
                        if (c.scattering == MODE_MIXED) {
                                const double X = Xi -
                                    coulomb_ehs_length(p, t.ix, material, t.energy) *
                                        lm::log(t.rng.uniform());
                                if ((*grammage_max == 0.) ||
                                    (X < *grammage_max)) {
                                        *grammage_max = X;
                                        t.step_foreseen = EVENT_VERTEX_COULOMB;
                                        return;
                                }
                        }
                        return;
                }

                const double s = sgn();
                int foreseen = EVENT_NONE;
                double kinetic_limit = 0.;
                if (scheme == MODE_MIXED) {
                        for (int i = 0; i < 1000; i++) {
                                const double zeta = t.rng.uniform();
                                if ((zeta <= 0.) || (zeta >= 1.)) continue;
                                const double nI = del_interaction_length(p, t.ix, material, t.energy) + s * lm::log(zeta);
                                if (nI <= 0.) break;
                                const double k =
                                    del_kinetic_from_interaction_length(p, t.ix, material, nI);
                                if (c.direction == MODE_FORWARD) {
                                        if (k >= t.energy) continue;
                                } else {
                                        if (k <= t.energy) continue;
                                }
                                if ((c.direction == MODE_BACKWARD) ||
                                    (k > p.Kt[material])) {
                                        kinetic_limit = k;
                                        foreseen = EVENT_VERTEX_DEL;
                                }
                                break;
                        }
                }

                if ((scheme < MODE_STRAGGLED) &&
                    (c.scattering == MODE_MIXED)) {
                        const double nI = ehs_interaction_length(p, t.ix, scheme, material, t.energy) +
                            s * lm::log(t.rng.uniform());
                        if (nI > 0.) {
                                const double k =
                                    ehs_kinetic_from_interaction_length(p, t.ix, scheme, material, nI);
                                if ((kinetic_limit <= 0.) ||
                                    ((c.direction == MODE_FORWARD) &&
                                        (k > kinetic_limit)) ||
                                    ((c.direction == MODE_BACKWARD) &&
                                        (k < kinetic_limit))) {
                                        kinetic_limit = k;
                                        foreseen = EVENT_VERTEX_COULOMB;
                                }
                        }
                }

                if (foreseen == EVENT_NONE) return;

                const double X = Xi +
                    s * (Xf - cel_grammage(st, t.ix, scheme, material, kinetic_limit));
                if ((*grammage_max <= 0) || (X < *grammage_max)) {
                        *grammage_max = X;
                        t.step_foreseen = foreseen;
                }
        }

        /* step_fluctuate, src/pumas.c:6943-7005. */
        PG_FN void step_fluctuate(Track & t, int material, double Xtot,
            double dX, double * kf, double * ratio) const
        {
                const int scheme = c.energy_loss;
                const double s = sgn();
                double k1, r = 0.;
                k1 = cel_kinetic_energy(st, t.ix, scheme, material, Xtot - s * dX);
                const double dk0 = fabs(t.energy - k1);
                if ((k1 > 0.) && (dk0 > 0.)) {
                        const double Omega0 =
                            bulk_straggling(t, material, t.energy);
                        const double Omega1 = bulk_straggling(t, material, k1);
                        double dk12 = 0.5 * dX * (Omega0 + Omega1);
                        const double tmp = dX / Xtot;
                        if (tmp > 5E-02) {
                                const double de1 =
                                    bulk_energy_loss(t, scheme, material, k1);
                                const double de0 = bulk_energy_loss(
                                    t, scheme, material, t.energy);
                                dk12 *= 1. + s * dX / dk0 * (de1 - de0);
                        }
                        const double dk1 = sqrt(dk12);
                        if (dk0 >= 3. * dk1) {
                                double u;
                                int tries = 0;
                                do
                                        u = randn(t);
                                while ((fabs(u) > 3.) &&
                                    !loop_stalled(t, tries, c.loop_max));
                                u /= 1.015387;
                                k1 += u * dk1;
                        } else if (dk0 >= 1.7320508 * dk1) {
                                const double u =
                                    1.7320508 * (1. - 2. * t.rng.uniform());
                                k1 += u * dk1;
                        } else {
                                const double dk32 = 3. * dk12;
                                const double dk02 = dk0 * dk0;
                                const double a =
                                    1. - (dk32 - dk02) / (dk32 + 3. * dk02);
                                if (t.rng.uniform() <= a) {
                                        const double b =
                                            0.5 * (dk32 + 3. * dk02) / dk0;
                                        const double u = t.rng.uniform();
                                        k1 = t.energy - s * b * u;
                                        if (k1 < 0.) k1 = 0.;
                                } else {
                                        k1 = t.energy;
                                }
                        }
                        r = fabs(t.energy - k1) / dk0;
                }
                if (r == 0.) r = 1.;
                *kf = k1;
                *ratio = r;
        }

        /* step_transport, src/pumas.c:6296-6925 (no magnetic field).
         * `density_io` holds the local density at the start of the step and
         * receives the one at its end; `step_max_locals` the locals step,
         * updated at the end of the step. Returns the end medium through
         * `out_medium`. */
        PG_FN void step_transport(Track & t, int straight, int medium,
            int material, double & density_io, double & step_max_locals,
            double grammage_max, double step_max_medium, int step_max_type,
            int * out_medium) const
        {
                const int scheme = c.energy_loss;
                double * const direction = t.dir;
                double * const position = t.pos;
                const double density = density_io;
                const double density_i = 1. / density;
                double locals_density = density;
                const double momentum =
                    sqrt(t.energy * (t.energy + 2. * p.mass));

                const int tmp_scheme =
                    (scheme == MODE_DISABLED) ? MODE_CSDA : scheme;
                const double Xtot =
                    bulk_grammage(t, tmp_scheme, material, t.energy);

                double step_loc;
                double invlb1 = 0.;
                if (!straight) {
                        double r = c.accuracy;
                        const double k_threshold = 1E+09;
                        if ((scheme == MODE_STRAGGLED) &&
                            (t.energy > k_threshold)) {
                                double f = k_threshold / t.energy;
                                if (f < 0.1) f = 0.1;
                                r *= f;
                        }
                        step_loc = r * density_i * Xtot;

                        if (c.scattering == MODE_MIXED) {
                                if (t.step_first != 0) {
                                        double mu0;
                                        bulk_msc(t, scheme, material,
                                            t.energy, &mu0, &invlb1);
                                        invlb1 *= density;
                                } else
                                        invlb1 = t.step_invlb1;
                                const double stepT = c.accuracy / invlb1;
                                if (stepT < step_loc) step_loc = stepT;
                        }
                } else {
                        if (grammage_max <= 0.)
                                step_loc = 1E+09 * density_i;
                        else
                                step_loc = grammage_max * density_i;
                }

                t.step_first = 0;

                if (step_loc < STEP_MIN) step_loc = STEP_MIN;
                double step = (step_max_medium <= 0) ?
                    step_loc :
                    (step_max_medium < step_loc ? step_max_medium : step_loc);
                if ((step_max_locals > 0.) && (step > step_max_locals))
                        step = step_max_locals;

                *out_medium = medium;
                int event = EVENT_NONE;
                if (c.event & EVENT_LIMIT_DISTANCE) {
                        const double d = c.limit_distance - t.distance;
                        if (d <= step) {
                                step = d;
                                event = EVENT_LIMIT_DISTANCE;
                        }
                }

                const double sg = sgn();
                position[0] += step * sg * direction[0];
                position[1] += step * sg * direction[1];
                position[2] += step * sg * direction[2];

                int end_medium = medium_at(position, t);
                double end_position[3] = { position[0], position[1],
                        position[2] };
                if (end_medium != medium) {
                        if (step_max_type == STEP_CHECK) {
                                const double step_min =
                                    (step < STEP_MIN) ? step : STEP_MIN;
                                double pi[3] = { position[0], position[1],
                                        position[2] };
                                double s1 = 0., s2 = -step_min;
                                position[0] = pi[0] + s2 * sg * direction[0];
                                position[1] = pi[1] + s2 * sg * direction[1];
                                position[2] = pi[2] + s2 * sg * direction[2];
                                int tmp_medium = medium_at(position, t);
                                if (tmp_medium != medium) {
                                        if (tmp_medium != end_medium)
                                                end_medium = tmp_medium;
                                        s1 = s2;
                                        s2 = -step;
                                        int halvings = 0;
                                        while ((fabs(s1 - s2) > STEP_MIN) &&
                                            !loop_stalled(t, halvings, c.loop_max)) {
                                                const double s3 =
                                                    0.5 * (s1 + s2);
                                                position[0] = pi[0] +
                                                    s3 * sg * direction[0];
                                                position[1] = pi[1] +
                                                    s3 * sg * direction[1];
                                                position[2] = pi[2] +
                                                    s3 * sg * direction[2];
                                                tmp_medium =
                                                    medium_at(position, t);
                                                if (tmp_medium == medium) {
                                                        s2 = s3;
                                                } else {
                                                        s1 = s3;
                                                        if (tmp_medium !=
                                                            end_medium)
                                                                end_medium =
                                                                    tmp_medium;
                                                }
                                        }
                                        position[0] =
                                            pi[0] + s2 * sg * direction[0];
                                        position[1] =
                                            pi[1] + s2 * sg * direction[1];
                                        position[2] =
                                            pi[2] + s2 * sg * direction[2];
                                        step += s1;
                                        end_position[0] =
                                            pi[0] + s1 * sg * direction[0];
                                        end_position[1] =
                                            pi[1] + s1 * sg * direction[1];
                                        end_position[2] =
                                            pi[2] + s1 * sg * direction[2];
                                }
                        }
                        event = EVENT_MEDIUM;
                        *out_medium = end_medium;
                }

                /* The end step locals, src/pumas.c:6494-6503: at the last
                 * point inside the medium. */
                if ((step_max_locals > 0.) &&
                    ((step_max_type != STEP_RAW) || (event != EVENT_MEDIUM))) {
                        int unused;
                        step_max_locals =
                            set_locals(t, medium, &unused, &locals_density);
                }

                if (event & EVENT_MEDIUM) {
                        position[0] = end_position[0];
                        position[1] = end_position[1];
                        position[2] = end_position[2];
                }

                double k1 = t.energy, dk = 0.;
                const double dX = 0.5 * step * (density + locals_density);
                if ((scheme >= MODE_CSDA) && (scheme <= MODE_MIXED)) {
                        const double X = Xtot - sg * dX;
                        if ((c.direction == MODE_FORWARD) &&
                            (X <= t.step_X_limit)) {
                                k1 = (c.event & EVENT_LIMIT_ENERGY) ?
                                    c.limit_energy :
                                    0.;
                                const double grammage =
                                    t.grammage + Xtot - t.step_X_limit;
                                if ((grammage_max <= 0.) ||
                                    (grammage < grammage_max)) {
                                        grammage_max = grammage;
                                        event = EVENT_LIMIT_ENERGY;
                                }
                        } else if ((c.direction == MODE_BACKWARD) &&
                            (t.step_X_limit > 0.) && (X > t.step_X_limit)) {
                                k1 = (c.event & EVENT_LIMIT_ENERGY) ?
                                    c.limit_energy :
                                    0.;
                                const double grammage =
                                    t.grammage + t.step_X_limit - Xtot;
                                if ((grammage_max <= 0.) ||
                                    (grammage < grammage_max)) {
                                        grammage_max = grammage;
                                        event = EVENT_LIMIT_ENERGY;
                                }
                        } else if (dX < EPSILON_X * Xtot) {
                                k1 -= sg *
                                    bulk_energy_loss(t, scheme, material, k1) *
                                    dX;
                                if ((c.event & EVENT_LIMIT_ENERGY) &&
                                    (c.limit_energy > 0.)) {
                                        if (((c.direction == MODE_FORWARD) &&
                                                (k1 < c.limit_energy)) ||
                                            ((c.direction == MODE_BACKWARD) &&
                                                (k1 > c.limit_energy))) {
                                                k1 = c.limit_energy;
                                                event = EVENT_LIMIT_ENERGY;
                                        }
                                } else if (k1 < 0.) {
                                        k1 = 0.;
                                }
                        } else {
                                k1 = cel_kinetic_energy(st, t.ix, scheme, material, X);
                        }
                } else if (scheme == MODE_STRAGGLED) {
                        double ratio;
                        step_fluctuate(t, material, Xtot, dX, &k1, &ratio);
                        dk = fabs(t.energy - k1);

                        double kinetic_limit = -1.;
                        if (c.direction == MODE_FORWARD) {
                                const double kinetic_min =
                                    (c.event & EVENT_LIMIT_ENERGY) ?
                                    c.limit_energy :
                                    0.;
                                if (k1 <= kinetic_min)
                                        kinetic_limit = kinetic_min;
                        } else {
                                if ((c.event & EVENT_LIMIT_ENERGY) &&
                                    (k1 >= c.limit_energy))
                                        kinetic_limit = c.limit_energy;
                        }
                        if (kinetic_limit >= 0.) {
                                const double grammage = t.grammage +
                                    fabs(t.energy - kinetic_limit) / dk * dX;
                                k1 = kinetic_limit;
                                if ((grammage_max <= 0.) ||
                                    (grammage < grammage_max)) {
                                        grammage_max = grammage;
                                        event = EVENT_LIMIT_ENERGY;
                                }
                        }

                        double Xmax = dX;
                        if (grammage_max > 0.) {
                                const double dX1 = grammage_max - t.grammage;
                                if (Xmax > dX1) Xmax = dX1;
                        }
                        double k_x = -1.;

                        /* In-step inelastic DEL. */
                        do {
                                double xs_del =
                                    bulk_del_cross_section(t, material, t.energy);
                                const double tmp_del =
                                    bulk_del_cross_section(t, material, k1);
                                if (tmp_del > xs_del) xs_del = tmp_del;
                                if (xs_del <= 0.) break;
                                const double X_del =
                                    bulk_neglog_uniform(t) / xs_del;
                                if (X_del >= Xmax) break;
                                const double k_h = cel_kinetic_energy(st, t.ix, scheme, material, Xtot - sg * X_del);
                                const double k_del = t.energy -
                                    sg * ratio * fabs(t.energy - k_h);
                                if (k_del <= 0.) break;
                                const double r =
                                    bulk_del_cross_section(t, material, k_del) /
                                    xs_del;
                                if (t.rng.uniform() > r) break;
                                k_x = k_del;
                                Xmax = X_del;
                                event = EVENT_VERTEX_DEL;
                        } while (0);

                        /* In-step elastic hard scattering. */
                        if (c.scattering == MODE_MIXED) do {
                                double kmin, kmax;
                                if (k1 <= t.energy)
                                        kmin = k1, kmax = t.energy;
                                else
                                        kmax = k1, kmin = t.energy;
                                if (kmin < p.K[1]) {
                                        const double tmp = kmin;
                                        kmin = kmax, kmax = tmp;
                                }
                                const double lb_ehs =
                                    bulk_ehs_length(t, material, kmin);
                                if (lb_ehs <= 0.) break;
                                const double X_ehs =
                                    bulk_neglog_uniform(t) * lb_ehs;
                                if (X_ehs >= Xmax) break;
                                const double k_h = cel_kinetic_energy(st, t.ix, scheme, material, Xtot - sg * X_ehs);
                                const double k_ehs = t.energy -
                                    sg * ratio * fabs(t.energy - k_h);
                                if (k_ehs <= 0.) break;
                                const double r =
                                    bulk_ehs_length(t, material, k_ehs) /
                                    lb_ehs;
                                if ((r <= 0.) || (t.rng.uniform() > 1. / r))
                                        break;
                                k_x = k_ehs;
                                Xmax = X_ehs;
                                event = EVENT_VERTEX_COULOMB;
                        } while (0);

                        if (k_x >= 0.) {
                                k1 = k_x;
                                grammage_max = t.grammage + Xmax;
                        }
                }

                const double Xi = t.grammage;
                t.grammage += dX;
                const double sf0 = step;
                if ((grammage_max > 0.) && (t.grammage >= grammage_max)) {
                        const double dX_ = grammage_max - Xi;
                        if ((fabs(density - locals_density) <= FLT_EPS) ||
                            (sf0 <= FLT_EPS)) {
                                step = dX_ * density_i;
                        } else {
                                const double drho = locals_density - density;
                                double tmp =
                                    density * density + 2. * dX_ * drho / sf0;
                                tmp = (tmp > FLT_EPS) ? sqrt(tmp) : 0;
                                const double h_int = (tmp - density) / drho;
                                step *= h_int;
                        }
                        t.grammage = grammage_max;
                        if (!(event & EVENT_LIMIT_ENERGY)) {
                                if (scheme <= MODE_MIXED) {
                                        if (scheme != MODE_DISABLED)
                                                k1 = cel_kinetic_energy(st, t.ix,
                                                    scheme, material,
                                                    Xtot - sg * (t.grammage -
                                                                    Xi));
                                        event = t.step_foreseen;
                                } else if (!(event & (EVENT_VERTEX_COULOMB |
                                                         EVENT_VERTEX_DEL))) {
                                        k1 = t.energy -
                                            sg * (t.grammage - Xi) * dk / dX;
                                        if (k1 < 0.) k1 = 0.;
                                        event = EVENT_LIMIT_GRAMMAGE;
                                }
                        }

                        const double ds_ = step - sf0;
                        position[0] += ds_ * sg * direction[0];
                        position[1] += ds_ * sg * direction[1];
                        position[2] += ds_ * sg * direction[2];
                }

                int decayed = 0;
                double time_max =
                    (c.event & EVENT_LIMIT_TIME) ? c.limit_time : 0.;
                if (c.decay == MODE_RANDOMISED) {
                        if ((time_max <= 0.) || (t.lifetime < time_max)) {
                                time_max = t.lifetime;
                                decayed = 1;
                        }
                }

                const double sf1 = step;
                if (straight && (scheme != MODE_DISABLED)) {
                        const double Ti =
                            cel_proper_time(p, t.ix, scheme, material, t.energy);
                        if (time_max > 0.) {
                                const double Tf = Ti -
                                    sg * (time_max - t.time) * density;
                                if (Tf > 0.) {
                                        const double dxT = fabs(Xtot -
                                            cel_grammage_as_time(p, t.ix, scheme, material, Tf));
                                        if (Xi + dxT < t.grammage) {
                                                event = EVENT_LIMIT_TIME;
                                                t.time = time_max;
                                                t.decayed = decayed;
                                                step = dxT * density_i;
                                                if (step > sf1) step = sf1;
                                                t.grammage = Xi + dxT;
                                                const double xf =
                                                    Xtot - sg * dxT;
                                                k1 = cel_kinetic_energy(st, t.ix, scheme, material, xf);
                                        }
                                }
                        }
                        if (event != EVENT_LIMIT_TIME) {
                                const double Tf =
                                    cel_proper_time(p, t.ix, scheme, material, k1);
                                t.time += fabs(Tf - Ti) * density_i;
                        }
                } else {
                        const double p_f = (k1 <= 0) ?
                            momentum :
                            sqrt(k1 * (k1 + 2. * p.mass));
                        const double ti = t.time;
                        t.time +=
                            0.5 * step * p.mass * (1. / momentum + 1. / p_f);
                        if ((time_max > 0.) && (t.time >= time_max)) {
                                event = EVENT_LIMIT_TIME;
                                t.time = time_max;
                                t.decayed = decayed;

                                const double a = (momentum / p_f - 1.) / sf1;
                                const double cc =
                                    -2. * (time_max - ti) * momentum / p.mass;
                                if (a != 0.) {
                                        double delta = 1. - a * cc;
                                        if (delta < 0.) delta = 0.;
                                        step = 1. / a * (sqrt(delta) - 1.);
                                } else {
                                        step = -0.5 * cc;
                                }
                                if (step < 0.) step = 0.;

                                if (scheme != MODE_DISABLED) {
                                        const double p1 =
                                            momentum / (1 + a * step);
                                        k1 = sqrt(p1 * p1 - p.mass * p.mass) -
                                            p.mass;
                                        if (k1 < 0.) k1 = 0.;
                                }

                                const double Xf = Xi + dX;
                                if (density == locals_density)
                                        t.grammage = Xi + step * density;
                                else
                                        t.grammage = Xi +
                                            step * (density + 0.5 * step /
                                                sf0 * (locals_density -
                                                          density));
                                if (t.grammage > Xf) t.grammage = Xf;
                        }
                }

                if (event == EVENT_LIMIT_TIME) {
                        const double ds_ = step - sf1;
                        position[0] += ds_ * sg * direction[0];
                        position[1] += ds_ * sg * direction[1];
                        position[2] += ds_ * sg * direction[2];
                }

                t.step_event = event;
                t.energy = k1;
                t.distance += step;

                if (c.scattering == MODE_MIXED) {
                        double mu0, invlb1_;
                        if (t.energy <= 0.) {
                                t.step_invlb1 = invlb1;
                        } else {
                                bulk_msc(t, scheme, material, t.energy,
                                    &mu0, &invlb1_);
                                t.step_invlb1 = locals_density * invlb1_;
                        }
                }

                if ((invlb1 > 0.) || (t.step_invlb1 > 0.)) {
                        double ilb1 = 0.25 * step * (invlb1 + t.step_invlb1);
                        if (ilb1 > 1.) ilb1 = 1.;
                        double mu;
                        int tries = 0;
                        do {
                                mu = ilb1 * bulk_neglog_uniform(t);
                        } while ((mu > 1.) && !loop_stalled(t, tries, c.loop_max));
                        rotate(t, mu);
                }
                density_io = locals_density;
        }

        /* Position of a material's first component in the global list. */
        PG_FN int component0(int material) const
        {
                return p.material_component0[material];
        }

        /*
         * del_randomise_target, src/pumas.c:6154-6270. Returns the process
         * and sets the element (global index) and, backward, the weight
         * (`weight` holds Q on input in backward mode).
         */
        PG_FN void del_randomise_target(Track & t, int material,
            double * weight, int * process, int * element) const
        {
                int i1, i2;
                double h, dk;
                i1 = index_K(p, t.ix, t.energy);
                if (i1 < 0) {
                        i1 = i2 = 0;
                        h = dk = 0.;
                } else if (i1 >= p.n_energies - 1) {
                        i1 = i2 = p.n_energies - 1;
                        h = dk = 0.;
                } else {
                        i2 = i1 + 1;
                        const double K1 = p.K[i1];
                        dk = p.K[i2] - K1;
                        h = (t.energy - K1) / dk;
                }
                (void)i2;

                const int ic0 = component0(material);
                const int nc = p.elements_in[material];
                if (c.direction == MODE_FORWARD) {
                        const double zeta = t.rng.uniform();
                        for (int ic = ic0; ic < ic0 + nc; ic++)
                                for (int ip = 0; ip < N_DEL_PROCESSES; ip++) {
                                        const double * f = p.CSf +
                                            (ip * p.n_components + ic) *
                                                p.n_energies + i1;
                                        const double * df = p.CSf_dK +
                                            (ip * p.n_components + ic) *
                                                p.n_energies + i1;
                                        const double csf = pchip(h, f[0], f[1],
                                            df[0] * dk, df[1] * dk);
                                        if (!(zeta > csf)) {
                                                const int iel =
                                                    p.component_element[ic];
                                                const double * n = p.CSn +
                                                    (ip * p.n_elements + iel) *
                                                        p.n_energies + i1;
                                                const double * dn = p.CSn_dK +
                                                    (ip * p.n_elements + iel) *
                                                        p.n_energies + i1;
                                                const double csn = pchip(h,
                                                    n[0], n[1], dn[0] * dk,
                                                    dn[1] * dk);
                                                *weight = 1. / csn;
                                                *element = iel;
                                                *process = ip;
                                                return;
                                        }
                                }
                        /* The reference asserts (src/pumas.c:6207); not
                         * reachable for normalised fractions. */
                        *element = p.component_element[ic0 + nc - 1];
                        *process = N_DEL_PROCESSES - 1;
                        if (t.max_steps >= 0) t.max_steps = MARK_NO_TARGET;
                        return;
                } else {
                        const double Q = *weight;
                        double stot = 0.;
                        for (int ip = 0; ip < N_DEL_PROCESSES; ip++) {
                                for (int ic = ic0; ic < ic0 + nc; ic++) {
                                        const int iel = p.component_element[ic];
                                        const double d = dcs_evaluate(p, mv, t.ix, ip, iel, t.energy, Q);
                                        stot += d * p.component_fraction[ic];
                                }
                        }
                        if (stot == 0.) {
                                *weight = 0.;
                                return;
                        } else {
                                *weight = stot;
                        }

                        if (c.scattering == MODE_DISABLED) {
                                *element = p.n_elements;
                                *process = N_DEL_PROCESSES;
                                return;
                        }

                        double zeta;
                        int tries = 0;
                        for (;;) {
                                zeta = t.rng.uniform() * stot;
                                if ((zeta > 0.) && (zeta < stot)) break;
                                if (loop_stalled(t, tries, c.loop_max)) break;
                        }
                        /* The reference keeps the energy transfer and the
                         * weight in one union (struct del_info,
                         * src/pumas.c:612-626); once the weight is set to
                         * the summed DCS, the selection pass below reads
                         * that value as the energy transfer
                         * (src/pumas.c:6229-6243). Kept for parity. */
                        const double Q2 = stot;
                        double s = 0.;
                        for (int ic = ic0; ic < ic0 + nc; ic++)
                                for (int ip = 0; ip < N_DEL_PROCESSES; ip++) {
                                        const int iel = p.component_element[ic];
                                        const double si = dcs_evaluate(p, mv, t.ix,
                                            ip, iel, t.energy, Q2) *
                                            p.component_fraction[ic];
                                        s += si;
                                        if (!(zeta > s)) {
                                                *element = iel;
                                                *process = ip;
                                                return;
                                        }
                                }
                        /* The reference asserts (src/pumas.c:6264): the
                         * history is marked and ends at its next step. */
                        *element = p.component_element[ic0 + nc - 1];
                        *process = N_DEL_PROCESSES - 1;
                        if (t.max_steps >= 0) t.max_steps = MARK_NO_TARGET;
                }
        }

        /* dcs_ionisation_randomise, src/pumas.c:11185-11226. */
        PG_FN double ionisation_randomise(
            Track & t, int iel, double K, double xlow) const
        {
                const double P2 = K * (K + 2. * p.mass);
                const double E = K + p.mass;
                const double Wmax = 2. * ELECTRON_MASS * P2 /
                    (p.mass * p.mass +
                        ELECTRON_MASS * (ELECTRON_MASS + 2. * E));
                if (Wmax < p.cutoff * K) return K;
                double Wmin = (13.6 / 19.2) * p.element_I[iel];
                const double qlow = K * xlow;
                if (qlow >= Wmin) Wmin = qlow;

                const double E2 = E * E;
                const double beta2 = P2 / E2;
                const double a0 = 0.5 / E2;
                const double a1 = beta2 / Wmax;

                const double p0 = a0 * (Wmax - Wmin);
                const double p2 = 1. / Wmin - 1. / Wmax;
                double q;
                int tries = 0;
                for (;;) {
                        double z = t.rng.uniform() * (p0 + p2);
                        if (z <= p0) {
                                z /= p0;
                                q = (Wmax - Wmin) * z;
                        } else {
                                z = (z - p0) / p2;
                                q = Wmin * Wmax / (Wmax - z * (Wmax - Wmin));
                        }
                        const double r0 = a0 + 1 / (q * q);
                        const double r1 = r0 - a1 / q;
                        if (t.rng.uniform() * r0 <= r1) break;
                        if (loop_stalled(t, tries, c.loop_max)) break;
                }
                return K - q;
        }

        /* del_randomise_forward, src/pumas.c:5912-6017. Sets the process
         * and element once a target is drawn; returns 1 when a polar angle
         * is to be sampled (a non NULL polar function in the reference). */
        PG_NOINLINE int del_randomise_forward(
            Track & t, int material, int * process_out, int * element) const
        {
                if (t.energy <= p.Kt[material]) return 0;

                double weight;
                int process, iel;
                del_randomise_target(t, material, &weight, &process, &iel);
                *element = iel;
                *process_out = process;

                if (process == 3) {
                        t.energy =
                            ionisation_randomise(t, iel, t.energy, p.cutoff);
                        return 1;
                }

                double qmin, qmax;
                if (process == 2) {
                        dcs_photonuclear_range(p.mass, t.energy, &qmin, &qmax);
                } else {
                        dcs_pair_production_range(
                            mv, iel, p.mass, t.energy, &qmin, &qmax);
                        if (process == 0) qmin = 0.;
                }
                double xmin = qmin / t.energy;
                if (xmin < p.cutoff) xmin = p.cutoff;
                const double xmax = qmax / t.energy;
                if (xmin >= xmax) return 0;

                const int eidx = p.element_index[iel];
                double a0, a1, b0, b1, p0;
                const int imax = p.n_energies - 1;
                if (t.energy >= p.K[imax]) {
                        const float * envelope = p.DCS_envelope +
                            2 * ((process * p.n_elements + eidx) *
                                        p.n_energies + imax);
                        a0 = (double)envelope[0];
                        b0 = (double)envelope[1];
                        p0 = 1.;
                        a1 = b1 = 0.;
                } else {
                        const int row = index_K(p, t.ix, t.energy);
                        const float * envelope0 = p.DCS_envelope +
                            2 * ((process * p.n_elements + eidx) *
                                        p.n_energies + row);
                        const float * envelope1 = envelope0 + 2;
                        a0 = (double)envelope0[0];
                        b0 = (double)envelope0[1];
                        a1 = (double)envelope1[0];
                        b1 = (double)envelope1[1];

                        const double k0 = p.K[row];
                        const double k1 = p.K[row + 1];
                        const double h1 = (t.energy - k0) / (k1 - k0);
                        const double h0 = 1. - h1;
                        b0 *= h0;
                        b1 *= h1;

                        p0 = envelope_norm(a0, xmin, xmax) * b0;
                        const double pp1 = envelope_norm(a1, xmin, xmax) * b1;
                        p0 /= p0 + pp1;
                }

                double x = 0.;
                int i;
                for (i = 0; i < 100; i++) {
                        double u = t.rng.uniform();
                        double a;
                        if (u < p0) {
                                a = a0;
                                u = u / p0;
                        } else {
                                a = a1;
                                u = (u - p0) / (1. - p0);
                        }
                        if (a == 1) {
                                x = xmin * lm::exp(u * lm::log(xmax / xmin));
                        } else {
                                const double am = a - 1.;
                                const double pmin = lm::pow(xmin, -am);
                                const double pmax = lm::pow(xmax, -am);
                                x = lm::pow(pmin - u * (pmin - pmax), -1. / am);
                        }

                        const double penv =
                            b0 * lm::pow(x, -a0) + b1 * lm::pow(x, -a1);
                        const double d = dcs_evaluate(p, mv, t.ix, process, iel, t.energy, t.energy * x);
                        if (t.rng.uniform() * penv <= d) break;
                }
                if (i == 100) return 0;
                t.energy *= (1. - x);
                return 1;
        }

        /* del_randomise_power_law, src/pumas.c:6109-6138, for its one use
         * with alpha = BMC_ALPHA = 2 (src/pumas.c:6066-6068). The exponent
         * 1 - alpha = -1 is then a constant, and the reference's compiler
         * rewrites pow(x, -1) as 1 / x (gcc expands powers -1, 0, 1 and 2
         * exactly); so does this. */
        PG_FN void del_randomise_power_law(Track & t, double xmin,
            double xmax, double * p_r, double * p_w) const
        {
                static_assert(BMC_ALPHA == 2.0, "power law of exponent -1");
                double r = 0., w_bias;
                const double x0 = 1. / xmin;
                const double x1 = 1. / xmax;
                int tries = 0;
                for (;;) {
                        const double z = t.rng.uniform();
                        const double tmp = x0 + z * (x1 - x0);
                        r = 1. / tmp;
                        if ((r < xmin) || (r >= xmax)) {
                                if (!loop_stalled(t, tries, c.loop_max)) continue;
                                w_bias = 0.;
                                break;
                        }
                        w_bias = (x1 - x0) * r / (-tmp);
                        break;
                }
                *p_r = 1. - r;
                *p_w = w_bias;
        }

        /* del_randomise_reverse, src/pumas.c:6036-6093. Returns the process
         * index with a polar angle to sample, -1 otherwise. */
        PG_NOINLINE int del_randomise_reverse(
            Track & t, int material, int * process_out, int * element) const
        {
                const double lnq0 = -lm::log(p.cutoff);
                const double kt = p.Kt[material];
                const double kf = t.energy;
                const double pCEL =
                    (kf >= kt) ? 0. : lnq0 / (lnq0 + lm::log(kt / (kt - kf)));
                if (pCEL && (t.rng.uniform() < pCEL)) {
                        t.weight /= pCEL;
                        return -1;
                }
                const double xf = (kf >= kt * (1. - p.cutoff)) ?
                    p.cutoff :
                    1. - kf / kt;

                double xmax;
                const double m1 = p.mass - ELECTRON_MASS;
                if (t.energy < 0.5 * m1 * m1 / ELECTRON_MASS) {
                        const double m2 = p.mass + ELECTRON_MASS;
                        xmax = 2. * ELECTRON_MASS * (t.energy + 2. * p.mass) /
                            (m2 * m2);
                        if (xmax < xf)
                                return -1;
                        else if (xmax > 1.)
                                xmax = 1.;
                } else {
                        xmax = 1.;
                }

                double r, w_bias;
                del_randomise_power_law(t, xf, xmax, &r, &w_bias);
                w_bias /= r;
                t.energy /= r;

                double weight = t.energy - kf;
                int process, iel;
                del_randomise_target(t, material, &weight, &process, &iel);
                if (weight <= 0) {
                        t.weight = 0.;
                        return -1;
                }

                t.weight *= w_bias * weight /
                    (del_cross_section(p, t.ix, material, kf) * (1. - pCEL));
                *process_out = process;

                if (c.scattering == MODE_DISABLED) return -1;
                *element = iel;
                return process;
        }

        /* transport_do_del, src/pumas.c:5706-5757. */
        PG_NOINLINE void transport_do_del(Track & t, int material) const
        {
                double ki, kf;
                int process = -1, polar, iel = -1;
                if (c.direction == MODE_FORWARD) {
                        ki = t.energy;
                        polar = del_randomise_forward(
                                    t, material, &process, &iel) ?
                            process :
                            -1;
                        kf = t.energy;
                } else {
                        kf = t.energy;
                        polar = del_randomise_reverse(
                            t, material, &process, &iel);
                        ki = t.energy;
                }

                if (process < 0) {
                        if (t.weight > 0)
                                t.step_event = EVENT_NONE;
                        else
                                t.step_event = EVENT_WEIGHT;
                } else {
                        t.step_event = (process == 0) ?
                            EVENT_VERTEX_BREMSSTRAHLUNG :
                            (process == 1) ? EVENT_VERTEX_PAIR_CREATION :
                            (process == 2) ? EVENT_VERTEX_PHOTONUCLEAR :
                            (process == 3) ? EVENT_VERTEX_DELTA_RAY :
                                             EVENT_VERTEX;
                }

                if ((c.scattering == MODE_MIXED) && (polar >= 0)) {
                        if (kf > ki) {
                                const double tmp = kf;
                                kf = ki;
                                ki = tmp;
                        }
                        const double mu =
                            polar_sample(p, t.rng, polar, iel, ki, kf);
                        rotate(t, mu);
                }
        }

        /* Scattering parameters of component i of a material: CM frame,
         * spin factor and screening, as set up per element by
         * transport_do_ehs (src/pumas.c:5786-5803). */
        PG_FN void coulomb_parameters(CoulombData & d, int material, int i,
            double kinetic, double * kinetic0) const
        {
                const int ic = component0(material) + i;
                const int iel = p.component_element[ic];
                const double Z = p.element_Z[iel];
                const double A = p.element_A[iel];
                coulomb_frame_parameters(
                    Z, A, p.mass, kinetic, kinetic0, d.fCM);
                d.fspin = coulomb_spin_factor(p.mass, kinetic);
                coulomb_screening_parameters(p, iel, kinetic, *kinetic0,
                    &d.n_parameters, d.amplitude, d.screening);
        }

        /* Restricted hard cross section of component i, the remainder of
         * the per element set up (src/pumas.c:5804-5815). */
        PG_FN double coulomb_cs_hard(int material, int i, double kinetic,
            double mu0) const
        {
                CoulombData d;
                double kinetic0;
                coulomb_parameters(d, material, i, kinetic, &kinetic0);
                coulomb_pole_reduction(d);
                const int ic = component0(material) + i;
                const int iel = p.component_element[ic];
                d.normalisation = p.component_fraction[ic] /
                    coulomb_normalisation(p.element_Z[iel], p.element_A[iel],
                        p.mass, kinetic, kinetic0);
                return d.normalisation * coulomb_restricted_cs(mu0, d);
        }

        /*
         * transport_do_ehs, src/pumas.c:5768-5885. Cross sections are only
         * needed to choose among several elements: a single element is
         * chosen whatever they are (the reference's draw is still made).
         * The scattering parameters of the chosen element are computed
         * again instead of being kept for all elements; the recomputation
         * repeats the same operations on the same inputs.
         */
        /* The rejection sampling of transport_do_ehs (src/pumas.c:
         * 5834-5864) in single precision, for the FP32-bulk engine: the same
         * draws and acceptance test, with 1 + A - zeta (1 - mu0) summed as
         * (1 - zeta) + A + zeta mu0 (no cancellation in float) and the
         * nuclear form factor by its series below x = 0.5. The screening
         * set up and the CM to lab transform stay in double precision. */
        PG_FN double ehs_sample_f32(
            Track & t, const CoulombData & d, double A, double mu0) const
        {
                const int n = d.n_parameters - 1;
                const float Af = (float)A, mu0f = (float)mu0;
                const float fspin = (float)d.fspin;
                const float N = (float)screening_at(d, n);
                float amp[3], scr[3];
                for (int i = 0; i < 3; i++) {
                        amp[i] = (i < n) ? (float)d.amplitude[i] : 0.f;
                        scr[i] = (i < n) ? (float)d.screening[i] : 1.f;
                }
                float mu1 = mu0f;
                int tries = 0;
                for (;;) {
                        const float zeta = t.rng.uniformf();
                        const float tmp = (1.f - zeta) + Af + zeta * mu0f;
                        if (tmp <= 0.f) {
                                if (!loop_stalled(t, tries, c.loop_max))
                                        continue;
                                mu1 = mu0f;
                                break;
                        }
                        mu1 = (Af + mu0f) * (Af + 1.f) / tmp - Af;
                        if (mu1 < mu0f)
                                mu1 = mu0f;
                        else if (mu1 > 1.f)
                                mu1 = 1.f;

                        float ratio = 0.f;
                        for (int i = 0; i < 3; i++) {
                                if (i < n)
                                        ratio += amp[i] * (Af + mu1) /
                                            (scr[i] + mu1);
                        }
                        ratio *= ehs_form_factor_f32(mu1, N);
                        ratio *= ratio * (1.f - fspin * mu1);
                        if (t.rng.uniformf() <= ratio) break;
                        if (loop_stalled(t, tries, c.loop_max)) break;
                }
                return mu1;
        }

        /* coulomb_nuclear_form_factor (src/pumas.c:7396-7417) in single
         * precision: 3 (sin x - x cos x) / x^3 squared, by its Taylor
         * series below x = 0.5 (where the difference cancels). */
        static PG_FN float ehs_form_factor_f32(float mu, float N)
        {
                if (mu <= 0.f) return 1.f;
                if (mu > 1.f) return 0.f;
                const float x2 = 10.f * mu / N;
                float d;
                if (x2 <= 1E-04f) {
                        d = 1.f / (1.f + 0.1f * x2);
                } else if (x2 < 0.25f) {
                        d = 1.f + x2 * (-1.f / 10.f + x2 * (1.f / 280.f +
                                      x2 * (-1.f / 15120.f +
                                               x2 * (1.f / 1330560.f))));
                } else {
                        const float x = sqrtf(x2);
                        float sx, cx;
                        sincosf(x, &sx, &cx);
                        d = 3.f * (sx - x * cx) / (x2 * x);
                }
                return d * d;
        }

        PG_FN void transport_do_ehs(Track & t, int material) const
        {
                const double kinetic = t.energy;
                double mu0 = 0., invlb1 = 0.;
                /* The FP64 step tables (st) hold the MSC columns in FP64
                 * mode; FP32-bulk reads its single precision mirrors. */
                if constexpr (F32) {
                        float m, l;
                        ftable_get_msc(ft, t.fx, c.energy_loss, material,
                            (float)kinetic, &m, &l);
                        mu0 = m;
                        invlb1 = l;
                } else
                        table_get_msc(st, t.ix, c.energy_loss, material,
                            kinetic, &mu0, &invlb1);

                const int nc = p.elements_in[material];
                int ihard = nc - 1;
                if (nc == 1) {
                        (void)t.rng.uniform();
                } else {
                        double cs_hard[MAX_COMPONENTS];
                        double cs_tot = 0.;
                        for (int i = 0; i < nc; i++) {
                                cs_hard[i] =
                                    coulomb_cs_hard(material, i, kinetic, mu0);
                                cs_tot += cs_hard[i];
                        }
                        const double cc0 = t.rng.uniform() * cs_tot;
                        double cs = 0.;
                        for (int i = 0; i < nc; i++) {
                                cs += cs_hard[i];
                                if (cs >= cc0) {
                                        ihard = i;
                                        break;
                                }
                        }
                }

                CoulombData d;
                double kinetic0;
                coulomb_parameters(d, material, ihard, kinetic, &kinetic0);
                const int n = d.n_parameters - 1;
                double A = d.screening[0];
#pragma unroll
                for (int i = 1; i < 3; i++) {
                        if (i >= n) break;
                        const double Ai = d.screening[i];
                        if (Ai < A) A = Ai;
                }
                double mu1;
                if constexpr (F32)
                        mu1 = ehs_sample_f32(t, d, A, mu0);
                else {
                int tries = 0;
                for (;;) {
                        const double zeta = t.rng.uniform();
                        const double tmp = 1. + A - zeta * (1 - mu0);
                        if (tmp <= 0.) {
                                if (!loop_stalled(t, tries, c.loop_max))
                                        continue;
                                mu1 = mu0;
                                break;
                        }
                        mu1 = (A + mu0) * (A + 1.) / tmp - A;
                        if (mu1 < mu0)
                                mu1 = mu0;
                        else if (mu1 > 1.)
                                mu1 = 1;

                        double ratio = 0.;
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                                if (i >= n) break;
                                ratio += d.amplitude[i] * (A + mu1) /
                                    (d.screening[i] + mu1);
                        }
                        ratio *= coulomb_nuclear_form_factor(
                            mu1, screening_at(d, n));
                        ratio *= ratio * (1. - d.fspin * mu1);
                        if (t.rng.uniform() <= ratio) break;
                        if (loop_stalled(t, tries, c.loop_max)) break;
                }
                }

                const double gamma = d.fCM[0];
                const double tau = d.fCM[1];
                double mu;
                if (mu1 > 1E-06) {
                        const double a = gamma * (tau + 1. - 2. * mu1);
                        const double ct_h =
                            a / sqrt(4. * mu1 * (1. - mu1) + a * a);
                        mu = 0.5 * (1. - ct_h);
                } else {
                        const double dd = gamma * (1. + tau);
                        mu = mu1 / (dd * dd);
                }
                rotate(t, mu);
        }

        /*
         * The stepping transport of transport_with_stepping
         * (src/pumas.c:5194-5514) and the entry logic of
         * pumas_context_transport (src/pumas.c:3114-3251), cut at the
         * points where the reference's step loop branches on an event, so
         * that a history advances one phase at a time: INIT (entry checks,
         * closed form CSDA, stepping set up), STEP (one Monte Carlo step),
         * DEL and EHS (a pending discrete event), DONE. The operations and
         * random draws of a history are the reference's, in its order; a
         * scheduler may interleave the phases of many histories freely.
         *
         * The phase functions are inlined into the scheduler kernels so
         * that a Walker stays in registers. The rare, heavy DEL sampling is
         * out of line and works on copies of the Track and of the engine:
         * passing the address of the Walker or of the kernel's engine to a
         * call would pin them to scratch memory. The closed form CSDA runs
         * only in the entry kernel (one lane per history, not persistent)
         * and is inlined there.
         */
        template <bool CLOSED = false>
        PG_FN void walker_init(Walker & w) const
        {
                Track & t = w.t;
                w.event = EVENT_NONE;
                w.rc = RETURN_SUCCESS;
                w.media0 = w.medium = -1;
                w.phase = PHASE_DONE;
                {
                        const uint64_t id =
                            ((uint64_t)t.rng.h1 << 32) | t.rng.h0;
                        w.record = (rec.frames != nullptr) &&
                            (id >= rec.first) && (id < rec.last);
                }
                if (t.decayed) {
                        w.event = EVENT_VERTEX_DECAY;
                        w.end_medium = -1;
                        return;
                }
                {
                        const double * const u = t.dir;
                        const double norm2 =
                            u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                        if (fabs(norm2 - 1) > FLT_EPS) {
                                w.rc = RETURN_DIRECTION_ERROR;
                                w.end_medium = -1;
                                return;
                        }
                }

                int medium;
                double step_max_medium;
                int step_max_type = medium_step(t, &medium, &step_max_medium);
                w.media0 = medium;
                w.end_medium = -1;
                if (medium < 0) {
                        w.event = EVENT_MEDIUM;
                        record_frame(w, -1,
                            EVENT_MEDIUM | EVENT_START | EVENT_STOP);
                        return;
                } else if ((step_max_medium > 0.) &&
                    (step_max_type == STEP_CHECK))
                        step_max_medium += 0.5 * STEP_MIN;
                int material;
                double density;
                const double step_max_locals =
                    set_locals(t, medium, &material, &density);
                if ((step_max_locals > 0.) &&
                    (step_max_locals < step_max_medium))
                        step_max_medium = step_max_locals;
                if (density <= 0.) {
                        w.rc = RETURN_DENSITY_ERROR;
                        return;
                }

                if (c.decay == MODE_RANDOMISED) {
                        const double u = t.rng.uniform();
                        t.lifetime = t.time - p.ctau * lm::log(u);
                }

                int do_stepping = 1;
                if ((step_max_medium <= 0.) && (step_max_locals <= 0.) &&
                    (c.energy_loss <= MODE_CSDA)) {
                        if ((c.energy_loss == MODE_DISABLED) &&
                            ((c.event & EVENT_LIMIT) == 0)) {
                                w.rc = RETURN_MISSING_LIMIT;
                                return;
                        } else if ((c.scattering == MODE_DISABLED) &&
                            (c.energy_loss == MODE_CSDA)) {
                                do_stepping = 0;
                        }
                }
                /* The end medium of a history that stops before stepping
                 * is its start medium. */
                w.end_medium = medium;
                /* CLOSED: a launch for contexts and geometries where every
                 * history takes the closed form (CSDA, no scattering, a
                 * geometry without step limits), without the stepping set
                 * up below. */
                if (CLOSED || !do_stepping) {
                        record_frame(w, medium, EVENT_START);
                        w.event = transport_with_csda(t, material, density);
                        record_frame(w, medium, w.event | EVENT_STOP);
                        return;
                }

                /* transport_with_stepping set up, src/pumas.c:5210-5273. */
                record_frame(w, medium, EVENT_START);
                w.medium = medium;
                w.material = material;
                w.density = density;
                w.step_max_medium = step_max_medium;
                w.step_max_locals = step_max_locals;
                w.step_max_type = step_max_type;
                const int scheme = c.energy_loss;
                w.straight = ((c.scattering == MODE_DISABLED) &&
                                 (scheme <= MODE_MIXED) &&
                                 (step_max_locals <= 0.)) ?
                    1 :
                    0;
                t.step_event = EVENT_NONE;
                w.ti = t.time;
                w.wi = t.weight;
                w.Xi = t.grammage;
                if (scheme > MODE_DISABLED) {
                        const double ki = t.energy;
                        w.Xf = cel_grammage(p, t.ix, scheme, material, ki);
                        w.dei = 1. /
                            cel_energy_loss(p, t.ix, scheme, material, ki);
                } else {
                        w.Xf = w.dei = 0.;
                }

                if ((c.event & EVENT_LIMIT_DISTANCE) &&
                    (t.distance >= c.limit_distance)) {
                        w.event = EVENT_LIMIT_DISTANCE;
                        return;
                }
                if ((c.event & EVENT_LIMIT_GRAMMAGE) &&
                    (t.grammage >= c.limit_grammage)) {
                        w.event = EVENT_LIMIT_GRAMMAGE;
                        return;
                }
                if ((c.event & EVENT_LIMIT_TIME) &&
                    (t.time >= c.limit_time)) {
                        w.event = EVENT_LIMIT_TIME;
                        return;
                }
                if (t.weight <= 0.) {
                        w.event = EVENT_WEIGHT;
                        return;
                }

                w.grammage_max = 0.;
                t.step_event = EVENT_NONE;
                t.step_first = 1;
                t.step_X_limit = (c.event & EVENT_LIMIT_ENERGY) ?
                    cel_grammage(p, t.ix,
                        (scheme > MODE_DISABLED) ? scheme : MODE_CSDA,
                        material, c.limit_energy) :
                    0.;
                t.step_invlb1 = 0;
                transport_limit(t, material, w.Xi, w.Xf, &w.grammage_max);
                if (t.step_event) {
                        w.event = t.step_event;
                        return;
                }
                w.step_index = 1;
                w.phase = PHASE_STEP;
        }

        /* End of the step loop, src/pumas.c:5492-5513. */
        PG_FN void walker_finish(Walker & w) const
        {
                Track & t = w.t;
                if (c.direction == MODE_FORWARD) {
                        const double kinetic_min =
                            (c.event & EVENT_LIMIT_ENERGY) ? c.limit_energy :
                                                             0.;
                        if (fabs(t.energy - kinetic_min) < FLT_EPS)
                                t.energy = kinetic_min;
                } else {
                        if ((c.event & EVENT_LIMIT_ENERGY) &&
                            (fabs(t.energy - c.limit_energy) < FLT_EPS))
                                t.energy = c.limit_energy;
                }
                w.end_medium = w.medium;
                w.event = t.step_event;
                w.phase = PHASE_DONE;
                if ((t.max_steps < 0) && (w.rc == RETURN_SUCCESS))
                        w.rc = (t.max_steps == MARK_LOOP_LIMIT) ?
                            HISTORY_LOOP_LIMIT :
                            HISTORY_NO_TARGET;
                record_frame(w, w.medium, t.step_event | EVENT_STOP);
        }

        /* Common tail of an event: new initial conditions and limits,
         * src/pumas.c:5473-5488. */
        PG_FN void walker_post(Walker & w) const
        {
                Track & t = w.t;
                const int scheme = c.energy_loss;
                w.ti = t.time;
                w.wi = t.weight;
                w.Xi = t.grammage;
                if (c.energy_loss >= MODE_CSDA) {
                        const double ki = t.energy;
                        w.Xf = cel_grammage(p, t.ix, scheme, w.material, ki);
                        w.dei = 1. /
                            cel_energy_loss(p, t.ix, scheme, w.material, ki);
                }
                transport_limit(t, w.material, w.Xi, w.Xf, &w.grammage_max);
                if (t.step_event)
                        walker_finish(w);
                else
                        w.phase = PHASE_STEP;
        }

        /* One iteration of the step loop up to the event dispatch,
         * src/pumas.c:5277-5467. */
        PG_FN void walker_step(Walker & w) const
        {
                Track & t = w.t;
                const int scheme = c.energy_loss;
                if (w.step_index > 1) {
                        int unused;
                        w.step_max_type =
                            medium_step(t, &unused, &w.step_max_medium);
                        if ((w.step_max_medium > 0.) &&
                            (w.step_max_type == STEP_CHECK))
                                w.step_max_medium += 0.5 * STEP_MIN;
                }
                if (!(t.energy <= DBL_MAXV) ||
                    !(fabs(t.pos[0]) + fabs(t.pos[1]) + fabs(t.pos[2]) <=
                        DBL_MAXV)) {
                        /* A kinetic energy or a position that is no longer
                         * a finite number (from a non finite or overflowing
                         * input state) would make the history step until
                         * the cap below, as in the reference, which has no
                         * end for it: the history ends at once with a value
                         * error. */
                        w.rc = RETURN_VALUE_ERROR;
                        t.step_event = EVENT_LIMIT;
                        walker_finish(w);
                        return;
                }
                if ((int)t.steps >= t.max_steps) {
                        /* The step cap, or a history marked by a loop that
                         * gave up or by a failed target selection: it ends
                         * here with no event and a status code. */
                        if (t.max_steps >= 0) w.rc = HISTORY_STEP_LIMIT;
                        t.step_event = EVENT_NONE;
                        walker_finish(w);
                        return;
                }

                int new_medium = -1;
                step_transport(t, w.straight, w.medium, w.material, w.density,
                    w.step_max_locals, w.grammage_max, w.step_max_medium,
                    w.step_max_type, &new_medium);
                w.step_index++;
                /* Recorded steps, src/pumas.c:5296-5328: every period-th
                 * step also gets the event-time weight update. */
                const int record_step = w.record && (rec.period > 0) &&
                    (w.step_index != 1) && ((w.step_index % rec.period) == 0);
                if (!t.step_event && !record_step) return;

                if ((c.direction == MODE_BACKWARD) &&
                    (c.energy_loss >= MODE_CSDA)) {
                        const double wt = (c.decay == MODE_WEIGHTED) ?
                            lm::exp(-fabs(t.time - w.ti) / p.ctau) :
                            1.;
                        t.weight = w.wi * wt * w.dei *
                            cel_energy_loss(p, t.ix, scheme, w.material,
                                t.energy);
                } else if (c.decay == MODE_WEIGHTED) {
                        t.weight = w.wi * lm::exp(-fabs(t.time - w.ti) / p.ctau);
                }

                if (!t.step_event) {
                        record_frame(w, w.medium, EVENT_NONE);
                } else if (t.step_event & (EVENT_LIMIT | EVENT_WEIGHT |
                                       EVENT_VERTEX_DECAY)) {
                        walker_finish(w);
                } else if (t.step_event & EVENT_VERTEX_DEL) {
                        w.phase = PHASE_DEL;
                } else if (t.step_event & EVENT_VERTEX_COULOMB) {
                        w.phase = PHASE_EHS;
                } else if (t.step_event & EVENT_MEDIUM) {
                        w.medium = new_medium;
                        if ((w.medium < 0) || (c.event & EVENT_MEDIUM)) {
                                walker_finish(w);
                                return;
                        }
                        w.step_max_locals =
                            set_locals(t, w.medium, &w.material, &w.density);
                        w.straight = ((c.scattering == MODE_DISABLED) &&
                                         (scheme <= MODE_MIXED) &&
                                         (w.step_max_locals <= 0.)) ?
                            1 :
                            0;
                        const int tmp_scheme =
                            scheme > MODE_DISABLED ? scheme : MODE_CSDA;
                        t.step_X_limit = (c.event & EVENT_LIMIT_ENERGY) ?
                            cel_grammage(p, t.ix, tmp_scheme, w.material,
                                c.limit_energy) :
                            0.;
                        t.step_first = 1;
                        t.step_invlb1 = 0;
                        if (!(c.event & EVENT_MEDIUM))
                                record_frame(w, w.medium, t.step_event);
                        walker_post(w);
                }
        }

        /* The locals after a discrete vertex, src/pumas.c:5404-5415: the
         * direction, and with it the proposed step, may have changed. */
        PG_FN void walker_locals(Walker & w) const
        {
                if (w.step_max_locals > 0.)
                        w.step_max_locals = set_locals(
                            w.t, w.medium, &w.material, &w.density);
        }

        /* A pending inelastic DEL, src/pumas.c:5339-5386. */
        PG_FN void walker_del(Walker & w) const
        {
                Track & t = w.t;
                /* Pre and post DEL points when recording, src/pumas.c:
                 * 5343-5387: the pre point is the state after the DEL with
                 * its energy and direction before it. */
                const int rec_del = w.record && (rec.period > 0);
                const double ki = t.energy;
                const double ui[3] = { t.dir[0], t.dir[1], t.dir[2] };
                {
                        Track tmp = t;
                        const EngineT cold = *this;
                        cold.transport_do_del(tmp, w.material);
                        t = tmp;
                }
                if (rec_del) {
                        Walker pre = w;
                        pre.t.energy = ki;
                        pre.t.dir[0] = ui[0];
                        pre.t.dir[1] = ui[1];
                        pre.t.dir[2] = ui[2];
                        record_frame(pre, w.medium, t.step_event);
                }
                if ((c.event & t.step_event) ||
                    t.step_event == EVENT_WEIGHT) {
                        walker_finish(w);
                        return;
                }
                if (rec_del) record_frame(w, w.medium, EVENT_NONE);
                t.step_first = 1;
                t.step_invlb1 = 0;
                walker_locals(w);
                walker_post(w);
        }

        /* A pending elastic hard scattering, src/pumas.c:5393-5398. */
        PG_FN void walker_ehs(Walker & w) const
        {
                Track & t = w.t;
                {
                        /* Inlined: the EHS kernel holds nothing else. */
                        transport_do_ehs(t, w.material);
                }
                if (c.event & t.step_event) {
                        walker_finish(w);
                        return;
                }
                walker_locals(w);
                walker_post(w);
        }

        /* Run the phases of one history to the end. */
        PG_FN void walker_run(Walker & w) const
        {
                walker_init(w);
                while (w.phase != PHASE_DONE) {
                        if (w.phase == PHASE_STEP)
                                walker_step(w);
                        else if (w.phase == PHASE_DEL)
                                walker_del(w);
                        else
                                walker_ehs(w);
                }
        }
};

using Engine = EngineT<false>;
using EngineF32 = EngineT<true>;

} /* namespace pg */

#endif
