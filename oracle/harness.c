/*
 * Oracle harness (test infrastructure only): the reference PUMAS library,
 * compiled from its unmodified source file (src/pumas.c) in this translation
 * unit, driven over batches of histories.
 *
 * What the harness adds around pumas_context_transport
 * (src/pumas.c:3114-3251):
 *   - `context->random`: a Philox4x32-10 stream, key = seed, counter =
 *     (draw index / 2, stream 1, history id), uniform ((w >> 11) + 0.5) 2^-53,
 *     the stream the device engine draws from;
 *   - `context->medium` and `pumas_medium.locals`: host callbacks built on
 *     the geometry of pumas_amd/csrc/geometry.h, so that both sides locate
 *     media with the same arithmetic;
 *   - per history resets of the context's Gaussian spare
 *     (src/pumas.c:455-462), so that every history is independent of the
 *     ones transported before it on the same context;
 *   - step and draw counters per history;
 *   - one context per thread (include/pumas.h:545-549) for the CPU baseline.
 *
 * It is never linked into libpumas_hip.
 */
#include "pumas_export.c"

#include "geometry.h"
#include "pumas_gpu.h"

#include <pthread.h>
#include <setjmp.h>
#include <signal.h>
#include <stdint.h>
#include <time.h>

/*
 * Histories on which the reference aborts (an assert in
 * del_randomise_target, src/pumas.c:6264, reached in backward transport
 * when the selection pass finds no target) are caught and reported with
 * this return code instead of ending the process.
 */
#define REF_ABORTED (-1)

static __thread sigjmp_buf * ref_jump = NULL;

static void ref_abort_handler(int sig)
{
        if (ref_jump != NULL) siglongjmp(*ref_jump, 1);
        signal(sig, SIG_DFL);
        raise(sig);
}

/* ------------------------------------------------------------------------
 * Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11).
 * ------------------------------------------------------------------------ */
void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
    uint32_t out[4])
{
        uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
        uint32_t k0 = key[0], k1 = key[1];
        int r;
        for (r = 0; r < 10; r++) {
                const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
                const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
                const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
                const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
                c1 = (uint32_t)p1;
                c3 = (uint32_t)p0;
                c0 = n0;
                c2 = n2;
                k0 += 0x9E3779B9u;
                k1 += 0xBB67AE85u;
        }
        out[0] = c0;
        out[1] = c1;
        out[2] = c2;
        out[3] = c3;
}

/* Uniform number `index` of history `history` in stream `stream`. */
double ref_uniform(uint64_t seed, uint32_t stream, uint64_t history,
    uint32_t index)
{
        const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
        const uint32_t ctr[4] = { index >> 1, stream, (uint32_t)history,
                (uint32_t)(history >> 32) };
        uint32_t out[4];
        ref_philox4x32_10(ctr, key, out);
        const uint64_t w = (index & 1u) ?
            (((uint64_t)out[3] << 32) | out[2]) :
            (((uint64_t)out[1] << 32) | out[0]);
        return ((double)(w >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------
 * Physics.
 * ------------------------------------------------------------------------ */
static void silent_handler(
    enum pumas_return rc, pumas_function_t * caller, const char * message)
{
        (void)rc;
        (void)caller;
        (void)message;
}

int ref_physics_create(const char * mdf, const char * dedx, double cutoff,
    void ** physics)
{
        pumas_error_handler_set(&silent_handler);
        struct pumas_physics_settings settings = { 0 };
        settings.cutoff = cutoff;
        struct pumas_physics * p = NULL;
        const enum pumas_return rc = pumas_physics_create(
            &p, PUMAS_PARTICLE_MUON, mdf, dedx, &settings);
        *physics = p;
        return rc;
}

void ref_physics_destroy(void * physics)
{
        struct pumas_physics * p = physics;
        pumas_physics_destroy(&p);
}

int ref_physics_export(void * physics, const char * path)
{
        return pumas_physics_export(physics, path);
}

/* ------------------------------------------------------------------------
 * Geometry callbacks.
 * ------------------------------------------------------------------------ */
struct ref_medium {
        struct pumas_medium api;
        double density;
        /* Exponential profile (pumas_gpu_medium): scale length, reference
         * altitude and minimum direction cosine. */
        double scale, z_ref, cos_min;
};

struct ref_geometry {
        struct pg_geometry g;
        struct ref_medium * media;
        int64_t n_media;
        int * material;
        double * density;
        const struct pumas_gpu_medium * desc_media; /* copy of the table */
        int lazy;
};

/* Uniform medium with an explicit density (`pumas_locals_cb`). */
static double ref_locals(struct pumas_medium * medium,
    struct pumas_state * state, struct pumas_locals * locals)
{
        (void)state;
        locals->density = ((struct ref_medium *)medium)->density;
        return 0.;
}

/* Non uniform medium: an exponential density profile along z, written as
 * the atmospheres of the reference's example and test
 * (examples/pumas/geometry.c:78-97, tests/test-pumas.c:1871-1884). */
static double ref_locals_exponential(struct pumas_medium * medium,
    struct pumas_state * state, struct pumas_locals * locals)
{
        const struct ref_medium * m = (const struct ref_medium *)medium;
        locals->density =
            m->density * exp(-(state->position[2] - m->z_ref) / m->scale);
        const double uz = fabs(state->direction[2]);
        return m->scale / ((uz <= m->cos_min) ? m->cos_min : uz);
}

/* The table entry of a device medium index when it has its own locals: the
 * media of infinite and slab geometries and the air of a terrain. */
static int ref_table_entry(const struct ref_geometry * r, int64_t k)
{
        if (r->g.kind == GEO_TERRAIN) return (k == 0) ? 0 : -1;
        if (r->g.kind == GEO_VOXELS) return -1;
        return (int)k;
}

/* The reference medium of device medium index k (its material, and a
 * locals callback with its density when that is set). The density is
 * stored last with release semantics: a medium filled by two threads at
 * once gets the same values twice. */
static void ref_medium_fill(struct ref_geometry * r, int64_t k)
{
        if ((r->g.kind == GEO_VOXELS) && (k == 0)) return;
        int material;
        double density;
        geo_medium_properties(&r->g, (int)k, &material, &density);
        r->media[k].api.material = material;
        r->media[k].api.locals = (density > 0.) ? &ref_locals : NULL;
        const int e = ref_table_entry(r, k);
        if ((e >= 0) &&
            (r->desc_media[e].locals == PUMAS_GPU_LOCALS_EXPONENTIAL)) {
                r->media[k].scale = r->desc_media[e].scale;
                r->media[k].z_ref = r->desc_media[e].z_ref;
                r->media[k].cos_min = r->desc_media[e].cos_min;
                r->media[k].api.locals = &ref_locals_exponential;
        }
        __atomic_store(&r->media[k].density, &density, __ATOMIC_RELEASE);
}

static struct pumas_medium * ref_medium_at(struct ref_geometry * r, int m)
{
        if (m < 0) return NULL;
        if (r->lazy) {
                double density;
                __atomic_load(&r->media[m].density, &density, __ATOMIC_ACQUIRE);
                if (density == 0.) ref_medium_fill(r, m);
        }
        return &r->media[m].api;
}

void * ref_geometry_create(const struct pumas_gpu_geometry_desc * d)
{
        struct ref_geometry * r = calloc(1, sizeof(*r));
        struct pg_geometry * g = &r->g;
        g->kind = d->kind;
        g->step_type = d->step_type;
        g->n_media = d->n_media;
        r->material = calloc(d->n_media, sizeof(int));
        r->density = calloc(d->n_media, sizeof(double));
        struct pumas_gpu_medium * table =
            calloc(d->n_media, sizeof(struct pumas_gpu_medium));
        memcpy(table, d->media, d->n_media * sizeof(struct pumas_gpu_medium));
        r->desc_media = table;
        int i;
        for (i = 0; i < d->n_media; i++) {
                r->material[i] = d->media[i].material;
                r->density[i] = d->media[i].density;
        }
        g->material = r->material;
        g->density = r->density;
        g->n_layers = d->n_layers;
        g->z = d->z;
        g->nx = d->nx;
        g->ny = d->ny;
        g->nz = d->nz;
        g->x0 = d->x0;
        g->y0 = d->y0;
        g->z0 = d->z0;
        g->dx = d->dx;
        g->dy = d->dy;
        g->dz = d->dz;
        g->height = d->height;
        g->cell_density = d->density;
        g->label = d->label;
        g->z_top = d->z_top;

        /* One reference medium per device medium index. */
        int64_t n;
        if (d->kind == GEO_TERRAIN)
                n = 1 + (int64_t)(d->nx - 1) * (d->ny - 1);
        else if (d->kind == GEO_VOXELS)
                n = 1 + (int64_t)d->nx * d->ny * d->nz;
        else
                n = d->n_media;
        r->n_media = n;
        r->media = calloc(n, sizeof(struct ref_medium));
        /* Grid media (one per terrain column or voxel, up to 1.3e8 of them
         * for 512^3 voxels) are filled on first use (ref_medium_at); the
         * untouched part of the calloc'ed table is never mapped. */
        r->lazy = (d->kind == GEO_TERRAIN) || (d->kind == GEO_VOXELS);
        if (!r->lazy) {
                int64_t k;
                for (k = 0; k < n; k++) ref_medium_fill(r, k);
        }
        return r;
}

void ref_geometry_destroy(void * geometry)
{
        struct ref_geometry * r = geometry;
        if (r == NULL) return;
        free(r->media);
        free(r->material);
        free(r->density);
        free((void *)r->desc_media);
        free(r);
}

/* Per context data (the `user_data` extra memory of pumas_context_create). */
struct ref_user {
        const struct ref_geometry * geometry;
        uint64_t seed;
        uint64_t history;
        uint32_t draws;
        uint32_t steps;
};

static double ref_random(struct pumas_context * context)
{
        struct ref_user * u = context->user_data;
        return ref_uniform(u->seed, 1, u->history, u->draws++);
}

static enum pumas_step ref_medium_cb(struct pumas_context * context,
    struct pumas_state * state, struct pumas_medium ** medium, double * step)
{
        struct ref_user * u = context->user_data;
        struct ref_geometry * r = (struct ref_geometry *)u->geometry;
        const double sgn =
            (context->mode.direction == PUMAS_MODE_FORWARD) ? 1. : -1.;
        int m;
        double s = 0.;
        const int type = geo_locate(
            &r->g, state->position, state->direction, sgn, &m, step ? &s : 0);
        if (step != NULL) {
                *step = s;
                u->steps++;
        }
        if (medium != NULL) *medium = ref_medium_at(r, m);
        return (type == GEO_STEP_RAW) ? PUMAS_STEP_RAW : PUMAS_STEP_CHECK;
}

/* ------------------------------------------------------------------------
 * Batch transport.
 * ------------------------------------------------------------------------ */
struct ref_job {
        const struct pumas_physics * physics;
        const struct pumas_gpu_context * config;
        const struct ref_geometry * geometry;
        struct pumas_state * states;
        int64_t n;
        uint64_t seed;
        uint64_t first;
        int * events;
        int * media;
        uint32_t * steps;
        uint32_t * draws;
        int * rc;
        int thread;
        int n_threads;
        uint64_t total_steps;
        /* Recording (single thread): frames of every history, as the
         * reference's default in-memory recorder (src/pumas.c:7738-7780). */
        int rec_period;
        int64_t rec_capacity, rec_count;
        struct pumas_state * rec_states;
        int * rec_media;
        int * rec_events;
        uint64_t * rec_history;
};

static int ref_medium_index(
    const struct ref_geometry * g, const struct pumas_medium * m)
{
        return (m == NULL) ? -1 : (int)((const struct ref_medium *)m - g->media);
}

static void * ref_worker(void * arg)
{
        struct ref_job * job = arg;
        struct pumas_context * context = NULL;
        if (pumas_context_create(&context, job->physics,
                sizeof(struct ref_user)) != PUMAS_RETURN_SUCCESS)
                return NULL;
        struct simulation_context * ctx = (struct simulation_context *)context;
        context->medium = &ref_medium_cb;
        context->random = &ref_random;
        context->mode.energy_loss = job->config->mode.energy_loss;
        context->mode.decay = job->config->mode.decay;
        context->mode.direction = job->config->mode.direction;
        context->mode.scattering = job->config->mode.scattering;
        context->event = job->config->event;
        context->limit.energy = job->config->limit.energy;
        context->limit.distance = job->config->limit.distance;
        context->limit.grammage = job->config->limit.grammage;
        context->limit.time = job->config->limit.time;
        context->accuracy = job->config->accuracy;
        struct ref_user * u = context->user_data;
        u->geometry = job->geometry;
        u->seed = job->seed;
        struct pumas_recorder * recorder = NULL;
        if (job->rec_states != NULL) {
                if (pumas_recorder_create(&recorder, 0) !=
                    PUMAS_RETURN_SUCCESS) {
                        pumas_context_destroy(&context);
                        return NULL;
                }
                recorder->period = job->rec_period;
                context->recorder = recorder;
        }

        int64_t i;
        for (i = job->thread; i < job->n; i += job->n_threads) {
                u->history = job->first + (uint64_t)i;
                u->draws = 0;
                u->steps = 0;
                ctx->randn_done = 0;
                ctx->randn_next = 0.;
                ctx->lifetime = 0.;
                enum pumas_event event = PUMAS_EVENT_NONE;
                struct pumas_medium * media[2] = { NULL, NULL };
                struct pumas_state backup = job->states[i];
                sigjmp_buf jump;
                int rc;
                if (sigsetjmp(jump, 1) == 0) {
                        ref_jump = &jump;
                        rc = pumas_context_transport(
                            context, job->states + i, &event, media);
                        ref_jump = NULL;
                } else {
                        ref_jump = NULL;
                        rc = REF_ABORTED;
                        job->states[i] = backup;
                        media[0] = media[1] = NULL;
                }
                if (job->events) job->events[i] = event;
                if (job->media) {
                        int k;
                        for (k = 0; k < 2; k++)
                                job->media[2 * i + k] = (media[k] == NULL) ?
                                    -1 :
                                    (int)((struct ref_medium *)media[k] -
                                        job->geometry->media);
                }
                if (job->steps) job->steps[i] = u->steps;
                if (job->draws) job->draws[i] = u->draws;
                if (job->rc) job->rc[i] = rc;
                job->total_steps += u->steps;
                if (recorder != NULL) {
                        const struct pumas_frame * f;
                        for (f = recorder->first; f != NULL; f = f->next) {
                                const int64_t k = job->rec_count++;
                                if (k >= job->rec_capacity) continue;
                                job->rec_states[k] = f->state;
                                job->rec_media[k] =
                                    ref_medium_index(job->geometry, f->medium);
                                job->rec_events[k] = f->event;
                                job->rec_history[k] = u->history;
                        }
                        pumas_recorder_clear(recorder);
                }
        }
        if (recorder != NULL) pumas_recorder_destroy(&recorder);
        pumas_context_destroy(&context);
        return NULL;
}

/*
 * Transport `n` histories on `n_threads` threads (one context each, history
 * ids interleaved). Returns the wall time in seconds; `total_steps` receives
 * the number of Monte Carlo steps.
 */
double ref_transport(const void * physics,
    const struct pumas_gpu_context * config, const void * geometry,
    struct pumas_state * states, int64_t n, uint64_t seed, uint64_t first,
    int * events, int * media, uint32_t * steps, uint32_t * draws, int * rc,
    int n_threads, uint64_t * total_steps)
{
        if (n_threads < 1) n_threads = 1;
        struct sigaction sa, previous;
        memset(&sa, 0x0, sizeof(sa));
        sa.sa_handler = &ref_abort_handler;
        sigemptyset(&sa.sa_mask);
        sigaction(SIGABRT, &sa, &previous);
        struct ref_job * jobs = calloc(n_threads, sizeof(*jobs));
        pthread_t * threads = calloc(n_threads, sizeof(*threads));
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        int t;
        for (t = 0; t < n_threads; t++) {
                struct ref_job * j = jobs + t;
                j->physics = physics;
                j->config = config;
                j->geometry = geometry;
                j->states = states;
                j->n = n;
                j->seed = seed;
                j->first = first;
                j->events = events;
                j->media = media;
                j->steps = steps;
                j->draws = draws;
                j->rc = rc;
                j->thread = t;
                j->n_threads = n_threads;
                if (n_threads > 1)
                        pthread_create(threads + t, NULL, &ref_worker, j);
                else
                        ref_worker(j);
        }
        uint64_t total = 0;
        for (t = 0; t < n_threads; t++) {
                if (n_threads > 1) pthread_join(threads[t], NULL);
                total += jobs[t].total_steps;
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        if (total_steps) *total_steps = total;
        sigaction(SIGABRT, &previous, NULL);
        free(jobs);
        free(threads);
        return (t1.tv_sec - t0.tv_sec) + 1E-09 * (t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------------
 * Direct access to reference routines, for unit level parity checks.
 * ------------------------------------------------------------------------ */
double ref_cel(const void * physics, int which, int scheme, int material,
    double x)
{
        const struct pumas_physics * p = physics;
        switch (which) {
        case 0:
                return cel_grammage(p, NULL, scheme, material, x);
        case 1:
                return cel_proper_time(p, NULL, scheme, material, x);
        case 2:
                return cel_kinetic_energy(p, NULL, scheme, material, x);
        case 3:
                return cel_energy_loss(p, NULL, scheme, material, x);
        case 4:
                return cel_straggling(p, NULL, material, x);
        case 5:
                return cel_grammage_as_time(p, NULL, scheme, material, x);
        case 6:
                return del_cross_section(p, NULL, material, x);
        case 7:
                return del_interaction_length(p, NULL, material, x);
        case 8:
                return del_kinetic_from_interaction_length(
                    p, NULL, material, x);
        case 9:
                return ehs_interaction_length(p, NULL, scheme, material, x);
        case 10:
                return ehs_kinetic_from_interaction_length(
                    p, NULL, scheme, material, x);
        case 11:
                return coulomb_ehs_length(p, NULL, material, x);
        }
        return 0.;
}

/* dcs_evaluate (src/pumas.c:11243) for process `ip` on element `iel`. */
double ref_dcs_evaluate(const void * physics, int ip, int iel, double K,
    double q)
{
        const struct pumas_physics * p = physics;
        return dcs_evaluate(p, NULL, dcs_get(ip), p->element[iel], K, q);
}

/* Exact DCS (no tabulation) for process `ip` on element `iel`. */
double ref_dcs_exact(const void * physics, int ip, int iel, double K, double q)
{
        const struct pumas_physics * p = physics;
        return dcs_get(ip)(p, p->element[iel], K, q);
}

/* math_dilog, src/pumas.c:12609. */
double ref_dilog(double x) { return math_dilog(x); }

/*
 * Polar angle of a DEL (polar_get, src/pumas.c:11736-11976) for process `ip`
 * on element `iel`, sample i with kinetic energies ki[i] -> kf[i] drawing
 * from the transport Philox stream of history i under `seed`; the number of
 * draws of sample i goes to draws[i]. Returns 0, or -1 without a context.
 */
int ref_polar(const void * physics, int ip, int iel, const double * ki,
    const double * kf, int64_t n, uint64_t seed, double * mu,
    uint32_t * draws)
{
        struct pumas_context * context = NULL;
        if (pumas_context_create(&context, physics,
                sizeof(struct ref_user)) != PUMAS_RETURN_SUCCESS)
                return -1;
        context->random = &ref_random;
        struct ref_user * u = context->user_data;
        const struct pumas_physics * p = physics;
        polar_function_t * polar = polar_get(ip);
        u->seed = seed;
        for (int64_t i = 0; i < n; i++) {
                u->history = (uint64_t)i;
                u->draws = 0;
                mu[i] = polar(p, context, p->element[iel], ki[i], kf[i]);
                draws[i] = u->draws;
        }
        pumas_context_destroy(&context);
        return 0;
}

/*
 * Transport `n` histories on one thread with the reference's default
 * recorder at `period`, writing up to `capacity` frames (state, medium index,
 * event, history id) in the order recorded. Returns the number of frames
 * recorded (more than `capacity` if some were dropped).
 */
int64_t ref_transport_record(const void * physics,
    const struct pumas_gpu_context * config, const void * geometry,
    struct pumas_state * states, int64_t n, uint64_t seed, uint64_t first,
    int period, int64_t capacity, struct pumas_state * frame_states,
    int * frame_media, int * frame_events, uint64_t * frame_history)
{
        struct sigaction sa, previous;
        memset(&sa, 0x0, sizeof(sa));
        sa.sa_handler = &ref_abort_handler;
        sigemptyset(&sa.sa_mask);
        sigaction(SIGABRT, &sa, &previous);
        struct ref_job job;
        memset(&job, 0x0, sizeof(job));
        job.physics = physics;
        job.config = config;
        job.geometry = geometry;
        job.states = states;
        job.n = n;
        job.seed = seed;
        job.first = first;
        job.n_threads = 1;
        job.rec_period = period;
        job.rec_capacity = capacity;
        job.rec_states = frame_states;
        job.rec_media = frame_media;
        job.rec_events = frame_events;
        job.rec_history = frame_history;
        ref_worker(&job);
        sigaction(SIGABRT, &previous, NULL);
        return job.rec_count;
}
