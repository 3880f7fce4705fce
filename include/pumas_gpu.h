/*
 * libpumas_hip: MI355X (gfx950) transport of batches of independent muon
 * histories, behind a plain C ABI.
 *
 * The device path replaces, for a whole batch at once, the per-history call
 *     pumas_context_transport(context, state, event, media)
 * of the reference library (include/pumas.h:1046-1048, src/pumas.c:3114-3251).
 * What the reference takes as host callbacks becomes data:
 *   - `context->medium` / `pumas_medium.locals`  -> a device geometry
 *     (struct pumas_gpu_geometry_desc), media identified by index;
 *   - `context->random`                           -> a counter based
 *     Philox4x32-10 stream keyed by (seed, history id);
 *   - `pumas_physics` (opaque, src/pumas.c:660-806) -> a PTAB table file
 *     exported from it once and uploaded once.
 * The state, mode, event, limit and return-code surface is the reference's,
 * with identical layouts and enum values, so that a `struct pumas_state`
 * array filled for the reference can be handed over unchanged.
 */
#ifndef PUMAS_GPU_H
#define PUMAS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef pumas_h
/*
 * Types shared with the reference public header. When include/pumas.h of the
 * reference is included first (it defines `pumas_h`), its own definitions are
 * used instead; layouts and values are identical.
 */

/* Transport modes, include/pumas.h:90-140. */
enum pumas_mode {
        PUMAS_MODE_DISABLED = -1,
        PUMAS_MODE_CSDA = 0,
        PUMAS_MODE_MIXED = 1,
        PUMAS_MODE_STRAGGLED = 2,
        PUMAS_MODE_WEIGHTED = 0,
        PUMAS_MODE_RANDOMISED = 1,
        PUMAS_MODE_FORWARD = 0,
        PUMAS_MODE_BACKWARD = 1
};

/* Return codes, include/pumas.h:143-196. */
enum pumas_return {
        PUMAS_RETURN_SUCCESS = 0,
        PUMAS_RETURN_ACCURACY_ERROR,
        PUMAS_RETURN_END_OF_FILE,
        PUMAS_RETURN_DECAY_ERROR,
        PUMAS_RETURN_DENSITY_ERROR,
        PUMAS_RETURN_DIRECTION_ERROR,
        PUMAS_RETURN_INCOMPLETE_FILE,
        PUMAS_RETURN_INDEX_ERROR,
        PUMAS_RETURN_PHYSICS_ERROR,
        PUMAS_RETURN_INTERNAL_ERROR,
        PUMAS_RETURN_IO_ERROR,
        PUMAS_RETURN_FORMAT_ERROR,
        PUMAS_RETURN_MEDIUM_ERROR,
        PUMAS_RETURN_MEMORY_ERROR,
        PUMAS_RETURN_MODEL_ERROR,
        PUMAS_RETURN_MISSING_LIMIT,
        PUMAS_RETURN_MISSING_RANDOM,
        PUMAS_RETURN_PATH_ERROR,
        PUMAS_RETURN_RAISE_ERROR,
        PUMAS_RETURN_TOO_LONG,
        PUMAS_RETURN_UNDEFINED_MDF,
        PUMAS_RETURN_UNKNOWN_ELEMENT,
        PUMAS_RETURN_UNKNOWN_MATERIAL,
        PUMAS_RETURN_UNKNOWN_PARTICLE,
        PUMAS_RETURN_VALUE_ERROR,
        PUMAS_N_RETURNS
};

/* Transport events, include/pumas.h:199-236. */
enum pumas_event {
        PUMAS_EVENT_NONE = 0,
        PUMAS_EVENT_LIMIT_ENERGY = 1,
        PUMAS_EVENT_LIMIT_DISTANCE = 2,
        PUMAS_EVENT_LIMIT_GRAMMAGE = 4,
        PUMAS_EVENT_LIMIT_TIME = 8,
        PUMAS_EVENT_LIMIT = 15,
        PUMAS_EVENT_MEDIUM = 16,
        PUMAS_EVENT_VERTEX_BREMSSTRAHLUNG = 32,
        PUMAS_EVENT_VERTEX_PAIR_CREATION = 64,
        PUMAS_EVENT_VERTEX_PHOTONUCLEAR = 128,
        PUMAS_EVENT_VERTEX_DELTA_RAY = 256,
        PUMAS_EVENT_VERTEX_DEL = 480,
        PUMAS_EVENT_VERTEX_COULOMB = 512,
        PUMAS_EVENT_VERTEX_DECAY = 1024,
        PUMAS_EVENT_VERTEX = 2016,
        PUMAS_EVENT_WEIGHT = 2048,
        PUMAS_EVENT_START = 4096,
        PUMAS_EVENT_STOP = 8192
};

/* Monte Carlo state, include/pumas.h:289-309 (104 bytes). */
struct pumas_state {
        double charge;
        double energy;
        double distance;
        double grammage;
        double time;
        double weight;
        double position[3];
        double direction[3];
        int decayed;
};

/* Local medium properties, include/pumas.h:314-321. */
struct pumas_locals {
        double density;
        double magnet[3];
};

/* Geometry step checking modes, include/pumas.h:461-476. */
enum pumas_step { PUMAS_STEP_CHECK = 0, PUMAS_STEP_RAW };

/* Context modes and limits, include/pumas.h:554-594. */
struct pumas_context_mode {
        enum pumas_mode energy_loss;
        enum pumas_mode decay;
        enum pumas_mode direction;
        enum pumas_mode scattering;
};

struct pumas_context_limit {
        double energy;
        double distance;
        double grammage;
        double time;
};
#endif /* pumas_h */

/*
 * The value part of `struct pumas_context` (include/pumas.h:636-672): the
 * callbacks, recorder and user data of the reference context are host
 * pointers and have no device counterpart (see the geometry and the Philox
 * stream below).
 */
struct pumas_gpu_context {
        struct pumas_context_mode mode;
        enum pumas_event event;
        struct pumas_context_limit limit;
        double accuracy;
};

/* Defaults of pumas_context_create (src/pumas.c:2582-2605) for a muon. */
void pumas_gpu_context_default(struct pumas_gpu_context * context);

/* The reference's simulation context (include/pumas.h:636-672); only its
 * value part is read here. */
struct pumas_context;

/*
 * Copies the mode, event, limit and accuracy of a reference context
 * (created by pumas_context_create, include/pumas.h:1214, or any struct of
 * its public layout) into `out`. Its callbacks, recorder and user data are
 * host pointers without a device counterpart: media come from the device
 * geometry and random numbers from the Philox stream.
 */
void pumas_gpu_context_from_reference(const struct pumas_context * context,
    struct pumas_gpu_context * out);

/* ------------------------------------------------------------------------
 * Physics tables.
 * ------------------------------------------------------------------------ */
typedef struct pumas_gpu_physics pumas_gpu_physics;
typedef struct pumas_gpu_recorder pumas_gpu_recorder;

/*
 * Upload the physics tables of a PTAB file, once. The file is written by
 * `pumas_physics_export` from a reference `struct pumas_physics` created with
 * pumas_physics_create (include/pumas.h:898-906). FP64 tables are uploaded
 * as is, together with FP32 mirrors for the FP32-bulk precision mode.
 *
 * Replaces: the `physics` member of the reference simulation context
 * (src/pumas.c:426-430), bound by pumas_context_create (include/pumas.h:1214).
 */
enum pumas_return pumas_gpu_physics_upload(
    const char * ptab_path, pumas_gpu_physics ** physics);

/* Same, from a PTAB image already in host memory. */
enum pumas_return pumas_gpu_physics_upload_memory(
    const void * ptab, size_t size, pumas_gpu_physics ** physics);

/* Release the device tables; replaces pumas_physics_destroy (pumas.h:924). */
void pumas_gpu_physics_destroy(pumas_gpu_physics ** physics);

/* Table geometry queries (the PTAB header). */
int pumas_gpu_physics_n_materials(const pumas_gpu_physics * physics);
int pumas_gpu_physics_n_energies(const pumas_gpu_physics * physics);
double pumas_gpu_physics_cutoff(const pumas_gpu_physics * physics);

/* ------------------------------------------------------------------------
 * Geometry: device media replacing the `pumas_medium_cb` callback
 * (include/pumas.h:509-511) and the `pumas_locals_cb` callback
 * (include/pumas.h:348-349). A medium is a material index and its local
 * properties, uniform or an exponential density profile; media are
 * identified by index (-1 = outside, i.e. a NULL medium).
 * ------------------------------------------------------------------------ */
enum pumas_gpu_geometry_kind {
        /* One infinite uniform medium: `medium[0]`, no boundary. */
        PUMAS_GPU_GEOMETRY_INFINITE = 0,
        /* Horizontal layers: medium i fills z in [z[i], z[i+1]). */
        PUMAS_GPU_GEOMETRY_SLABS = 1,
        /* Heightmap terrain: medium 0 is the air above the ground, medium
         * 1 + (iy * (nx - 1) + ix) the rock column of cell (ix, iy). */
        PUMAS_GPU_GEOMETRY_TERRAIN = 2,
        /* Regular voxel grid: medium 1 + voxel index; material per voxel. */
        PUMAS_GPU_GEOMETRY_VOXELS = 3
};

/* Local properties of a medium: what its `pumas_locals_cb` returns. */
enum pumas_gpu_locals {
        /* Uniform density, a null proposed step (a uniform medium). */
        PUMAS_GPU_LOCALS_UNIFORM = 0,
        /* Density profile along z, as the non uniform atmospheres of the
         * reference's examples and tests (examples/pumas/geometry.c:78-97,
         * tests/test-pumas.c:1871-1884):
         *   density = density * exp(-(z - z_ref) / scale)
         *   step    = scale / max(|u_z|, cos_min)
         * with u the direction of the state. The reference multiplies the
         * returned step by the context accuracy (src/pumas.c:5528-5552) and
         * limits every Monte Carlo step to it. cos_min = 1 proposes `scale`
         * in every direction. Allowed for the media of INFINITE and SLABS
         * geometries and for the air (media[0]) of TERRAIN. */
        PUMAS_GPU_LOCALS_EXPONENTIAL = 1
};

struct pumas_gpu_medium {
        int material;
        /* Density in kg/m^3; zero or less selects the material default.
         * EXPONENTIAL: the density at z = z_ref. */
        double density;
        /* Fields below: zero for a uniform medium. */
        enum pumas_gpu_locals locals;
        double scale;   /* EXPONENTIAL: scale length, in m (> 0) */
        double z_ref;   /* EXPONENTIAL: reference altitude, in m */
        double cos_min; /* EXPONENTIAL: in (0, 1] */
};

struct pumas_gpu_geometry_desc {
        enum pumas_gpu_geometry_kind kind;
        /* Step checking mode reported by the geometry for proposed steps. */
        enum pumas_step step_type;
        /* INFINITE: media[0]. SLABS: media[0..n_layers-1]. TERRAIN: media[0]
         * is the air, media[1] the ground material (density scaled per
         * column). VOXELS: media[0..n_materials-1], one per voxel class. */
        const struct pumas_gpu_medium * media;
        int n_media;
        /* SLABS: n_layers + 1 ascending z boundaries, in m. */
        const double * z;
        int n_layers;
        /* TERRAIN / VOXELS grid. */
        int nx, ny, nz;
        double x0, y0, z0;       /* grid origin, in m */
        double dx, dy, dz;       /* cell sizes, in m */
        const float * height;    /* TERRAIN: [ny][nx] node heights, in m */
        const float * density;   /* TERRAIN: [ny-1][nx-1] cell densities;
                                    VOXELS: [nz][ny][nx] densities, kg/m^3 */
        const uint8_t * label;   /* VOXELS: [nz][ny][nx] medium class */
        double z_top;            /* TERRAIN: top of the air layer, in m */
};

typedef struct pumas_gpu_geometry pumas_gpu_geometry;

enum pumas_return pumas_gpu_geometry_create(const pumas_gpu_physics * physics,
    const struct pumas_gpu_geometry_desc * desc, pumas_gpu_geometry ** geometry);
void pumas_gpu_geometry_destroy(pumas_gpu_geometry ** geometry);

/* ------------------------------------------------------------------------
 * Transport.
 * ------------------------------------------------------------------------ */
enum pumas_gpu_precision {
        /* FP64 everywhere, bit-level parity target with the reference. */
        PUMAS_GPU_PRECISION_FP64 = 0,
        /* FP32 table interpolation and sampling, FP64 accumulators. */
        PUMAS_GPU_PRECISION_FP32_BULK = 1
};

struct pumas_gpu_options {
        enum pumas_gpu_precision precision;
        /* Threads per block (multiple of 64, at most 512; 0 = 512). */
        int block_size;
        /* Resident blocks per CU for the persistent kernel (0 = 512 /
         * block_size: 8 wavefronts per CU). */
        int blocks_per_cu;
        /* Safety cap on Monte Carlo steps per history (0 = 10^7). */
        int max_steps;
        /* Scheduler rounds launched between two reads of the queue
         * lengths (0 = default; rounded up to an even number). */
        int rounds_per_check;
        /* Non zero: launch the rounds one by one instead of replaying them
         * as a captured HIP graph. */
        int eager;
        /* Lanes of a stepping wavefront that wait for an elastic hard
         * scattering before the wavefront runs them (0 = 16, at most 64). */
        int ehs_batch;
        /* Frame recorder, or NULL (see below). */
        const struct pumas_gpu_recorder * recorder;
        /* DELs of the last histories of a batch applied inside the
         * stepping wavefronts instead of in queued rounds: > 0 always,
         * < 0 never, 0 in the rounds whose histories fit in the resident
         * lanes at once (small requests, and the end of large batches). */
        int tail_del;
        /* Tables of the per-step lookups (the FP32 mirrors in FP32-bulk,
         * the FP64 step tables otherwise) read by the stepping kernel from
         * a copy in each block's LDS when `blocks_per_cu` copies fit in the
         * LDS of a CU: 0 or > 0 when they fit, < 0 never (the tables are
         * then read from HBM through the caches). Results do not depend on
         * it. */
        int lds_tables;
        /* Histories per wavefront of the stepping kernel (1 to 64). 0: a
         * batch that does not fill the resident wavefronts is spread over
         * all of them with fewer histories each (lower latency for small
         * requests); 64: full wavefronts on as few CUs as the batch fills.
         * Results do not depend on it. */
        int lanes_per_wave;
        /* Iterations after which a rejection or bisection loop gives up
         * and its history ends with PUMAS_GPU_HISTORY_LOOP_LIMIT (0 =
         * 2^16). The reference's loops have no limit: above its tables
         * (kinetic energies beyond 1e9 GeV, which backward discrete losses
         * can reach) its elastic hard scattering angle loop needs up to
         * ~3e7 trials per event, a wavefront held for seconds. Parity on
         * such histories needs 2^25. */
        int max_loop;
};

/*
 * Autotuner: times pumas_gpu_transport on a sample of `n` device-resident
 * states (left unchanged) for each launch shape (block size, resident
 * blocks per CU) and EHS batch size of a fixed candidate list, and writes
 * the fastest into `best` (its precision is `precision`), with its rate in
 * Monte Carlo steps per second. The sample should hold at least ~10^6
 * histories of the workload to tune (several times the 131072 resident
 * lanes), so that the tail of a launch does not dominate the timing.
 */
enum pumas_return pumas_gpu_autotune(const pumas_gpu_physics * physics,
    const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, const struct pumas_state * states,
    size_t n, uint64_t seed, enum pumas_gpu_precision precision,
    struct pumas_gpu_options * best, double * steps_per_s, void * stream);

/*
 * Per-history status codes (outputs->rc) beyond enum pumas_return, for
 * histories whose end state is not the reference's. They are not errors of
 * the call (pumas_gpu_transport still returns the first error code of
 * enum pumas_return, or success), they are counted in pumas_gpu_stats, and
 * such a history ends with event PUMAS_EVENT_NONE (a tally with an event
 * mask leaves it out).
 */
enum pumas_gpu_history_status {
        /* options->max_steps Monte Carlo steps were taken (the reference has
         * no cap). */
        PUMAS_GPU_HISTORY_STEP_LIMIT = 100,
        /* A rejection or bisection loop reached options->max_loop
         * iterations (2^16 by default): a non finite input, a sampling
         * range that rounds to nothing, or kinetic energies far above the
         * tables, where the reference would not return or needs more
         * trials. */
        PUMAS_GPU_HISTORY_LOOP_LIMIT = 101,
        /* The target selection of a discrete energy loss matched no target
         * (the reference aborts on assert(0), src/pumas.c:6207, 6264). */
        PUMAS_GPU_HISTORY_NO_TARGET = 102
};

/* Optional per-history outputs (device pointers, each may be NULL). */
struct pumas_gpu_outputs {
        enum pumas_event * events;  /* [n] end condition, as `*event` */
        int32_t * media;            /* [n][2] start / end medium index */
        uint32_t * steps;           /* [n] Monte Carlo steps taken */
        uint32_t * draws;           /* [n] random numbers drawn */
        int32_t * rc;               /* [n] per history enum pumas_return */
};

/* Aggregate statistics of one call (host struct). */
struct pumas_gpu_stats {
        uint64_t steps;         /* total Monte Carlo steps */
        uint64_t histories;     /* histories transported */
        double kernel_ms;       /* transport kernel time (HIP events) */
        double step_kernel_ms;  /* time in the stepping kernel (eager
                                   launches only; 0 with graph replays) */
        uint64_t step_launches; /* launches of the stepping kernel */
        uint64_t rounds;        /* host checks of the queue lengths */
        uint64_t lds_bytes;     /* LDS copy of the step tables per block of
                                   the stepping kernel (0: read from HBM) */
        /* Histories ended with a status of enum pumas_gpu_history_status
         * (their end state is not the reference's). */
        uint64_t step_limited;  /* PUMAS_GPU_HISTORY_STEP_LIMIT */
        uint64_t loop_limited;  /* PUMAS_GPU_HISTORY_LOOP_LIMIT */
        uint64_t no_target;     /* PUMAS_GPU_HISTORY_NO_TARGET */
        double init_kernel_ms;  /* time in the entry kernel (one lane per
                                   history; the whole transport of the
                                   closed form CSDA) */
};

/*
 * Transport `n` histories stored in device memory (`states`, AoS
 * struct pumas_state, updated in place), with the physics, context and
 * geometry given. History i draws from the Philox stream keyed by `seed` at
 * counter (first_history + i, draw index), the same stream the reference
 * receives through `context->random` in the oracle harness.
 *
 * Returns the first error raised by any history (catch semantics of
 * include/pumas.h:1146-1184); per-history codes go to outputs->rc.
 * `stream` is a hipStream_t (NULL = default stream). The call is
 * asynchronous unless `stats` is requested.
 *
 * Replaces: pumas_context_transport (include/pumas.h:1046-1048) called once
 * per history, e.g. examples/pumas/straight.c:133-179.
 */
enum pumas_return pumas_gpu_transport(const pumas_gpu_physics * physics,
    const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states,
    size_t n, uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream);

/*
 * Same as pumas_gpu_transport, allocating nothing: no reference
 * counterpart. For callers that cannot call hipMalloc (code under stream
 * capture, tight serving loops, harnesses that forbid allocations). The
 * transport's per-call device buffers come from a static array of the
 * library on each GPU (about 90 MB, held until the process ends), which
 * limits a call to PUMAS_GPU_NOALLOC_MAX_HISTORIES histories (larger `n`
 * is refused with PUMAS_RETURN_VALUE_ERROR). Launches are eager (no graph)
 * and `stats` holds counts, no timings. Calls on one GPU share its array:
 * calls that overlap on a GPU race and must be serialised by the caller.
 */
#define PUMAS_GPU_NOALLOC_MAX_HISTORIES (1 << 18)
enum pumas_return pumas_gpu_transport_noalloc(
    const pumas_gpu_physics * physics,
    const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states,
    size_t n, uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream);

/*
 * pumas_gpu_transport with the reference's own context type, the boundary
 * of include/pumas.h:1046-1048 for a batch: reads the mode, event, limit and
 * accuracy of `context` (pumas_gpu_context_from_reference) and transports
 * the device-resident `states` as pumas_gpu_transport does.
 */
enum pumas_return pumas_gpu_transport_context(
    const pumas_gpu_physics * physics, const struct pumas_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states,
    size_t n, uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream);

/*
 * Host-memory convenience: copies `states` (and requested outputs) to the
 * device, transports them and copies the results back. Output pointers in
 * `outputs` are host pointers here.
 */
enum pumas_return pumas_gpu_transport_host(const pumas_gpu_physics * physics,
    const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states,
    size_t n, uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats);

/* ------------------------------------------------------------------------
 * Frame recorder, the device counterpart of the reference's pumas_recorder
 * (include/pumas.h:378-458; src/pumas.c:2797-2851 create/clear/destroy,
 * 7738-7780 recording). Histories with ids in [first_history, first_history
 * + n_histories) record frames at the reference's points (track start and
 * stop, every `period`-th step, DEL pre and post points, medium changes;
 * only events and medium changes when `period` <= 0) into a device buffer,
 * in the order recorded; frames beyond the capacity are counted and
 * dropped. Pass the recorder in pumas_gpu_options.recorder. Replaces:
 * `context->recorder` (include/pumas.h:668).
 * ------------------------------------------------------------------------ */
struct pumas_gpu_frame {
        struct pumas_state state; /* 104 bytes */
        int32_t medium;   /* medium index, -1 = outside (a NULL medium) */
        int32_t event;    /* enum pumas_event of the frame */
        uint64_t history; /* history id */
        uint64_t reserved;
};

enum pumas_return pumas_gpu_recorder_create(size_t capacity, int period,
    uint64_t first_history, uint64_t n_histories,
    pumas_gpu_recorder ** recorder);
void pumas_gpu_recorder_destroy(pumas_gpu_recorder ** recorder);
/* Forget the recorded frames (replaces pumas_recorder_clear). */
enum pumas_return pumas_gpu_recorder_clear(pumas_gpu_recorder * recorder);
/* Copy up to `max_frames` frames to host memory; `length` receives the
 * number of frames recorded (above the capacity if some were dropped). */
enum pumas_return pumas_gpu_recorder_read(const pumas_gpu_recorder * recorder,
    struct pumas_gpu_frame * frames, size_t max_frames, uint64_t * length);

/* ------------------------------------------------------------------------
 * Tallies: flux estimators of transported histories, on the device.
 *
 * No counterpart in the reference library: its examples compute the same
 * estimator in a host loop over histories, sum of weight x flux model
 * (examples/pumas/straight.c:163-179 with examples/pumas/flux.c:39-89).
 * ------------------------------------------------------------------------ */
enum pumas_gpu_flux {
        /* Weights only (flux factor 1). */
        PUMAS_GPU_FLUX_NONE = 0,
        /* Gaisser's model, per m^2 s sr GeV (examples/pumas/flux.c:57-65). */
        PUMAS_GPU_FLUX_GAISSER = 1,
        /* GCCLY: Gaisser with Volkova's cos(theta*) and the low energy
         * correction of Guan et al. (examples/pumas/flux.c:70-89). */
        PUMAS_GPU_FLUX_GCCLY = 2
};

#define PUMAS_GPU_TALLY_MAX_BINS 256

struct pumas_gpu_tally_spec {
        enum pumas_gpu_flux flux;
        /* Non zero: multiply by the fraction of the flux with the state's
         * charge (flux.c:39-52); zero: the total flux of both charges. */
        int charge_fraction;
        /* Factor applied to the flux model (0 = 1), e.g. 1e-4 for a flux
         * per cm^2 instead of per m^2. */
        double scale;
        /* Non zero: cos(zenith) = -direction[2] of the final state (the
         * muon arrives downwards); zero: the fixed `cos_theta` below. */
        int cos_from_direction;
        double cos_theta;
        /* Histories whose event has a bit of the mask are tallied; 0, or no
         * events given, selects all. */
        unsigned int event_mask;
        /* Histograms: log bins of the final kinetic energy (GeV) and linear
         * bins of cos(zenith); 0 bins disables a histogram. */
        int n_energy_bins;
        double energy_min, energy_max;
        int n_cos_bins;
        double cos_min, cos_max;
};

struct pumas_gpu_tally_result {
        double sum;      /* sum over tallied histories of weight x flux */
        double sum2;     /* sum of squares, for the standard error */
        uint64_t count;  /* tallied histories */
        /* Host arrays given by the caller (or NULL): [2 * n_bins] holding
         * per bin the sum and the sum of squares. */
        double * energy;
        double * cos_theta;
};

/*
 * Tally `n` device-resident states (and their events, or NULL) on
 * `stream`. The sums are deterministic (fixed-order reductions). The call
 * waits for its stream.
 */
enum pumas_return pumas_gpu_tally(const struct pumas_state * states,
    const enum pumas_event * events, size_t n,
    const struct pumas_gpu_tally_spec * spec,
    struct pumas_gpu_tally_result * result, void * stream);

/* Same, with host arrays. */
enum pumas_return pumas_gpu_tally_host(const struct pumas_state * states,
    const enum pumas_event * events, size_t n,
    const struct pumas_gpu_tally_spec * spec,
    struct pumas_gpu_tally_result * result);

/* ------------------------------------------------------------------------
 * Request server: continuous batching of transport requests.
 *
 * No counterpart in the reference library (one history per call on the
 * calling thread, src/pumas.c:3114). For callers that serve many uneven
 * requests: each request is admitted as soon as a slot (a HIP stream) is
 * free and its device memory fits the HBM budget beside the requests in
 * flight, and runs concurrently with them. A request's results are the
 * same alone or in any mix.
 * ------------------------------------------------------------------------ */
typedef struct pumas_gpu_server pumas_gpu_server;

#define PUMAS_GPU_SERVER_MAX_SLOTS 32

struct pumas_gpu_server_options {
        /* Device memory for requests in flight, in bytes (0 = 200 GiB,
         * capped by the free memory at creation). */
        size_t hbm_budget;
        /* Requests in flight at once, one stream each (0 = 4). */
        int slots;
};

struct pumas_gpu_request {
        const pumas_gpu_physics * physics;
        const pumas_gpu_geometry * geometry;
        struct pumas_gpu_context context;
        enum pumas_gpu_precision precision;
        /* States, updated in place, and optional outputs: host memory, or
         * device memory when `device` is non zero. The caller keeps them
         * alive until pumas_gpu_server_wait returns for this request. */
        struct pumas_state * states;
        size_t n;
        uint64_t seed;
        uint64_t first_history;
        struct pumas_gpu_outputs outputs;
        int device;
};

struct pumas_gpu_request_status {
        enum pumas_return rc;
        double queue_ms;    /* submission to admission */
        double latency_ms;  /* submission to completion */
        double kernel_ms;   /* transport time on the device */
        uint64_t steps;
        char message[256];
};

enum pumas_return pumas_gpu_server_create(
    const struct pumas_gpu_server_options * options,
    pumas_gpu_server ** server);

/* Completes the requests submitted, then releases the server. */
void pumas_gpu_server_destroy(pumas_gpu_server ** server);

/* Queue a request; returns its ticket. The request structure is copied. */
enum pumas_return pumas_gpu_server_submit(pumas_gpu_server * server,
    const struct pumas_gpu_request * request, uint64_t * ticket);

/* Device bytes that a request of `n` histories holds while in flight, the
 * share of the HBM budget its admission needs (`device`: as the request's
 * field, states and outputs already in device memory). */
size_t pumas_gpu_server_request_bytes(size_t n, int device);

/* 1 if the request is done, 0 if pending, -1 for an unknown ticket. */
int pumas_gpu_server_poll(pumas_gpu_server * server, uint64_t ticket);

/* Wait for a request, fill its status and release its ticket; returns the
 * request's return code. */
enum pumas_return pumas_gpu_server_wait(pumas_gpu_server * server,
    uint64_t ticket, struct pumas_gpu_request_status * status);

/* Message of the last error raised on the calling thread, or "". */
const char * pumas_gpu_error_message(void);

/* Library version string. */
const char * pumas_gpu_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PUMAS_GPU_H */
