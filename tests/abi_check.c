/*
 * Test infrastructure: a plain C caller of libpumas_hip.
 *
 * It includes the reference's public header first (include/pumas.h of the
 * reference, so that the shared types are the reference's own), then
 * pumas_gpu.h, and checks at compile time that the layouts and values the
 * library relies on are the reference's. At run time it builds a
 * reference `struct pumas_context` as a user of the reference would,
 * converts it with pumas_gpu_context_from_reference and transports a batch
 * of states read from a file through the C ABI (pumas_gpu_transport_host)
 * in the 3-layer slab stack (water [0, 200) m, standard rock [200, 700) m,
 * air [700, 800) m):
 *
 *   abi_check <table.ptab> <states.bin> <out.bin> <seed> <energy_loss>
 *             <scattering> <direction> <decay> <event> <limit_energy>
 *             <limit_distance> <limit_grammage> <limit_time>
 *
 * out.bin: the final states, then per history the event (int32), the start
 * and end media (2 x int32), the steps and draws (uint32) and rc (int32).
 */
#include "pumas.h"

#include "pumas_gpu.h"

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(c) _Static_assert(c, #c)

/* struct pumas_state, include/pumas.h:289-309. */
CHECK(sizeof(struct pumas_state) == 104);
CHECK(offsetof(struct pumas_state, charge) == 0);
CHECK(offsetof(struct pumas_state, energy) == 8);
CHECK(offsetof(struct pumas_state, distance) == 16);
CHECK(offsetof(struct pumas_state, grammage) == 24);
CHECK(offsetof(struct pumas_state, time) == 32);
CHECK(offsetof(struct pumas_state, weight) == 40);
CHECK(offsetof(struct pumas_state, position) == 48);
CHECK(offsetof(struct pumas_state, direction) == 72);
CHECK(offsetof(struct pumas_state, decayed) == 96);
CHECK(sizeof(struct pumas_locals) == 32);

/* The value part of struct pumas_context (include/pumas.h:636-672) is laid
 * out as struct pumas_gpu_context, after four pointers. */
CHECK(offsetof(struct pumas_context, mode) == 4 * sizeof(void *));
CHECK(offsetof(struct pumas_context, event) -
        offsetof(struct pumas_context, mode) ==
    offsetof(struct pumas_gpu_context, event));
CHECK(offsetof(struct pumas_context, limit) -
        offsetof(struct pumas_context, mode) ==
    offsetof(struct pumas_gpu_context, limit));
CHECK(offsetof(struct pumas_context, accuracy) -
        offsetof(struct pumas_context, mode) ==
    offsetof(struct pumas_gpu_context, accuracy));
CHECK(sizeof(struct pumas_context) - offsetof(struct pumas_context, mode) ==
    sizeof(struct pumas_gpu_context));

/* Enum values used by pumas_gpu.h when the reference header is absent. */
CHECK(PUMAS_MODE_DISABLED == -1 && PUMAS_MODE_CSDA == 0 &&
    PUMAS_MODE_MIXED == 1 && PUMAS_MODE_STRAGGLED == 2);
CHECK(PUMAS_MODE_WEIGHTED == 0 && PUMAS_MODE_RANDOMISED == 1);
CHECK(PUMAS_MODE_FORWARD == 0 && PUMAS_MODE_BACKWARD == 1);
CHECK(PUMAS_EVENT_NONE == 0 && PUMAS_EVENT_LIMIT_ENERGY == 1 &&
    PUMAS_EVENT_LIMIT_DISTANCE == 2 && PUMAS_EVENT_LIMIT_GRAMMAGE == 4 &&
    PUMAS_EVENT_LIMIT_TIME == 8 && PUMAS_EVENT_LIMIT == 15 &&
    PUMAS_EVENT_MEDIUM == 16 && PUMAS_EVENT_VERTEX_BREMSSTRAHLUNG == 32 &&
    PUMAS_EVENT_VERTEX_PAIR_CREATION == 64 &&
    PUMAS_EVENT_VERTEX_PHOTONUCLEAR == 128 &&
    PUMAS_EVENT_VERTEX_DELTA_RAY == 256 && PUMAS_EVENT_VERTEX_DEL == 480 &&
    PUMAS_EVENT_VERTEX_COULOMB == 512 && PUMAS_EVENT_VERTEX_DECAY == 1024 &&
    PUMAS_EVENT_VERTEX == 2016 && PUMAS_EVENT_WEIGHT == 2048 &&
    PUMAS_EVENT_START == 4096 && PUMAS_EVENT_STOP == 8192);
CHECK(PUMAS_RETURN_SUCCESS == 0 && PUMAS_RETURN_ACCURACY_ERROR == 1 &&
    PUMAS_RETURN_DENSITY_ERROR == 4 && PUMAS_RETURN_DIRECTION_ERROR == 5 &&
    PUMAS_RETURN_MEDIUM_ERROR == 12 && PUMAS_RETURN_MISSING_LIMIT == 15 &&
    PUMAS_RETURN_VALUE_ERROR == 24 && PUMAS_N_RETURNS == 25);
CHECK(PUMAS_STEP_CHECK == 0 && PUMAS_STEP_RAW == 1);

static void * read_file(const char * path, size_t * size)
{
        FILE * f = fopen(path, "rb");
        if (f == NULL) return NULL;
        fseek(f, 0, SEEK_END);
        *size = (size_t)ftell(f);
        fseek(f, 0, SEEK_SET);
        void * data = malloc(*size > 0 ? *size : 1);
        if (fread(data, 1, *size, f) != *size) {
                free(data);
                data = NULL;
        }
        fclose(f);
        return data;
}

int main(int argc, char * argv[])
{
        if (argc != 14) {
                fprintf(stderr, "usage: see the header of abi_check.c\n");
                return 2;
        }
        size_t size = 0;
        struct pumas_state * states = read_file(argv[2], &size);
        if ((states == NULL) || (size % sizeof(*states))) {
                fprintf(stderr, "abi_check: cannot read %s\n", argv[2]);
                return 1;
        }
        const size_t n = size / sizeof(*states);
        const uint64_t seed = strtoull(argv[4], NULL, 10);

        /* A reference context, as a user of the reference fills it. */
        struct pumas_context context;
        memset(&context, 0x0, sizeof(context));
        context.mode.energy_loss = (enum pumas_mode)atoi(argv[5]);
        context.mode.scattering = (enum pumas_mode)atoi(argv[6]);
        context.mode.direction = (enum pumas_mode)atoi(argv[7]);
        context.mode.decay = (enum pumas_mode)atoi(argv[8]);
        context.event = (enum pumas_event)atoi(argv[9]);
        context.limit.energy = atof(argv[10]);
        context.limit.distance = atof(argv[11]);
        context.limit.grammage = atof(argv[12]);
        context.limit.time = atof(argv[13]);
        context.accuracy = 1E-02; /* pumas_context_create default */

        struct pumas_gpu_context values;
        pumas_gpu_context_from_reference(&context, &values);
        if ((values.mode.energy_loss != context.mode.energy_loss) ||
            (values.mode.scattering != context.mode.scattering) ||
            (values.mode.direction != context.mode.direction) ||
            (values.mode.decay != context.mode.decay) ||
            (values.event != context.event) ||
            (values.limit.distance != context.limit.distance) ||
            (values.accuracy != context.accuracy)) {
                fprintf(stderr, "abi_check: context conversion\n");
                return 1;
        }

        pumas_gpu_physics * physics = NULL;
        if (pumas_gpu_physics_upload(argv[1], &physics) !=
            PUMAS_RETURN_SUCCESS) {
                fprintf(stderr, "abi_check: %s\n", pumas_gpu_error_message());
                return 1;
        }
        const struct pumas_gpu_medium media[3] = {
                { 1, 0. }, /* water */
                { 0, 0. }, /* standard rock */
                { 2, 0. }  /* air */
        };
        const double z[4] = { 0., 200., 700., 800. };
        struct pumas_gpu_geometry_desc desc;
        memset(&desc, 0x0, sizeof(desc));
        desc.kind = PUMAS_GPU_GEOMETRY_SLABS;
        desc.step_type = PUMAS_STEP_RAW;
        desc.media = media;
        desc.n_media = 3;
        desc.z = z;
        desc.n_layers = 3;
        pumas_gpu_geometry * geometry = NULL;
        if (pumas_gpu_geometry_create(physics, &desc, &geometry) !=
            PUMAS_RETURN_SUCCESS) {
                fprintf(stderr, "abi_check: %s\n", pumas_gpu_error_message());
                return 1;
        }

        const size_t nn = n > 0 ? n : 1;
        int32_t * events = calloc(nn, sizeof(int32_t));
        int32_t * ends = calloc(2 * nn, sizeof(int32_t));
        uint32_t * steps = calloc(nn, sizeof(uint32_t));
        uint32_t * draws = calloc(nn, sizeof(uint32_t));
        int32_t * rc = calloc(nn, sizeof(int32_t));
        struct pumas_gpu_outputs outputs = { (enum pumas_event *)events, ends,
                steps, draws, rc };
        struct pumas_gpu_stats stats;
        const enum pumas_return r = pumas_gpu_transport_host(physics, &values,
            geometry, states, n, seed, 0, &outputs, NULL, &stats);
        if (r != PUMAS_RETURN_SUCCESS) {
                fprintf(stderr, "abi_check: %s\n", pumas_gpu_error_message());
                return 1;
        }

        FILE * f = fopen(argv[3], "wb");
        if (f == NULL) return 1;
        fwrite(states, sizeof(*states), n, f);
        fwrite(events, sizeof(int32_t), n, f);
        fwrite(ends, sizeof(int32_t), 2 * n, f);
        fwrite(steps, sizeof(uint32_t), n, f);
        fwrite(draws, sizeof(uint32_t), n, f);
        fwrite(rc, sizeof(int32_t), n, f);
        fclose(f);
        printf("abi_check: %zu histories, %llu steps\n", n,
            (unsigned long long)stats.steps);

        pumas_gpu_geometry_destroy(&geometry);
        pumas_gpu_physics_destroy(&physics);
        free(states);
        free(events);
        free(ends);
        free(steps);
        free(draws);
        free(rc);
        return 0;
}
