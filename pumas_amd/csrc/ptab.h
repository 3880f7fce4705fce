/*
 * PTAB: a pointer-free, versioned container for PUMAS physics tables.
 *
 * The reference keeps its physics as one opaque heap blob whose 53 data
 * pointers are relocated on load (src/pumas.c:1554-1783 memory map,
 * src/pumas.c:2096-2122 relocation, src/pumas.c:2167-2196 dump). That blob is
 * platform dependent and holds raw pointers, so it cannot be shipped to a
 * device. PTAB stores the same arrays as named, typed, 256-byte aligned
 * sections of a little-endian file. Array contents are bit-identical to the
 * reference's in-memory tables; index layouts are the reference's
 * (src/pumas.c:4260-4752), documented per section in the exporter.
 *
 * Layout:
 *   struct ptab_header
 *   struct ptab_entry[n_entries]
 *   section data (each section starts at a multiple of PTAB_ALIGN)
 */
#ifndef PUMAS_PTAB_H
#define PUMAS_PTAB_H

#include <stdint.h>

#define PTAB_MAGIC "PUMASTAB"
#define PTAB_VERSION 1u
#define PTAB_ALIGN 256u
#define PTAB_NAME_LEN 32

enum ptab_dtype { PTAB_F64 = 0, PTAB_F32 = 1, PTAB_I32 = 2 };

/* Integer metadata slots. */
enum ptab_imeta {
        PTAB_N_ENERGIES = 0,
        PTAB_N_MATERIALS,
        PTAB_N_COMPOSITES,
        PTAB_N_ELEMENTS,
        PTAB_N_COMPONENTS,
        PTAB_MAX_COMPONENTS,
        PTAB_N_TABLE_DCS,
        PTAB_PARTICLE,
        PTAB_N_IMETA = 16
};

/* Floating point metadata slots. */
enum ptab_fmeta {
        PTAB_CTAU = 0,
        PTAB_MASS,
        PTAB_CUTOFF,
        PTAB_ELASTIC_RATIO,
        PTAB_N_FMETA = 8
};

struct ptab_header {
        char magic[8];
        uint32_t version;
        uint32_t n_entries;
        int32_t imeta[PTAB_N_IMETA];
        double fmeta[PTAB_N_FMETA];
};

struct ptab_entry {
        char name[PTAB_NAME_LEN];
        uint32_t dtype;
        uint32_t reserved;
        uint64_t count;
        uint64_t offset;
};

#endif
