/*
 * Test infrastructure: the device transport engine (pumas_amd/csrc/engine_*.h)
 * compiled for the host CPU with g++, so that its arithmetic can be compared
 * bit for bit with the oracle on the CPU. The GPU build differs from this one
 * only in its libm (OCML instead of glibc). This file is not part of the
 * library and is not reachable from its transport path.
 */
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/pumas_gpu.h"
#include "engine_transport.h"
#include "ptab.h"

using namespace pg;

static std::vector<uint8_t> g_image;
static PhysicsView g_view;
static ModelView g_mv;
static std::vector<int> g_c0;
static std::vector<int16_t> g_coarse;

static const void *sec(const char *name) {
    const ptab_header *h = (const ptab_header *)g_image.data();
    const ptab_entry *e = (const ptab_entry *)(g_image.data() + sizeof(ptab_header));
    for (uint32_t i = 0; i < h->n_entries; i++)
        if (!strncmp(e[i].name, name, PTAB_NAME_LEN)) return g_image.data() + e[i].offset;
    fprintf(stderr, "missing %s\n", name); return nullptr;
}

extern "C" int host_load(const char *path) {
    FILE *f = fopen(path, "rb"); if (!f) return -1;
    fseek(f, 0, SEEK_END); long n = ftell(f); fseek(f, 0, SEEK_SET);
    g_image.resize(n); if (fread(g_image.data(), 1, n, f) != (size_t)n) return -1; fclose(f);
    const ptab_header *h = (const ptab_header *)g_image.data();
    PhysicsView &v = g_view;
    v.n_energies = h->imeta[PTAB_N_ENERGIES]; v.n_materials = h->imeta[PTAB_N_MATERIALS];
    v.n_elements = h->imeta[PTAB_N_ELEMENTS]; v.n_components = h->imeta[PTAB_N_COMPONENTS];
    v.n_table_dcs = h->imeta[PTAB_N_TABLE_DCS]; v.max_components = h->imeta[PTAB_MAX_COMPONENTS];
    v.mass = h->fmeta[PTAB_MASS]; v.ctau = h->fmeta[PTAB_CTAU]; v.cutoff = h->fmeta[PTAB_CUTOFF];
#define D(f, n) v.f = (const double *)sec(n)
    D(K,"K"); D(K_dX,"K_dX"); D(K_dNI_in,"K_dNI_in"); D(K_dNI_el,"K_dNI_el"); D(X,"X"); D(X_dK,"X_dK");
    D(X_dT,"X_dT"); D(T,"T"); D(T_dK,"T_dK"); D(dE,"dE"); D(dE_dK,"dE_dK"); D(Omega,"Omega");
    D(Omega_dK,"Omega_dK"); D(NI_el,"NI_el"); D(NI_el_dK,"NI_el_dK"); D(NI_in,"NI_in"); D(NI_in_dK,"NI_in_dK");
    D(CS,"CS"); D(CS_dK,"CS_dK"); D(CSf,"CSf"); D(CSf_dK,"CSf_dK"); D(CSn,"CSn"); D(CSn_dK,"CSn_dK");
    D(Kt,"Kt"); D(a_max,"a_max"); D(b_max,"b_max"); D(Mu0,"Mu0"); D(Mu0_dK,"Mu0_dK"); D(Lb,"Lb");
    D(Lb_dK,"Lb_dK"); D(Ms1,"Ms1"); D(Ms1_dK,"Ms1_dK"); D(material_density,"material_density");
    D(component_fraction,"component_fraction"); D(element_Z,"element_Z"); D(element_A,"element_A");
    D(element_I,"element_I"); D(element_nscr,"element_n_screening"); D(element_amplitude,"element_amplitude");
    D(element_exponent,"element_exponent"); D(element_rn,"element_nuclear_radius"); D(element_lrad,"element_radiation_log");
    v.DCS = (const float*)sec("DCS"); v.DCS_x = (const float*)sec("DCS_x"); v.DCS_envelope = (const float*)sec("DCS_envelope");
    v.elements_in = (const int*)sec("elements_in"); v.component_element = (const int*)sec("component_element");
    v.element_index = (const int*)sec("element_index");
    g_c0.resize(v.n_materials); int acc = 0;
    for (int m = 0; m < v.n_materials; m++) { g_c0[m] = acc; acc += v.elements_in[m]; }
    v.material_component0 = g_c0.data();
    g_coarse.assign(4096, 0);
    v.k_nbins = k_coarse_build(v.K, v.n_energies, g_coarse.data(), (int)g_coarse.size(), &v.k_bin0);
    v.k_coarse = v.k_nbins ? g_coarse.data() : nullptr;
    g_mv.z13 = (const double*)sec("element_z13"); g_mv.zm13 = (const double*)sec("element_zm13");
    g_mv.z23 = (const double*)sec("element_z23"); g_mv.dn = (const double*)sec("element_dn");
    g_mv.lrad = v.element_lrad; g_mv.cst = (const double*)sec("constants");
    g_mv.gq9_x = (const double*)sec("gq9_x"); g_mv.gq9_w = (const double*)sec("gq9_w");
    g_mv.gq12_x = (const double*)sec("gq12_x"); g_mv.gq12_w = (const double*)sec("gq12_w");
    return 0;
}

static Recorder g_rec;
static int g_max_steps = 10000000;
static int g_loop_max = pg::LOOP_MAX_DEFAULT;

/* Step cap of the following host_transport calls. */
extern "C" void host_set_max_steps(int n) { g_max_steps = n; }
extern "C" void host_set_loop_max(int n) { g_loop_max = n; }

/* Frames of the following host_transport calls go to `frames` (NULL stops
 * recording), as the device recorder. */
extern "C" void host_record(Frame *frames, unsigned long long *count,
    unsigned long long capacity, int period, uint64_t first, uint64_t n) {
    g_rec.frames = frames; g_rec.count = count; g_rec.capacity = capacity;
    g_rec.period = period; g_rec.first = first; g_rec.last = first + n;
}

extern "C" int host_transport(const pumas_gpu_context *ctx, const pumas_gpu_geometry_desc *d,
    pumas_state *states, int64_t n, uint64_t seed, uint64_t first,
    int *events, int *media, uint32_t *steps, uint32_t *draws, int *rcs) {
    pg_geometry g; memset(&g, 0, sizeof(g));
    std::vector<int> mat(d->n_media); std::vector<double> den(d->n_media);
    for (int i = 0; i < d->n_media; i++) { mat[i] = d->media[i].material; den[i] = d->media[i].density; }
    g.kind = d->kind; g.step_type = d->step_type; g.n_media = d->n_media; g.material = mat.data(); g.density = den.data();
    std::vector<pg_locals> loc(d->n_media); bool uniform = true;
    for (int i = 0; i < d->n_media; i++) {
        loc[i].kind = d->media[i].locals; loc[i].scale = d->media[i].scale;
        loc[i].z_ref = d->media[i].z_ref; loc[i].cos_min = d->media[i].cos_min;
        if (loc[i].kind != GEO_LOCALS_UNIFORM) uniform = false;
    }
    g.locals = uniform ? nullptr : loc.data();
    g.n_layers = d->n_layers; g.z = d->z; g.nx = d->nx; g.ny = d->ny; g.nz = d->nz; g.x0 = d->x0; g.y0 = d->y0;
    g.z0 = d->z0; g.dx = d->dx; g.dy = d->dy; g.dz = d->dz; g.height = d->height; g.cell_density = d->density;
    g.label = d->label; g.z_top = d->z_top;
    Config c; c.energy_loss = ctx->mode.energy_loss; c.decay = ctx->mode.decay; c.direction = ctx->mode.direction;
    c.scattering = ctx->mode.scattering; c.event = ctx->event; c.limit_energy = ctx->limit.energy;
    c.limit_distance = ctx->limit.distance; c.limit_grammage = ctx->limit.grammage; c.limit_time = ctx->limit.time;
    c.accuracy = ctx->accuracy; c.loop_max = g_loop_max;
    static FloatTables g_ft; Engine eng = {g_view, g_mv, c, g, g_ft, g_view, g_rec};
    for (int64_t i = 0; i < n; i++) {
        pumas_state *s = states + i; Walker w; Track &t = w.t;
        t.charge = s->charge; t.energy = s->energy; t.distance = s->distance; t.grammage = s->grammage;
        t.time = s->time; t.weight = s->weight;
        for (int k = 0; k < 3; k++) { t.pos[k] = s->position[k]; t.dir[k] = s->direction[k]; }
        t.decayed = s->decayed; t.step_first = 0; t.step_event = t.step_foreseen = 0;
        t.step_X_limit = t.step_invlb1 = t.lifetime = 0.; t.randn_done = 0; t.randn_next = 0.;
        t.ix.reset(g_view.n_energies); t.fx.reset(g_view.n_energies); t.rng.init(seed, first + i, STREAM_TRANSPORT, 0); t.steps = 0; t.max_steps = g_max_steps;
        eng.walker_run(w);
        s->charge = t.charge; s->energy = t.energy; s->distance = t.distance; s->grammage = t.grammage;
        s->time = t.time; s->weight = t.weight;
        for (int k = 0; k < 3; k++) { s->position[k] = t.pos[k]; s->direction[k] = t.dir[k]; }
        s->decayed = t.decayed;
        events[i] = w.event; media[2*i] = w.media0; media[2*i+1] = w.end_medium; steps[i] = t.steps; draws[i] = t.rng.count; rcs[i] = w.rc;
    }
    return 0;
}

/* Number of values in [K[0], K[n - 1]) whose coarse index row differs from
 * the bisection's: every grid value, its neighbours in double precision,
 * and `n_random` log-uniform values (a fixed LCG). */
extern "C" long host_check_k_index(long n_random) {
    const PhysicsView &p = g_view;
    const int n = p.n_energies;
    long bad = 0;
    auto check = [&](double x) {
        if (!(x >= p.K[0]) || !(x < p.K[n - 1])) return;
        const int a = k_index_fast(p, x);
        const int b = table_index(p.K, n, x);
        if ((a >= 0) && (a != b)) bad++;
        if ((a < 0) && (x >= p.K[1])) bad++; /* the index must cover the grid */
    };
    for (int i = 1; i < n; i++) {
        check(p.K[i]);
        check(nextafter(p.K[i], 0.));
        check(nextafter(p.K[i], 1e300));
        if (i + 1 < n) check(0.5 * (p.K[i] + p.K[i + 1]));
    }
    unsigned long long s = 12345;
    const double l0 = log(p.K[1]), l1 = log(p.K[n - 1]);
    for (long k = 0; k < n_random; k++) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        const double u = (double)(s >> 11) / 9007199254740992.0;
        check(exp(l0 + u * (l1 - l0)));
    }
    return bad;
}

/* Function-level entry points on the loaded tables (see probe.hip for the
 * device build of the same routines): which 0 exact DCS, 1 tabulated DCS,
 * 2 dilog of a[i], 3 polar angle of a DEL with the random stream of history
 * i under `seed` (draws of sample i in draws[i]). */
extern "C" void host_functions(int which, int ip, int iel, const double *a,
    const double *b, double *out, uint32_t *draws, uint64_t seed, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        if (which == 0) out[i] = dcs_exact(g_view, g_mv, ip, iel, a[i], b[i]);
        else if (which == 1) { Cache c; c.reset(g_view.n_energies);
            out[i] = dcs_evaluate(g_view, g_mv, c, ip, iel, a[i], b[i]); }
        else if (which == 2) out[i] = dilog(a[i]);
        else { Rng rng; rng.init(seed, (uint64_t)i, STREAM_TRANSPORT, 0);
            out[i] = polar_sample(g_view, rng, ip, iel, a[i], b[i]);
            if (draws) draws[i] = rng.count; }
    }
}

/* Row lookups without a cached row: table_row_wide on table[0..n) (which
 * 0), k_row on the loaded kinetic energy grid (which 1), table_row_uniform
 * on table[0..n) (which 2). */
extern "C" void host_rows(int which, const double *table, int n,
    const double *x, int *out, int64_t m) {
    for (int64_t i = 0; i < m; i++)
        out[i] = (which == 1) ? k_row(g_view, x[i]) :
            (which == 2) ? table_row_uniform(table, n, x[i]) :
            table_row_wide(table, n, x[i]);
}

/* Copies the loaded kinetic energy grid into out[cap]; returns its size. */
extern "C" int host_k_grid(double *out, int cap) {
    for (int i = 0; (i < g_view.n_energies) && (i < cap); i++)
        out[i] = g_view.K[i];
    return g_view.n_energies;
}

/* The table accessors on the loaded tables, numbered as the oracle's
 * ref_cel: 0 cel_grammage, 1 cel_proper_time, 2 cel_kinetic_energy,
 * 3 cel_energy_loss, 4 cel_straggling, 5 cel_grammage_as_time,
 * 6 del_cross_section, 7 del_interaction_length,
 * 8 del_kinetic_from_interaction_length, 9 ehs_interaction_length,
 * 10 ehs_kinetic_from_interaction_length, 11 coulomb_ehs_length.
 * Each value uses fresh table index caches, as the oracle's calls. */
extern "C" void host_cel(int which, int scheme, int material, const double *x,
    double *out, int64_t n) {
    const PhysicsView &p = g_view;
    for (int64_t i = 0; i < n; i++) {
        Cache c; c.reset(p.n_energies);
        const double v = x[i];
        switch (which) {
        case 0: out[i] = cel_grammage(p, c, scheme, material, v); break;
        case 1: out[i] = cel_proper_time(p, c, scheme, material, v); break;
        case 2: out[i] = cel_kinetic_energy(p, c, scheme, material, v); break;
        case 3: out[i] = cel_energy_loss(p, c, scheme, material, v); break;
        case 4: out[i] = cel_straggling(p, c, material, v); break;
        case 5: out[i] = cel_grammage_as_time(p, c, scheme, material, v); break;
        case 6: out[i] = del_cross_section(p, c, material, v); break;
        case 7: out[i] = del_interaction_length(p, c, material, v); break;
        case 8: out[i] = del_kinetic_from_interaction_length(p, c, material, v); break;
        case 9: out[i] = ehs_interaction_length(p, c, scheme, material, v); break;
        case 10: out[i] = ehs_kinetic_from_interaction_length(p, c, scheme, material, v); break;
        case 11: out[i] = coulomb_ehs_length(p, c, material, v); break;
        default: out[i] = 0.;
        }
    }
}
