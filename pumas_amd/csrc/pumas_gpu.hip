/*
 * libpumas_hip: C ABI, table upload, geometry upload and the transport
 * kernels for gfx950 (a wavefront scheduler over per-phase queues, below).
 */
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/pumas_gpu.h"
#include "engine_transport.h"
#include "ptab.h"
#include "status.h"

using namespace pg;

/* Dynamic LDS of the FP32-bulk stepping kernel: a copy of the single
 * precision tables, made by each block at its start. */
extern __shared__ float pg_lds_floats[];

/* A table column in that copy, as an offset in floats. Reads compile to LDS
 * loads (ds_read) addressed from the symbol, so the table view holding these
 * offsets stays in the kernel arguments (scalar registers). */
struct LdsFloats {
        uint32_t off;

        PG_FN float operator[](int i) const { return pg_lds_floats[off + i]; }
        PG_FN LdsFloats operator+(int k) const
        {
                return LdsFloats{ off + (uint32_t)k };
        }
};

using LdsTables = FloatTablesT<LdsFloats>;

/* The same for the FP64 stepping kernel: the tables read by every step. */
extern __shared__ double pg_lds_doubles[];

struct LdsDoubles {
        uint32_t off;

        PG_FN double operator[](int i) const { return pg_lds_doubles[off + i]; }
        PG_FN double operator*() const { return pg_lds_doubles[off]; }
        PG_FN LdsDoubles operator+(int k) const
        {
                return LdsDoubles{ off + (uint32_t)k };
        }
        PG_FN bool operator==(const LdsDoubles & o) const
        {
                return off == o.off;
        }
};

/* The members of PhysicsView that the per-step lookups read (CEL range,
 * stopping power and its inverse, straggling, DEL cross section, MSC and
 * EHS parameters), as columns of the LDS copy. */
struct LdsStepTables {
        int n_energies, n_materials;
        double mass;
        LdsDoubles K, X, X_dK, K_dX, dE, dE_dK, Omega, Omega_dK;
        LdsDoubles CS, CS_dK, Mu0, Mu0_dK, Ms1, Ms1_dK, Lb, a_max, b_max;
        const int16_t * k_coarse; /* in HBM, as PhysicsView::k_coarse */
        int k_bin0, k_nbins;
        /* The stepping kernel's lookups rarely miss far: no coarse index
         * code in it. */
        static constexpr bool coarse_index = false;
};

/* ------------------------------------------------------------------------
 * Error reporting.
 * ------------------------------------------------------------------------ */
static thread_local char g_error[1024] = "";

enum pumas_return pg_fail(enum pumas_return rc, const char * fmt, ...)
{
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(g_error, sizeof(g_error), fmt, ap);
        va_end(ap);
        return rc;
}

void pg_clear_error(void) { g_error[0] = '\0'; }

extern "C" const char * pumas_gpu_error_message(void) { return g_error; }

extern "C" const char * pumas_gpu_version(void) { return "pumas_hip 0.1.0"; }

extern "C" void pumas_gpu_context_default(struct pumas_gpu_context * context)
{
        memset(context, 0x0, sizeof(*context));
        context->mode.energy_loss = PUMAS_MODE_STRAGGLED;
        context->mode.decay = PUMAS_MODE_WEIGHTED;
        context->mode.direction = PUMAS_MODE_FORWARD;
        context->mode.scattering = PUMAS_MODE_MIXED;
        context->event = PUMAS_EVENT_NONE;
        context->accuracy = 1E-02;
}

/* Device frame buffer and counter of a recorder (see pumas_gpu_recorder_*
 * below). */
struct pumas_gpu_recorder {
        Frame * frames;
        unsigned long long * count;
        unsigned long long capacity;
        int period;
        uint64_t first, n;
};

/* ------------------------------------------------------------------------
 * Physics upload.
 * ------------------------------------------------------------------------ */
struct pumas_gpu_physics {
        void * device;          /* one allocation holding every section */
        size_t size;
        PhysicsView view;
        ModelView models;
        FloatTables floats;     /* FP32 mirrors, in the same allocation */
        LdsTables lds_floats;   /* the same, as offsets in the LDS copy */
        uint32_t n_floats;      /* floats in the mirrors (multiple of 64) */
        const double * step;    /* packed copy of the FP64 step tables */
        LdsStepTables lds_step; /* their columns, as offsets in that copy */
        uint32_t n_step;        /* doubles in that copy (multiple of 8) */
        /* FP32-bulk LDS copy: the mirrors, then the FP64 tables of the range
         * to energy inversion (K, X, X_dK, K_dX, a_max, b_max; the other
         * columns of lds_range are not in it and not read). */
        LdsStepTables lds_range;
        uint32_t range_head;    /* doubles of K..K_dX at the head of step */
        uint32_t range_tail_at; /* a_max..b_max: from here to n_step */
        int particle;
};

namespace {

struct Section {
        std::string name;
        uint32_t dtype;
        uint64_t count;
        uint64_t offset;
};

/* Parse a PTAB image; sections index into `image`. */
enum pumas_return parse_ptab(const uint8_t * image, size_t size,
    ptab_header * header, std::vector<Section> & sections)
{
        if (size < sizeof(ptab_header))
                return pg_fail(PUMAS_RETURN_FORMAT_ERROR, "PTAB image too small");
        memcpy(header, image, sizeof(*header));
        if (memcmp(header->magic, PTAB_MAGIC, 8) != 0)
                return pg_fail(PUMAS_RETURN_FORMAT_ERROR, "not a PTAB image");
        if (header->version != PTAB_VERSION)
                return pg_fail(PUMAS_RETURN_FORMAT_ERROR,
                    "unsupported PTAB version %u", header->version);
        const size_t dir_end = sizeof(ptab_header) +
            (size_t)header->n_entries * sizeof(ptab_entry);
        if (dir_end > size)
                return pg_fail(PUMAS_RETURN_INCOMPLETE_FILE, "truncated PTAB");
        const ptab_entry * e =
            reinterpret_cast<const ptab_entry *>(image + sizeof(ptab_header));
        for (uint32_t i = 0; i < header->n_entries; i++) {
                const size_t width = (e[i].dtype == PTAB_F64) ? 8 : 4;
                if (e[i].offset + e[i].count * width > size)
                        return pg_fail(PUMAS_RETURN_INCOMPLETE_FILE,
                            "truncated PTAB section %s", e[i].name);
                char name[PTAB_NAME_LEN + 1] = { 0 };
                memcpy(name, e[i].name, PTAB_NAME_LEN);
                sections.push_back({ name, e[i].dtype, e[i].count, e[i].offset });
        }
        return PUMAS_RETURN_SUCCESS;
}

const Section * find(const std::vector<Section> & s, const char * name)
{
        for (const auto & x : s)
                if (x.name == name) return &x;
        return nullptr;
}

} /* namespace */

extern "C" enum pumas_return pumas_gpu_physics_upload_memory(
    const void * ptab, size_t size, pumas_gpu_physics ** physics)
{
        if (physics == NULL)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "physics is NULL");
        *physics = NULL;
        const uint8_t * image = static_cast<const uint8_t *>(ptab);
        ptab_header h;
        std::vector<Section> sections;
        enum pumas_return rc = parse_ptab(image, size, &h, sections);
        if (rc != PUMAS_RETURN_SUCCESS) return rc;

        const int nM = h.imeta[PTAB_N_MATERIALS];
        if (h.imeta[PTAB_MAX_COMPONENTS] > MAX_COMPONENTS)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "materials with more than %d elements are not supported",
                    MAX_COMPONENTS);

        /* Derived section: first component of each material. */
        const Section * sin = find(sections, "elements_in");
        if (sin == nullptr)
                return pg_fail(PUMAS_RETURN_FORMAT_ERROR, "missing elements_in");
        std::vector<int32_t> component0(nM);
        {
                const int32_t * ein =
                    reinterpret_cast<const int32_t *>(image + sin->offset);
                int acc = 0;
                for (int m = 0; m < nM; m++) {
                        component0[m] = acc;
                        acc += ein[m];
                }
        }

        /* Single precision mirrors of the tables read every step by the
         * FP32-bulk engine, appended after the derived sections. */
        static const char * const mirrored[] = { "K", "X", "X_dK", "dE",
                "dE_dK", "Omega", "Omega_dK", "CS", "CS_dK", "Mu0", "Mu0_dK",
                "Ms1", "Ms1_dK", "Lb", "a_max", "b_max" };
        const int n_mirrored = sizeof(mirrored) / sizeof(mirrored[0]);
        std::vector<float> floats;
        std::vector<size_t> float_at(n_mirrored);
        for (int k = 0; k < n_mirrored; k++) {
                const Section * sec = find(sections, mirrored[k]);
                if ((sec == nullptr) || (sec->dtype != PTAB_F64))
                        return pg_fail(PUMAS_RETURN_FORMAT_ERROR,
                            "PTAB section `%s' is missing", mirrored[k]);
                float_at[k] = floats.size();
                const double * src =
                    reinterpret_cast<const double *>(image + sec->offset);
                for (uint64_t i = 0; i < sec->count; i++)
                        floats.push_back((float)src[i]);
                while (floats.size() % 64) floats.push_back(0.f);
        }

        /* The FP64 tables read by every step, packed for the LDS copy of
         * the FP64 stepping kernel (columns padded to 64 B). */
        static const char * const stepped[] = { "K", "X", "X_dK", "K_dX",
                "dE", "dE_dK", "Omega", "Omega_dK", "CS", "CS_dK", "Mu0",
                "Mu0_dK", "Ms1", "Ms1_dK", "Lb", "a_max", "b_max" };
        const int n_stepped = sizeof(stepped) / sizeof(stepped[0]);
        std::vector<double> step;
        std::vector<size_t> step_at(n_stepped);
        for (int k = 0; k < n_stepped; k++) {
                const Section * sec = find(sections, stepped[k]);
                if ((sec == nullptr) || (sec->dtype != PTAB_F64))
                        return pg_fail(PUMAS_RETURN_FORMAT_ERROR,
                            "PTAB section `%s' is missing", stepped[k]);
                step_at[k] = step.size();
                const double * src =
                    reinterpret_cast<const double *>(image + sec->offset);
                step.insert(step.end(), src, src + sec->count);
                while (step.size() % 8) step.push_back(0.);
        }

        /* Coarse index of the kinetic energy grid (k_index_fast). */
        std::vector<int16_t> coarse(4096);
        int k_bin0 = 0;
        {
                const Section * sk = find(sections, "K");
                const double * Kh =
                    reinterpret_cast<const double *>(image + sk->offset);
                const int nb = k_coarse_build(Kh, (int)sk->count,
                    coarse.data(), (int)coarse.size(), &k_bin0);
                coarse.resize(nb);
        }

        /* Device image: the PTAB file as is, then the derived sections. */
        const size_t extra_offset = (size + 255) / 256 * 256;
        const size_t float_offset = (extra_offset +
            component0.size() * sizeof(int32_t) + 255) / 256 * 256;
        const size_t step_offset = (float_offset +
            floats.size() * sizeof(float) + 255) / 256 * 256;
        const size_t coarse_offset = (step_offset +
            step.size() * sizeof(double) + 255) / 256 * 256;
        const size_t total = coarse_offset + coarse.size() * sizeof(int16_t);
        void * device = nullptr;
        HIP_TRY(hipMalloc(&device, total));
        hipError_t e1 = hipMemcpy(device, image, size, hipMemcpyHostToDevice);
        hipError_t e2 = hipMemcpy((uint8_t *)device + extra_offset,
            component0.data(), component0.size() * sizeof(int32_t),
            hipMemcpyHostToDevice);
        hipError_t e3 = hipMemcpy((uint8_t *)device + float_offset,
            floats.data(), floats.size() * sizeof(float),
            hipMemcpyHostToDevice);
        hipError_t e4 = hipMemcpy((uint8_t *)device + step_offset,
            step.data(), step.size() * sizeof(double), hipMemcpyHostToDevice);
        hipError_t e5 = coarse.empty() ? hipSuccess :
            hipMemcpy((uint8_t *)device + coarse_offset, coarse.data(),
                coarse.size() * sizeof(int16_t), hipMemcpyHostToDevice);
        if ((e1 != hipSuccess) || (e2 != hipSuccess) || (e3 != hipSuccess) ||
            (e4 != hipSuccess) || (e5 != hipSuccess)) {
                (void)hipFree(device);
                return pg_fail(PUMAS_RETURN_INTERNAL_ERROR,
                    "could not copy the tables to the device");
        }

        pumas_gpu_physics * ph = new pumas_gpu_physics();
        ph->device = device;
        ph->size = total;
        ph->particle = h.imeta[PTAB_PARTICLE];
        PhysicsView & v = ph->view;
        v.n_energies = h.imeta[PTAB_N_ENERGIES];
        v.n_materials = nM;
        v.n_elements = h.imeta[PTAB_N_ELEMENTS];
        v.n_components = h.imeta[PTAB_N_COMPONENTS];
        v.n_table_dcs = h.imeta[PTAB_N_TABLE_DCS];
        v.max_components = h.imeta[PTAB_MAX_COMPONENTS];
        v.mass = h.fmeta[PTAB_MASS];
        v.ctau = h.fmeta[PTAB_CTAU];
        v.cutoff = h.fmeta[PTAB_CUTOFF];

        const uint8_t * base = static_cast<const uint8_t *>(device);
        bool missing = false;
        std::string missing_name;
        auto dptr = [&](const char * name) -> const void * {
                const Section * s = find(sections, name);
                if (s == nullptr) {
                        missing = true;
                        missing_name = name;
                        return nullptr;
                }
                return base + s->offset;
        };
#define D64(field, name) v.field = (const double *)dptr(name)
#define F32(field, name) v.field = (const float *)dptr(name)
#define I32(field, name) v.field = (const int *)dptr(name)
        D64(K, "K");
        D64(K_dX, "K_dX");
        D64(K_dNI_in, "K_dNI_in");
        D64(K_dNI_el, "K_dNI_el");
        D64(X, "X");
        D64(X_dK, "X_dK");
        D64(X_dT, "X_dT");
        D64(T, "T");
        D64(T_dK, "T_dK");
        D64(dE, "dE");
        D64(dE_dK, "dE_dK");
        D64(Omega, "Omega");
        D64(Omega_dK, "Omega_dK");
        D64(NI_el, "NI_el");
        D64(NI_el_dK, "NI_el_dK");
        D64(NI_in, "NI_in");
        D64(NI_in_dK, "NI_in_dK");
        D64(CS, "CS");
        D64(CS_dK, "CS_dK");
        D64(CSf, "CSf");
        D64(CSf_dK, "CSf_dK");
        D64(CSn, "CSn");
        D64(CSn_dK, "CSn_dK");
        D64(Kt, "Kt");
        D64(a_max, "a_max");
        D64(b_max, "b_max");
        D64(Mu0, "Mu0");
        D64(Mu0_dK, "Mu0_dK");
        D64(Lb, "Lb");
        D64(Lb_dK, "Lb_dK");
        D64(Ms1, "Ms1");
        D64(Ms1_dK, "Ms1_dK");
        D64(material_density, "material_density");
        D64(component_fraction, "component_fraction");
        D64(element_Z, "element_Z");
        D64(element_A, "element_A");
        D64(element_I, "element_I");
        D64(element_nscr, "element_n_screening");
        D64(element_amplitude, "element_amplitude");
        D64(element_exponent, "element_exponent");
        D64(element_rn, "element_nuclear_radius");
        D64(element_lrad, "element_radiation_log");
        F32(DCS, "DCS");
        F32(DCS_x, "DCS_x");
        F32(DCS_envelope, "DCS_envelope");
        I32(elements_in, "elements_in");
        I32(component_element, "component_element");
        I32(element_index, "element_index");
        v.material_component0 = (const int *)(base + extra_offset);
        v.k_coarse = coarse.empty() ? nullptr :
            (const int16_t *)(base + coarse_offset);
        v.k_bin0 = k_bin0;
        v.k_nbins = (int)coarse.size();

        FloatTables & ft = ph->floats;
        {
                const float * fbase =
                    (const float *)((const uint8_t *)device + float_offset);
                const float * const * fields[] = { &ft.K, &ft.X, &ft.X_dK,
                        &ft.dE, &ft.dE_dK, &ft.Omega, &ft.Omega_dK, &ft.CS,
                        &ft.CS_dK, &ft.Mu0, &ft.Mu0_dK, &ft.Ms1, &ft.Ms1_dK,
                        &ft.Lb, &ft.a_max, &ft.b_max };
                for (int k = 0; k < n_mirrored; k++)
                        *const_cast<const float **>(fields[k]) =
                            fbase + float_at[k];
                ft.n_energies = v.n_energies;
                ft.n_materials = v.n_materials;
                ft.mass = (float)v.mass;

                LdsTables & lt = ph->lds_floats;
                LdsFloats * lfields[] = { &lt.K, &lt.X, &lt.X_dK, &lt.dE,
                        &lt.dE_dK, &lt.Omega, &lt.Omega_dK, &lt.CS,
                        &lt.CS_dK, &lt.Mu0, &lt.Mu0_dK, &lt.Ms1, &lt.Ms1_dK,
                        &lt.Lb, &lt.a_max, &lt.b_max };
                static_assert(sizeof(lfields) / sizeof(lfields[0]) ==
                        sizeof(fields) / sizeof(fields[0]),
                    "one LDS column per mirrored table");
                for (int k = 0; k < n_mirrored; k++)
                        lfields[k]->off = (uint32_t)float_at[k];
                lt.n_energies = ft.n_energies;
                lt.n_materials = ft.n_materials;
                lt.mass = ft.mass;
                ph->n_floats = (uint32_t)floats.size();
        }
        {
                LdsStepTables & st = ph->lds_step;
                LdsDoubles * columns[] = { &st.K, &st.X, &st.X_dK, &st.K_dX,
                        &st.dE, &st.dE_dK, &st.Omega, &st.Omega_dK, &st.CS,
                        &st.CS_dK, &st.Mu0, &st.Mu0_dK, &st.Ms1, &st.Ms1_dK,
                        &st.Lb, &st.a_max, &st.b_max };
                static_assert(sizeof(columns) / sizeof(columns[0]) ==
                        sizeof(stepped) / sizeof(stepped[0]),
                    "one LDS column per stepped table");
                for (int k = 0; k < n_stepped; k++)
                        columns[k]->off = (uint32_t)step_at[k];
                st.n_energies = v.n_energies;
                st.n_materials = v.n_materials;
                st.mass = v.mass;
                st.k_coarse = v.k_coarse;
                st.k_bin0 = v.k_bin0;
                st.k_nbins = v.k_nbins;
                ph->step = (const double *)((const uint8_t *)device +
                    step_offset);
                ph->n_step = (uint32_t)step.size();

                /* In the FP32-bulk copy, K..K_dX (the first four stepped
                 * columns) and a_max, b_max (the last two) follow the
                 * mirrors. */
                static_assert(n_stepped == 17, "stepped columns");
                ph->range_head = (uint32_t)step_at[4];
                ph->range_tail_at = (uint32_t)step_at[15];
                const uint32_t base = ph->n_floats / 2;
                LdsStepTables & rt = ph->lds_range;
                memset(&rt, 0x0, sizeof(rt));
                rt.n_energies = st.n_energies;
                rt.n_materials = st.n_materials;
                rt.mass = st.mass;
                rt.k_coarse = st.k_coarse;
                rt.k_bin0 = st.k_bin0;
                rt.k_nbins = st.k_nbins;
                rt.K.off = base + (uint32_t)step_at[0];
                rt.X.off = base + (uint32_t)step_at[1];
                rt.X_dK.off = base + (uint32_t)step_at[2];
                rt.K_dX.off = base + (uint32_t)step_at[3];
                rt.a_max.off = base + ph->range_head;
                rt.b_max.off = base + ph->range_head +
                    (uint32_t)(step_at[16] - step_at[15]);
        }

        ModelView & mv = ph->models;
        mv.z13 = (const double *)dptr("element_z13");
        mv.zm13 = (const double *)dptr("element_zm13");
        mv.z23 = (const double *)dptr("element_z23");
        mv.dn = (const double *)dptr("element_dn");
        mv.lrad = v.element_lrad;
        mv.cst = (const double *)dptr("constants");
        mv.gq9_x = (const double *)dptr("gq9_x");
        mv.gq9_w = (const double *)dptr("gq9_w");
        mv.gq12_x = (const double *)dptr("gq12_x");
        mv.gq12_w = (const double *)dptr("gq12_w");
#undef D64
#undef F32
#undef I32
        if (missing) {
                (void)hipFree(device);
                delete ph;
                return pg_fail(PUMAS_RETURN_FORMAT_ERROR,
                    "PTAB section `%s' is missing", missing_name.c_str());
        }
        *physics = ph;
        return PUMAS_RETURN_SUCCESS;
}

extern "C" enum pumas_return pumas_gpu_physics_upload(
    const char * path, pumas_gpu_physics ** physics)
{
        if (physics != NULL) *physics = NULL;
        FILE * f = fopen(path, "rb");
        if (f == NULL)
                return pg_fail(PUMAS_RETURN_PATH_ERROR, "could not open `%s'", path);
        fseek(f, 0, SEEK_END);
        const long size = ftell(f);
        fseek(f, 0, SEEK_SET);
        std::vector<uint8_t> image(size > 0 ? size : 0);
        const size_t n = fread(image.data(), 1, image.size(), f);
        fclose(f);
        if (n != image.size())
                return pg_fail(PUMAS_RETURN_IO_ERROR, "could not read `%s'", path);
        return pumas_gpu_physics_upload_memory(image.data(), image.size(),
            physics);
}

extern "C" void pumas_gpu_physics_destroy(pumas_gpu_physics ** physics)
{
        if ((physics == NULL) || (*physics == NULL)) return;
        (void)hipFree((*physics)->device);
        delete *physics;
        *physics = NULL;
}

extern "C" int pumas_gpu_physics_n_materials(const pumas_gpu_physics * p)
{
        return p->view.n_materials;
}

extern "C" int pumas_gpu_physics_n_energies(const pumas_gpu_physics * p)
{
        return p->view.n_energies;
}

extern "C" double pumas_gpu_physics_cutoff(const pumas_gpu_physics * p)
{
        return p->view.cutoff;
}

/* ------------------------------------------------------------------------
 * Geometry.
 * ------------------------------------------------------------------------ */
struct pumas_gpu_geometry {
        pg_geometry view;        /* device pointers */
        std::vector<void *> allocations;
};

namespace {

template <typename T>
enum pumas_return upload(pumas_gpu_geometry * g, const T * src, size_t n,
    const T ** dst)
{
        void * d = nullptr;
        HIP_TRY(hipMalloc(&d, n * sizeof(T) > 0 ? n * sizeof(T) : 1));
        g->allocations.push_back(d);
        HIP_TRY(hipMemcpy(d, src, n * sizeof(T), hipMemcpyHostToDevice));
        *dst = static_cast<const T *>(d);
        return PUMAS_RETURN_SUCCESS;
}

} /* namespace */

extern "C" void pumas_gpu_geometry_destroy(pumas_gpu_geometry ** geometry)
{
        if ((geometry == NULL) || (*geometry == NULL)) return;
        for (void * p : (*geometry)->allocations) (void)hipFree(p);
        delete *geometry;
        *geometry = NULL;
}

extern "C" enum pumas_return pumas_gpu_geometry_create(
    const pumas_gpu_physics * physics,
    const struct pumas_gpu_geometry_desc * desc, pumas_gpu_geometry ** out)
{
        if (out == NULL) return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL output");
        *out = NULL;
        if ((physics == NULL) || (desc == NULL))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        if ((desc->n_media <= 0) || (desc->media == NULL))
                return pg_fail(PUMAS_RETURN_MEDIUM_ERROR, "no media");
        bool uniform = true;
        for (int i = 0; i < desc->n_media; i++) {
                const struct pumas_gpu_medium & medium = desc->media[i];
                const int m = medium.material;
                if ((m < 0) || (m >= physics->view.n_materials))
                        return pg_fail(PUMAS_RETURN_INDEX_ERROR,
                            "invalid material index [%d] for medium %d", m, i);
                if (medium.locals == PUMAS_GPU_LOCALS_UNIFORM) continue;
                if (medium.locals != PUMAS_GPU_LOCALS_EXPONENTIAL)
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "invalid locals kind %d for medium %d",
                            (int)medium.locals, i);
                if (!(medium.scale > 0.) || !(medium.scale < INFINITY) ||
                    !(medium.cos_min > 0.) || !(medium.cos_min <= 1.) ||
                    !(fabs(medium.z_ref) < INFINITY))
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "medium %d: an exponential profile needs a finite "
                            "scale > 0, a finite z_ref and cos_min in (0, 1]",
                            i);
                const bool allowed =
                    (desc->kind == PUMAS_GPU_GEOMETRY_INFINITE) ||
                    (desc->kind == PUMAS_GPU_GEOMETRY_SLABS) ||
                    ((desc->kind == PUMAS_GPU_GEOMETRY_TERRAIN) && (i == 0));
                if (!allowed)
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "medium %d: non uniform locals apply to the media "
                            "of infinite and slab geometries and to the air "
                            "of a terrain", i);
                uniform = false;
        }

        pumas_gpu_geometry * g = new pumas_gpu_geometry();
        pg_geometry & v = g->view;
        memset(&v, 0x0, sizeof(v));
        v.kind = desc->kind;
        v.step_type = desc->step_type;
        v.n_media = desc->n_media;
        std::vector<int> material(desc->n_media);
        std::vector<double> density(desc->n_media);
        for (int i = 0; i < desc->n_media; i++) {
                material[i] = desc->media[i].material;
                density[i] = desc->media[i].density;
        }
        enum pumas_return rc;
#define UP(src, n, dst)                                                        \
        do {                                                                   \
                rc = upload(g, src, n, dst);                                   \
                if (rc != PUMAS_RETURN_SUCCESS) {                              \
                        pumas_gpu_geometry_destroy(&g);                        \
                        return rc;                                             \
                }                                                              \
        } while (0)
        UP(material.data(), material.size(), &v.material);
        UP(density.data(), density.size(), &v.density);
        if (!uniform) {
                std::vector<pg_locals> locals(desc->n_media);
                for (int i = 0; i < desc->n_media; i++) {
                        const struct pumas_gpu_medium & m = desc->media[i];
                        locals[i].kind = (int)m.locals;
                        locals[i].scale = m.scale;
                        locals[i].z_ref = m.z_ref;
                        locals[i].cos_min = m.cos_min;
                }
                UP(locals.data(), locals.size(), &v.locals);
        }

        switch (desc->kind) {
        case PUMAS_GPU_GEOMETRY_INFINITE:
                break;
        case PUMAS_GPU_GEOMETRY_SLABS:
                if ((desc->n_layers <= 0) || (desc->z == NULL) ||
                    (desc->n_media < desc->n_layers)) {
                        pumas_gpu_geometry_destroy(&g);
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "slabs need n_layers + 1 boundaries and media");
                }
                for (int i = 0; i < desc->n_layers; i++)
                        if (!(desc->z[i] < desc->z[i + 1])) {
                                pumas_gpu_geometry_destroy(&g);
                                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                                    "slab boundaries must increase");
                        }
                v.n_layers = desc->n_layers;
                UP(desc->z, (size_t)desc->n_layers + 1, &v.z);
                break;
        case PUMAS_GPU_GEOMETRY_TERRAIN: {
                if ((desc->nx < 2) || (desc->ny < 2) || (desc->height == NULL) ||
                    (desc->density == NULL) || (desc->n_media < 2) ||
                    (desc->dx <= 0.) || (desc->dy <= 0.)) {
                        pumas_gpu_geometry_destroy(&g);
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "invalid terrain description");
                }
                v.nx = desc->nx;
                v.ny = desc->ny;
                v.x0 = desc->x0;
                v.y0 = desc->y0;
                v.z0 = desc->z0;
                v.dx = desc->dx;
                v.dy = desc->dy;
                v.z_top = desc->z_top;
                UP(desc->height, (size_t)desc->nx * desc->ny, &v.height);
                UP(desc->density, (size_t)(desc->nx - 1) * (desc->ny - 1),
                    &v.cell_density);
                break;
        }
        case PUMAS_GPU_GEOMETRY_VOXELS: {
                if ((desc->nx < 1) || (desc->ny < 1) || (desc->nz < 1) ||
                    (desc->label == NULL) || (desc->density == NULL) ||
                    (desc->dx <= 0.) || (desc->dy <= 0.) || (desc->dz <= 0.)) {
                        pumas_gpu_geometry_destroy(&g);
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "invalid voxel description");
                }
                const size_t nv = (size_t)desc->nx * desc->ny * desc->nz;
                if (nv >= (size_t)0x7fffffff) {
                        pumas_gpu_geometry_destroy(&g);
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "too many voxels");
                }
                for (size_t i = 0; i < nv; i++)
                        if (desc->label[i] >= desc->n_media) {
                                pumas_gpu_geometry_destroy(&g);
                                return pg_fail(PUMAS_RETURN_INDEX_ERROR,
                                    "voxel %zu has an invalid label", i);
                        }
                v.nx = desc->nx;
                v.ny = desc->ny;
                v.nz = desc->nz;
                v.x0 = desc->x0;
                v.y0 = desc->y0;
                v.z0 = desc->z0;
                v.dx = desc->dx;
                v.dy = desc->dy;
                v.dz = desc->dz;
                UP(desc->label, nv, &v.label);
                UP(desc->density, nv, &v.cell_density);
                break;
        }
        default:
                pumas_gpu_geometry_destroy(&g);
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "unknown geometry kind");
        }
#undef UP
        *out = g;
        return PUMAS_RETURN_SUCCESS;
}

/* ------------------------------------------------------------------------
 * Transport kernels: a wavefront scheduler over per-phase queues.
 *
 * A batch runs in rounds. k_init loads every history and runs the entry
 * logic of pumas_context_transport; histories that need stepping get a
 * record (their Walker, in HBM) and a slot in the step queue. Each round
 * then runs two kernels:
 *   k_step  persistent; a lane takes a slot from the step queue and steps
 *           that history in registers, elastic hard scatterings (EHS)
 *           included, until it ends or reaches an inelastic DEL vertex,
 *           then writes it back and appends its slot to the DEL queue (one
 *           atomic per wavefront); free lanes refill by ballot;
 *   k_del   one lane per queued DEL, all lanes on the same sampling code;
 *           the histories go back to the next round's step queue.
 * EHS vertices (~16% of the steps of a muon in rock) are batched inside the
 * stepping wavefront: lanes that reach one wait until enough others do.
 * DEL vertices (~0.2% of the steps at a 5% cutoff) leave the wavefront: the
 * DEL samplers are the largest code of the path (exact DCS models, polar
 * angles, rejection loops) and would hold the stepping kernel's registers
 * for every step. A history's operations are the reference's, in its
 * order; only the interleaving across histories differs, and the random
 * stream of a history is keyed by its id, so results do not depend on the
 * schedule.
 * ------------------------------------------------------------------------ */
enum {
        C_TAKEN = 0, /* step queue entries taken in this round */
        C_STEPS,     /* Monte Carlo steps of finished histories */
        C_ERROR,     /* first error code raised by a history */
        C_STEP_N0,   /* step queue lengths (ping-pong) */
        C_STEP_N1,
        C_DEL_N,     /* DEL queue length */
        C_STATUS,    /* histories ended with HISTORY_STEP_LIMIT, then */
        C_STATUS_LOOP,   /* HISTORY_LOOP_LIMIT and */
        C_STATUS_TARGET, /* HISTORY_NO_TARGET */
        C_COUNT = 12
};

struct KernelArgs {
        PhysicsView p;
        ModelView mv;
        FloatTables ft;
        LdsTables ftl;          /* ft in the stepping kernel's LDS copy */
        LdsStepTables stl;      /* FP64 step tables in the LDS copy */
        /* The LDS copy of the stepping kernel: up to three pieces, each
         * `n4` 16 B words from `src` to LDS word `dst`. */
        struct {
                const float4 * src;
                uint32_t dst, n4;
        } lds_piece[3];
        int n_lds_pieces;
        Config c;
        pg_geometry g;
        struct pumas_state * states;
        uint64_t n;
        uint64_t seed;
        uint64_t first;
        enum pumas_event * events;
        int32_t * media;
        uint32_t * steps;
        uint32_t * draws;
        int32_t * rc;
        int max_steps;
        int ehs_batch;
        uint32_t tail_del_lanes; /* resident lanes of the stepping launch */
        int lanes_per_wave;      /* lanes of a wavefront taking histories */
        int tail_adaptive;       /* in-wavefront DELs chosen per round */
        Recorder rec;
        unsigned long long * counters;
        Walker * records;     /* [n] histories in flight */
        uint32_t * q_step[2]; /* [n] step queues (ping-pong) */
        uint32_t * q_del;     /* [n] */
};

/* The arguments are read in place from the kernarg segment (scalar loads of
 * uniform data): naming the by-value parameter's members by reference would
 * make a per-lane scratch copy of it. */
#define KERNEL_ARGS(args)                                                      \
        (void)(args);                                                          \
        const KernelArgs & a =                                                 \
            *(const KernelArgs *)__builtin_amdgcn_kernarg_segment_ptr()

/* Load history i, whose state is at `s`, into a walker. */
__device__ __forceinline__ void walker_load(const KernelArgs & a, uint64_t i,
    Walker & w, const struct pumas_state * s)
{
        Track & t = w.t;
        t.charge = s->charge;
        t.energy = s->energy;
        t.distance = s->distance;
        t.grammage = s->grammage;
        t.time = s->time;
        t.weight = s->weight;
        t.pos[0] = s->position[0];
        t.pos[1] = s->position[1];
        t.pos[2] = s->position[2];
        t.dir[0] = s->direction[0];
        t.dir[1] = s->direction[1];
        t.dir[2] = s->direction[2];
        t.decayed = s->decayed;
        t.step_first = 0;
        t.step_event = t.step_foreseen = 0;
        t.step_X_limit = t.step_invlb1 = t.lifetime = 0.;
        t.randn_done = 0;
        t.randn_next = 0.;
        t.ix.reset(a.p.n_energies);
        t.fx.reset(a.p.n_energies);
        t.rng.init(a.seed, a.first + i, STREAM_TRANSPORT, 0);
        t.steps = 0;
        t.max_steps = a.max_steps;
}

/* Store the end state of history i at `s` and its outputs. */
__device__ __forceinline__ void walker_store(const KernelArgs & a, uint64_t i,
    const Walker & w, struct pumas_state * s)
{
        const Track & t = w.t;
        s->charge = t.charge;
        s->energy = t.energy;
        s->distance = t.distance;
        s->grammage = t.grammage;
        s->time = t.time;
        s->weight = t.weight;
        s->position[0] = t.pos[0];
        s->position[1] = t.pos[1];
        s->position[2] = t.pos[2];
        s->direction[0] = t.dir[0];
        s->direction[1] = t.dir[1];
        s->direction[2] = t.dir[2];
        s->decayed = t.decayed;
        if (a.events) a.events[i] = (enum pumas_event)w.event;
        if (a.media) {
                a.media[2 * i] = w.media0;
                a.media[2 * i + 1] = w.end_medium;
        }
        if (a.steps) a.steps[i] = t.steps;
        if (a.draws) a.draws[i] = t.rng.count;
        if (a.rc) a.rc[i] = w.rc;
        if (w.rc >= HISTORY_STEP_LIMIT)
                atomicAdd(a.counters + C_STATUS + (w.rc - HISTORY_STEP_LIMIT),
                    1ull);
        else if (w.rc != RETURN_SUCCESS)
                atomicCAS(a.counters + C_ERROR, 0ull,
                    (unsigned long long)w.rc);
}

__device__ __forceinline__ uint64_t lanes_below()
{
        const int lane = threadIdx.x & 63;
        return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

/* Append `slot` to `queue` for the lanes where `pred` holds: one atomic per
 * wavefront, positions by prefix count. Called by all lanes of the wave. */
__device__ __forceinline__ void queue_append(unsigned long long * count,
    uint32_t * queue, bool pred, uint32_t slot)
{
        const uint64_t m = __ballot(pred);
        if (!m) return;
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned long long base = 0;
        if ((int)(threadIdx.x & 63) == leader)
                base = atomicAdd(count, (unsigned long long)__popcll(m));
        base = __shfl(base, leader, 64);
        if (pred) queue[base + __popcll(m & lanes_below())] = slot;
}

/* Sum a per-lane step tally over the wave into the global counter. */
__device__ __forceinline__ void steps_flush(
    const KernelArgs & a, unsigned long long acc)
{
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0 && acc) atomicAdd(a.counters + C_STEPS, acc);
}

/* Entry of every history of the batch; one lane per history. The
 * block's states (104 B records, 26 KB per block) are staged through LDS:
 * read from HBM with consecutive lanes on consecutive 8 B words, and
 * written back the same way (unchanged for the histories that go on to
 * the stepping kernel), instead of each lane touching its own record with
 * 104 B strides. */
constexpr int INIT_BLOCK = 256;

/* 16 B at an 8 B aligned address (global or LDS), by value. */
__device__ __forceinline__ double2 load16(const double * p)
{
        double2 v;
        __builtin_memcpy(&v, p, sizeof(v));
        return v;
}

__device__ __forceinline__ void store16(double * p, double2 v)
{
        __builtin_memcpy(p, &v, sizeof(v));
}

template <bool F32, bool CLOSED>
__global__ void __launch_bounds__(INIT_BLOCK)
    __attribute__((amdgpu_waves_per_eu(CLOSED ? 5 : 1))) k_init(KernelArgs args)
{
        KERNEL_ARGS(args);
        __shared__ struct pumas_state stage[INIT_BLOCK];
        const EngineT<F32> eng = { a.p, a.mv, a.c, a.g, a.ft, a.p, a.rec };
        const uint64_t base = (uint64_t)blockIdx.x * INIT_BLOCK;
        const uint64_t i = base + threadIdx.x;
        const bool active = i < a.n;
        const uint32_t words = (uint32_t)(
            ((a.n - base < INIT_BLOCK) ? a.n - base : INIT_BLOCK) *
            (sizeof(struct pumas_state) / sizeof(double)));
        double * const lds = reinterpret_cast<double *>(stage);
        double * const hbm = reinterpret_cast<double *>(a.states + base);
        constexpr uint32_t WORDS =
            INIT_BLOCK * (sizeof(struct pumas_state) / sizeof(double));
        /* A full block moves 16 B words (states are 8 B aligned; the
         * 16 B loads and stores need no more): 6.5 per lane. */
        constexpr uint32_t PAIRS = WORDS / 2;
        constexpr uint32_t PER_LANE = (PAIRS + INIT_BLOCK - 1) / INIT_BLOCK;
        static_assert(sizeof(struct pumas_state) % sizeof(double) == 0,
            "states are staged as 8 B words");
        static_assert(WORDS % 2 == 0, "a block holds whole 16 B words");
        if (words == WORDS) {
                /* Every load of the lane issued before the first wait (a
                 * loop with a runtime bound waits on each load in turn,
                 * one HBM round trip per word). */
                /* Words k < PER_LANE - 1 exist for every lane; the last
                 * for the first PAIRS % INIT_BLOCK lanes. */
                const bool tail = threadIdx.x + (PER_LANE - 1) * INIT_BLOCK <
                    PAIRS;
                double2 v[PER_LANE];
#pragma unroll
                for (uint32_t k = 0; k < PER_LANE - 1; k++)
                        v[k] = load16(hbm + 2 * (threadIdx.x + k * INIT_BLOCK));
                v[PER_LANE - 1] = tail ? load16(hbm + 2 * (threadIdx.x +
                    (PER_LANE - 1) * INIT_BLOCK)) : v[0];
#pragma unroll
                for (uint32_t k = 0; k < PER_LANE - 1; k++)
                        store16(lds + 2 * (threadIdx.x + k * INIT_BLOCK), v[k]);
                if (tail)
                        store16(lds + 2 * (threadIdx.x +
                            (PER_LANE - 1) * INIT_BLOCK), v[PER_LANE - 1]);
        } else {
                for (uint32_t k = threadIdx.x; k < words; k += INIT_BLOCK)
                        lds[k] = hbm[k];
        }
        __syncthreads();
        Walker w;
        w.phase = PHASE_EMPTY;
        if (active) {
                walker_load(a, i, w, stage + threadIdx.x);
                eng.template walker_init<CLOSED>(w);
                if (w.phase == PHASE_DONE)
                        walker_store(a, i, w, stage + threadIdx.x);
                if (!CLOSED && (w.phase == PHASE_STEP)) a.records[i] = w;
        }
        __syncthreads();
        if (words == WORDS) {
                const bool tail = threadIdx.x + (PER_LANE - 1) * INIT_BLOCK <
                    PAIRS;
                double2 v[PER_LANE];
#pragma unroll
                for (uint32_t k = 0; k < PER_LANE - 1; k++)
                        v[k] = load16(lds + 2 * (threadIdx.x + k * INIT_BLOCK));
                v[PER_LANE - 1] = tail ? load16(lds + 2 * (threadIdx.x +
                    (PER_LANE - 1) * INIT_BLOCK)) : v[0];
#pragma unroll
                for (uint32_t k = 0; k < PER_LANE - 1; k++)
                        store16(hbm + 2 * (threadIdx.x + k * INIT_BLOCK), v[k]);
                if (tail)
                        store16(hbm + 2 * (threadIdx.x +
                            (PER_LANE - 1) * INIT_BLOCK), v[PER_LANE - 1]);
        } else {
                for (uint32_t k = threadIdx.x; k < words; k += INIT_BLOCK)
                        hbm[k] = lds[k];
        }
        if (!CLOSED)
                queue_append(a.counters + C_STEP_N0, a.q_step[0],
                    w.phase == PHASE_STEP, (uint32_t)i);
}

/* Default maximum number of lanes of a wavefront waiting for an elastic
 * hard scattering before the wavefront processes them together. */
constexpr int EHS_BATCH = 16;

/* Clears the per-round counters before round `parity`. */
__global__ void k_round(unsigned long long * c, int parity)
{
        if (threadIdx.x == 0) {
                c[C_TAKEN] = 0;
                c[C_DEL_N] = 0;
                c[C_STEP_N0 + (parity ^ 1)] = 0;
        }
}

/*
 * Stepping: persistent over the step queue of round `parity`. With
 * TAIL_DEL, a wavefront that can no longer refill (the step queue is
 * exhausted) applies the DELs of its histories itself instead of queueing
 * them for the next round: the last histories of a batch then run to their
 * end without a round barrier per DEL, which sets the latency of small
 * batches. The out-of-line DEL call costs registers on every step, so
 * rounds that fill the device use the variant without it.
 */
template <bool F32, bool TAIL_DEL, bool LDS>
__global__ void __launch_bounds__(512, 2) k_step(KernelArgs args, int parity)
{
        KERNEL_ARGS(args);
        /* With LDS, the tables of the step lookups in the working precision
         * are read from a copy in LDS: the FP32 mirrors (FP32-bulk) or the
         * packed FP64 step tables. */
        using FT = typename std::conditional<LDS && F32, LdsTables,
            FloatTables>::type;
        using ST = typename std::conditional<LDS, LdsStepTables,
            PhysicsView>::type;
        const FT * ft;
        const ST * st;
        if constexpr (LDS && F32)
                ft = &a.ftl;
        else
                ft = &a.ft;
        if constexpr (LDS)
                st = &a.stl;
        else
                st = &a.p;
        const EngineT<F32, FT, ST> eng = { a.p, a.mv, a.c, a.g, *ft, *st,
                a.rec };
        const uint64_t below = lanes_below();
        /* Lanes of the wavefront that take histories: all of them, or the
         * first lanes_per_wave when a small batch is spread over the
         * device. */
        const bool lane_on = (int)(threadIdx.x & 63) < a.lanes_per_wave;
        const uint32_t * in = a.q_step[parity];
        const unsigned long long n_in = a.counters[C_STEP_N0 + parity];
        if (n_in == 0) return;
        /* Adaptive DEL mode: both variants are launched every round and
         * the one that does not fit the round's queue returns at once. The
         * in-wavefront DEL call costs registers on every step, so rounds
         * that fill the resident lanes run without it; the last rounds of
         * a batch, and small batches, run with it. */
        if (a.tail_adaptive && (TAIL_DEL != (n_in <= a.tail_del_lanes)))
                return;
        if constexpr (LDS) {
                /* The copy (tens of KB to 160 KB, read from L2 by every
                 * block of the launch), in 16 B words. */
                float4 * lds = (float4 *)pg_lds_floats;
                for (int k = 0; k < a.n_lds_pieces; k++) {
                        const float4 * src = a.lds_piece[k].src;
                        float4 * dst = lds + a.lds_piece[k].dst;
                        for (uint32_t i = threadIdx.x; i < a.lds_piece[k].n4;
                             i += blockDim.x)
                                dst[i] = src[i];
                }
                __syncthreads();
        }
        Walker w;
        w.phase = PHASE_EMPTY;
        uint32_t slot = 0;
        bool exhausted = false;
        unsigned long long steps_acc = 0;
        for (;;) {
                /* Hand over the histories that left the stepping phase. */
                const bool done = w.phase == PHASE_DONE;
                const bool to_del =
                    (w.phase == PHASE_DEL) && !(TAIL_DEL && exhausted);
                if (done) {
                        walker_store(a, slot, w, a.states + slot);
                        steps_acc += w.t.steps;
                }
                if (to_del) a.records[slot] = w;
                queue_append(a.counters + C_DEL_N, a.q_del, to_del, slot);
                if (done || to_del) w.phase = PHASE_EMPTY;

                /* Refill the free lanes from the step queue. */
                if (!exhausted) {
                        const uint64_t free_lanes =
                            __ballot((w.phase == PHASE_EMPTY) && lane_on);
                        if (free_lanes) {
                                const int count = __popcll(free_lanes);
                                const int leader = __ffsll(
                                    (unsigned long long)free_lanes) - 1;
                                unsigned long long base = 0;
                                if ((int)(threadIdx.x & 63) == leader)
                                        base = atomicAdd(
                                            a.counters + C_TAKEN, count);
                                base = __shfl(base, leader, 64);
                                if (base + count >= n_in) exhausted = true;
                                if ((w.phase == PHASE_EMPTY) && lane_on) {
                                        const unsigned long long k = base +
                                            __popcll(free_lanes & below);
                                        if (k < n_in) {
                                                slot = in[k];
                                                w = a.records[slot];
                                        }
                                }
                        }
                }
                if (!__ballot((w.phase == PHASE_STEP) ||
                        (w.phase == PHASE_EHS) ||
                        (TAIL_DEL && (w.phase == PHASE_DEL)))) {
                        if (exhausted) break;
                        continue;
                }
                if (w.phase == PHASE_STEP) eng.walker_step(w);

                /* Elastic hard scatterings, on ~16% of the steps of a muon
                 * in rock, run once enough lanes wait for one (half of the
                 * lanes in flight, at most a.ehs_batch) or none can step, so
                 * that the wavefront runs that code with most lanes busy. */
                const uint64_t ehs = __ballot(w.phase == PHASE_EHS);
                const uint64_t stepping = __ballot(w.phase == PHASE_STEP);
                const int waiting = __popcll(ehs);
                const int in_flight = waiting + __popcll(stepping);
                const int batch = (in_flight + 1) / 2 < a.ehs_batch ?
                    (in_flight + 1) / 2 :
                    a.ehs_batch;
                if (ehs && ((waiting >= batch) || !stepping)) {
                        if (w.phase == PHASE_EHS) eng.walker_ehs(w);
                }
                if (TAIL_DEL && exhausted && (w.phase == PHASE_DEL))
                        eng.walker_del(w);
        }
        steps_flush(a, steps_acc);
}

/* Inelastic DELs of round `parity`, one lane per queued history; the
 * histories go back to the next round's step queue. */
template <bool F32>
__global__ void __launch_bounds__(256, 2) k_del(KernelArgs args, int parity)
{
        KERNEL_ARGS(args);
        const EngineT<F32> eng = { a.p, a.mv, a.c, a.g, a.ft, a.p, a.rec };
        const unsigned long long n = a.counters[C_DEL_N];
        const uint32_t * queue = a.q_del;
        const int lane = threadIdx.x & 63;
        const unsigned long long stride = (unsigned long long)gridDim.x *
            blockDim.x;
        unsigned long long steps_acc = 0;
        for (unsigned long long base = (unsigned long long)blockIdx.x *
                     blockDim.x + (threadIdx.x & ~63u);
             base < n; base += stride) {
                const unsigned long long k = base + lane;
                const bool active = k < n;
                Walker w;
                w.phase = PHASE_EMPTY;
                uint32_t slot = 0;
                if (active) {
                        slot = queue[k];
                        w = a.records[slot];
                        eng.walker_del(w);
                        if (w.phase == PHASE_DONE) {
                                walker_store(a, slot, w, a.states + slot);
                                steps_acc += w.t.steps;
                        } else {
                                a.records[slot] = w;
                        }
                }
                queue_append(a.counters + C_STEP_N0 + (parity ^ 1),
                    a.q_step[parity ^ 1], w.phase == PHASE_STEP, slot);
        }
        steps_flush(a, steps_acc);
}

/* Clears the counters (a kernel, not a memset node, so that the launch
 * sequence can be captured in a graph). */
__global__ void zero_counters(unsigned long long * c, int n)
{
        if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

size_t pg_bytes_per_history(void) { return sizeof(Walker) + 3 * sizeof(uint32_t); }

/* ------------------------------------------------------------------------
 * C ABI transport.
 * ------------------------------------------------------------------------ */
namespace {

/* Histories per chunk: the records of a chunk take MAX_CHUNK * sizeof(Walker)
 * bytes of HBM (about 11 GB). */
constexpr size_t MAX_CHUNK = size_t(1) << 25;

/* Histories per chunk of a closed form launch, which keeps no records or
 * queues and indexes its histories in 64 bits: bounded by the entry
 * kernel's grid (2^30 lanes), so that a large batch is one launch and one
 * counter read instead of one per 2^25 histories. */
constexpr size_t MAX_CLOSED_CHUNK = size_t(1) << 30;

/* Default number of rounds launched between two reads of the queue
 * lengths, and its maximum. */
constexpr int ROUNDS_PER_CHECK = 16;
constexpr int MAX_ROUNDS_PER_CHECK = 256;

/* Launch geometry and optional timing of the stepping kernel. */
struct Launch {
        int n_cu, block, per_cu, rounds_per_check;
        bool graph;
        int tail_del; /* <0 never, >0 always, 0: rounds below the lanes */
        int lanes_per_wave; /* 1..64, or 0: spread small batches */
        unsigned lds_bytes; /* LDS copy of the step tables, or 0 */
        unsigned step_grid, del_grid;
        hipStream_t stream;
        hipEvent_t * marks; /* [2 * rounds_per_check] or NULL */
        hipEvent_t * init_marks; /* [2], around the entry kernel, or NULL */
        bool closed; /* every history takes the closed form CSDA */
};

template <bool F32, bool TAIL_DEL>
void launch_step(const KernelArgs & a, int parity, const Launch & L)
{
        if (L.lds_bytes)
                hipLaunchKernelGGL((k_step<F32, TAIL_DEL, true>),
                    dim3(L.step_grid), dim3(L.block), L.lds_bytes, L.stream,
                    a, parity);
        else
                hipLaunchKernelGGL((k_step<F32, TAIL_DEL, false>),
                    dim3(L.step_grid), dim3(L.block), 0, L.stream, a, parity);
}

template <bool F32>
enum pumas_return launch_round(
    const KernelArgs & a, int parity, const Launch & L, int r)
{
        hipStream_t stream = L.stream;
        hipLaunchKernelGGL(k_round, dim3(1), dim3(64), 0, stream, a.counters,
            parity);
        if (L.marks) HIP_TRY(hipEventRecord(L.marks[2 * r], stream));
        if (L.tail_del <= 0) launch_step<F32, false>(a, parity, L);
        if (L.tail_del >= 0) launch_step<F32, true>(a, parity, L);
        if (L.marks) HIP_TRY(hipEventRecord(L.marks[2 * r + 1], stream));
        hipLaunchKernelGGL(k_del<F32>, dim3(L.del_grid), dim3(256), 0,
            stream, a, parity);
        HIP_TRY(hipGetLastError());
        return PUMAS_RETURN_SUCCESS;
}

template <bool F32>
enum pumas_return allow_lds(int bytes)
{
        HIP_TRY(hipFuncSetAttribute((const void *)k_step<F32, true, true>,
            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        HIP_TRY(hipFuncSetAttribute((const void *)k_step<F32, false, true>,
            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        return PUMAS_RETURN_SUCCESS;
}

/* Bytes of the LDS copy of the step tables in the stepping kernel, or 0
 * when `per_cu` copies do not fit in the LDS of a CU. */
enum pumas_return lds_tables_size(const pumas_gpu_physics * physics,
    int device, int per_cu, bool f32, unsigned * bytes)
{
        *bytes = 0;
        const size_t need = f32 ?
            (size_t)physics->n_floats * sizeof(float) +
                (size_t)(physics->range_head + physics->n_step -
                    physics->range_tail_at) * sizeof(double) :
            (size_t)physics->n_step * sizeof(double);
        int per_block = 0, per_sm = 0;
        HIP_TRY(hipDeviceGetAttribute(
            &per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        HIP_TRY(hipDeviceGetAttribute(&per_sm,
            hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
        if ((need > (size_t)per_block) || (need * per_cu > (size_t)per_sm))
                return PUMAS_RETURN_SUCCESS;
        if (need > 65536) {
                const enum pumas_return rc = f32 ?
                    allow_lds<true>((int)need) :
                    allow_lds<false>((int)need);
                if (rc != PUMAS_RETURN_SUCCESS) return rc;
        }
        *bytes = (unsigned)need;
        return PUMAS_RETURN_SUCCESS;
}

/* Device memory and graph of a chunk, released on every return path. */
struct ChunkResources {
        hipStream_t stream;
        void * mem = nullptr;
        hipGraphExec_t exec = nullptr;

        ~ChunkResources()
        {
                if (exec) (void)hipGraphExecDestroy(exec);
                if (mem) (void)hipFreeAsync(mem, stream);
        }
};

/* Device bytes of the per-chunk buffers (counters, records, queues) of a
 * chunk of m histories. */
constexpr size_t chunk_bytes(size_t m)
{
        return 256 + (((m * sizeof(Walker)) + 255) & ~size_t(255)) +
            3 * (((m * sizeof(uint32_t)) + 255) & ~size_t(255));
}

/* Transport histories [base, base + m) of the batch described by `a0`.
 * The chunk's buffers are allocated on the stream, or taken from `fixed`
 * (at least chunk_bytes(m) bytes of device memory) when it is given. */
enum pumas_return transport_chunk(const KernelArgs & a0, size_t base,
    size_t m, bool f32, Launch L, unsigned long long * total_steps,
    unsigned long long * error, struct pumas_gpu_stats * stats,
    void * fixed = nullptr)
{
        /* No more resident blocks than the chunk fills with histories, so
         * that a small request takes full wavefronts on few CUs and leaves
         * the rest of the device to concurrent requests. */
        const size_t resident = (size_t)L.n_cu * L.per_cu;
        const size_t waves = (size_t)(L.block / 64);
        int lanes = L.lanes_per_wave;
        if (lanes <= 0) {
                /* Spread: as many histories per wavefront as it takes for
                 * the batch to fill every resident wavefront. */
                const size_t per = (m + resident * waves - 1) /
                    (resident * waves);
                lanes = (per < 1) ? 1 : (per > 64 ? 64 : (int)per);
        }
        const size_t per_block = waves * (size_t)lanes;
        const size_t fill = (m + per_block - 1) / per_block;
        L.step_grid = (unsigned)((fill < resident) ? (fill > 0 ? fill : 1) :
                                                     resident);
        const size_t del_fill = (m + 255) / 256;
        const size_t del_max = (size_t)L.n_cu * 2;
        L.del_grid = (unsigned)((del_fill < del_max) ?
                (del_fill > 0 ? del_fill : 1) :
                del_max);
        hipStream_t stream = L.stream;
        KernelArgs a = a0;
        a.lanes_per_wave = lanes;
        a.tail_del_lanes = L.step_grid * (uint32_t)(waves * lanes);
        a.tail_adaptive = (L.tail_del == 0);
        a.states = a0.states + base;
        a.n = m;
        a.first = a0.first + base;
        if (a0.events) a.events = a0.events + base;
        if (a0.media) a.media = a0.media + 2 * base;
        if (a0.steps) a.steps = a0.steps + base;
        if (a0.draws) a.draws = a0.draws + base;
        if (a0.rc) a.rc = a0.rc + base;

        ChunkResources res;
        res.stream = stream;
        const size_t b_counters = 256;
        /* A closed form launch needs no records or queues. */
        const size_t b_records = L.closed ? 0 :
            ((m * sizeof(Walker)) + 255) & ~size_t(255);
        const size_t b_queue = L.closed ? 0 :
            ((m * sizeof(uint32_t)) + 255) & ~size_t(255);
        char * ptr;
        if (fixed) {
                ptr = static_cast<char *>(fixed);
        } else {
                HIP_TRY(hipMallocAsync(
                    &res.mem, b_counters + b_records + 3 * b_queue, stream));
                ptr = static_cast<char *>(res.mem);
        }
        a.counters = reinterpret_cast<unsigned long long *>(ptr);
        a.records = reinterpret_cast<Walker *>(ptr + b_counters);
        char * q = ptr + b_counters + b_records;
        a.q_step[0] = reinterpret_cast<uint32_t *>(q);
        a.q_step[1] = reinterpret_cast<uint32_t *>(q + b_queue);
        a.q_del = reinterpret_cast<uint32_t *>(q + 2 * b_queue);

        enum pumas_return rc = PUMAS_RETURN_SUCCESS;
        unsigned long long host[C_COUNT];
        hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, stream,
            a.counters, C_COUNT);
        const unsigned blocks = (unsigned)((m + INIT_BLOCK - 1) / INIT_BLOCK);
        if (L.init_marks) HIP_TRY(hipEventRecord(L.init_marks[0], stream));
        if (f32 && L.closed)
                hipLaunchKernelGGL((k_init<true, true>), dim3(blocks),
                    dim3(INIT_BLOCK), 0, stream, a);
        else if (f32)
                hipLaunchKernelGGL((k_init<true, false>), dim3(blocks),
                    dim3(INIT_BLOCK), 0, stream, a);
        else if (L.closed)
                hipLaunchKernelGGL((k_init<false, true>), dim3(blocks),
                    dim3(INIT_BLOCK), 0, stream, a);
        else
                hipLaunchKernelGGL((k_init<false, false>), dim3(blocks),
                    dim3(INIT_BLOCK), 0, stream, a);
        if (L.init_marks) HIP_TRY(hipEventRecord(L.init_marks[1], stream));
        if (hipGetLastError() != hipSuccess)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "transport kernel launch failed");
        /* The sequence of rounds between two checks is captured once in a
         * HIP graph and replayed (an even number of rounds, so that every
         * replay starts on step queue 0): one launch per check instead of
         * three per round. Counters are cleared by kernels, not memset
         * nodes. */
        if (L.graph) {
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(
                    stream, hipStreamCaptureModeThreadLocal));
                int p = 0;
                for (int r = 0; r < L.rounds_per_check; r++, p ^= 1) {
                        rc = f32 ? launch_round<true>(a, p, L, r) :
                                   launch_round<false>(a, p, L, r);
                        if (rc != PUMAS_RETURN_SUCCESS) break;
                }
                const hipError_t e = hipStreamEndCapture(stream, &graph);
                if (rc != PUMAS_RETURN_SUCCESS) {
                        if (graph) (void)hipGraphDestroy(graph);
                        return rc;
                }
                HIP_TRY(e);
                const hipError_t ei =
                    hipGraphInstantiate(&res.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_TRY(ei);
        }
        int parity = 0, launched = 0;
        bool init_timed = false;
        for (;;) {
                HIP_TRY(hipMemcpyAsync(host, a.counters, sizeof(host),
                    hipMemcpyDeviceToHost, stream));
                HIP_TRY(hipStreamSynchronize(stream));
                if (L.init_marks && stats && (parity == 0) &&
                    (launched == 0) && !init_timed) {
                        float ms = 0.f;
                        HIP_TRY(hipEventElapsedTime(&ms, L.init_marks[0],
                            L.init_marks[1]));
                        stats->init_kernel_ms += ms;
                        init_timed = true;
                }
                if (L.marks && stats) {
                        for (int r = 0; r < launched; r++) {
                                float ms = 0.f;
                                HIP_TRY(hipEventElapsedTime(&ms,
                                    L.marks[2 * r], L.marks[2 * r + 1]));
                                stats->step_kernel_ms += ms;
                        }
                }
                if (stats) stats->step_launches += launched;
                launched = 0;
                if (host[C_STEP_N0 + parity] == 0) break;
                if (stats) stats->rounds += 1;
                if (res.exec) {
                        HIP_TRY(hipGraphLaunch(res.exec, stream));
                        launched = L.rounds_per_check;
                        continue;
                }
                for (int r = 0; r < L.rounds_per_check; r++) {
                        rc = f32 ? launch_round<true>(a, parity, L, r) :
                                   launch_round<false>(a, parity, L, r);
                        if (rc != PUMAS_RETURN_SUCCESS) break;
                        launched++;
                        parity ^= 1;
                }
                if (rc != PUMAS_RETURN_SUCCESS) break;
        }
        *total_steps += host[C_STEPS];
        if ((*error == 0) && (host[C_ERROR] != 0)) *error = host[C_ERROR];
        if (stats) {
                stats->step_limited += host[C_STATUS];
                stats->loop_limited += host[C_STATUS_LOOP];
                stats->no_target += host[C_STATUS_TARGET];
        }
        if (res.mem) {
                void * mem = res.mem;
                res.mem = nullptr;
                HIP_TRY(hipFreeAsync(mem, stream));
        }
        return rc;
}


enum pumas_return check_context(const pumas_gpu_physics * physics,
    const struct pumas_gpu_context * c)
{
        if ((c->accuracy <= 0) || (c->accuracy > 1))
                return pg_fail(PUMAS_RETURN_ACCURACY_ERROR,
                    "bad accuracy value (expected a value in ]0,1], got %g)",
                    c->accuracy);
        if ((c->mode.direction == PUMAS_MODE_BACKWARD) &&
            (c->mode.energy_loss > PUMAS_MODE_CSDA) &&
            (physics->view.cutoff < 1E-02))
                return pg_fail(PUMAS_RETURN_ACCURACY_ERROR,
                    "bad cutoff value for backward transport (expected a "
                    "value greater than or equal to 0.01, got %g)",
                    physics->view.cutoff);
        if ((physics->particle != 0) &&
            (c->mode.direction == PUMAS_MODE_FORWARD) &&
            (c->mode.decay == PUMAS_MODE_WEIGHTED))
                return pg_fail(PUMAS_RETURN_DECAY_ERROR,
                    "`PUMAS_MODE_WEIGHTED' mode is not valid for forward taus");
        return PUMAS_RETURN_SUCCESS;
}

} /* namespace */

/* The transport call. With `fixed` (device memory for chunk_bytes(n)
 * bytes) nothing is allocated: no chunk buffers, no graph (eager
 * launches), no timing events (stats then hold counts only). */
static enum pumas_return transport_impl(const pumas_gpu_physics * physics,
    const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states, size_t n,
    uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream_, void * fixed)
{
        if ((physics == NULL) || (context == NULL) || (geometry == NULL))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        if ((states == NULL) && (n > 0))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "no states (null)");
        enum pumas_return rc = check_context(physics, context);
        if (rc != PUMAS_RETURN_SUCCESS) return rc;
        const int precision =
            options ? options->precision : PUMAS_GPU_PRECISION_FP64;
        if ((precision != PUMAS_GPU_PRECISION_FP64) &&
            (precision != PUMAS_GPU_PRECISION_FP32_BULK))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "unknown precision mode %d", precision);
        hipStream_t stream = static_cast<hipStream_t>(stream_);

        KernelArgs a;
        memset(&a, 0x0, sizeof(a));
        a.p = physics->view;
        a.mv = physics->models;
        a.ft = physics->floats;
        a.ftl = physics->lds_floats;
        if (precision == PUMAS_GPU_PRECISION_FP32_BULK) {
                a.stl = physics->lds_range;
                const uint32_t nf4 = physics->n_floats / 4;
                const uint32_t head4 = physics->range_head / 2;
                a.lds_piece[0] = { (const float4 *)physics->floats.K, 0, nf4 };
                a.lds_piece[1] = { (const float4 *)physics->step, nf4, head4 };
                a.lds_piece[2] = { (const float4 *)(physics->step +
                                       physics->range_tail_at),
                        nf4 + head4,
                        (physics->n_step - physics->range_tail_at) / 2 };
                a.n_lds_pieces = 3;
        } else {
                a.stl = physics->lds_step;
                a.lds_piece[0] = { (const float4 *)physics->step, 0,
                        physics->n_step / 2 };
                a.n_lds_pieces = 1;
        }
        a.c.energy_loss = context->mode.energy_loss;
        a.c.decay = context->mode.decay;
        a.c.direction = context->mode.direction;
        a.c.scattering = context->mode.scattering;
        a.c.event = context->event;
        a.c.limit_energy = context->limit.energy;
        a.c.limit_distance = context->limit.distance;
        a.c.limit_grammage = context->limit.grammage;
        a.c.limit_time = context->limit.time;
        a.c.accuracy = context->accuracy;
        a.c.loop_max = (options && options->max_loop > 0) ? options->max_loop :
                                                            LOOP_MAX_DEFAULT;
        a.g = geometry->view;
        a.states = states;
        a.n = n;
        a.seed = seed;
        a.first = first_history;
        if (outputs) {
                a.events = outputs->events;
                a.media = outputs->media;
                a.steps = outputs->steps;
                a.draws = outputs->draws;
                a.rc = outputs->rc;
        }
        a.max_steps = (options && options->max_steps > 0) ? options->max_steps :
                                                            10000000;
        a.ehs_batch = (options && options->ehs_batch > 0) ? options->ehs_batch :
                                                            EHS_BATCH;
        if (options && options->recorder) {
                const pumas_gpu_recorder * r = options->recorder;
                a.rec.frames = r->frames;
                a.rec.count = r->count;
                a.rec.capacity = r->capacity;
                a.rec.period = r->period;
                a.rec.first = r->first;
                a.rec.last = r->first + r->n;
        }
        if (a.ehs_batch > 64)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "ehs_batch must be at most 64 (got %d)", a.ehs_batch);

        int device;
        HIP_TRY(hipGetDevice(&device));
        int n_cu = 0;
        HIP_TRY(hipDeviceGetAttribute(
            &n_cu, hipDeviceAttributeMultiprocessorCount, device));
        /* Default launch: 8 resident wavefronts per CU, as one block of
         * 512 lanes (one LDS copy of the step tables per CU). */
        const int block = (options && options->block_size > 0) ?
            options->block_size :
            512;
        if ((block % 64) || (block > 512))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "block size must be a multiple of 64, at most 512");
        const int per_cu = (options && options->blocks_per_cu > 0) ?
            options->blocks_per_cu :
            (512 / block > 0 ? 512 / block : 1);
        int rounds_per_check = (options && options->rounds_per_check > 0) ?
            options->rounds_per_check :
            ROUNDS_PER_CHECK;
        rounds_per_check += rounds_per_check & 1; /* even, for the graph */
        if (rounds_per_check > MAX_ROUNDS_PER_CHECK)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "rounds_per_check must be at most %d",
                    MAX_ROUNDS_PER_CHECK);

        /* Timing events, destroyed on every return path. */
        struct Events {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                hipEvent_t init[2] = { nullptr, nullptr };
                std::vector<hipEvent_t> marks;

                ~Events()
                {
                        if (e0) (void)hipEventDestroy(e0);
                        if (e1) (void)hipEventDestroy(e1);
                        for (auto & e : init)
                                if (e) (void)hipEventDestroy(e);
                        for (auto & e : marks)
                                if (e) (void)hipEventDestroy(e);
                }
        } ev;
        if (stats && !fixed) {
                HIP_TRY(hipEventCreate(&ev.e0));
                HIP_TRY(hipEventCreate(&ev.e1));
                HIP_TRY(hipEventCreate(&ev.init[0]));
                HIP_TRY(hipEventCreate(&ev.init[1]));
                HIP_TRY(hipEventRecord(ev.e0, stream));
        }
        const bool f32 = (precision == PUMAS_GPU_PRECISION_FP32_BULK);
        unsigned lds_bytes = 0;
        if (!(options && options->lds_tables < 0)) {
                rc = lds_tables_size(physics, device, per_cu, f32, &lds_bytes);
                if (rc != PUMAS_RETURN_SUCCESS) return rc;
        }
        unsigned long long total_steps = 0, error = 0;
        /* The legacy NULL stream cannot be captured. */
        const bool graph = !(options && options->eager) &&
            (stream != nullptr) && !fixed;
        if (stats) {
                memset(stats, 0x0, sizeof(*stats));
                if (!graph && !fixed) {
                        ev.marks.resize(2 * rounds_per_check, nullptr);
                        for (auto & e : ev.marks)
                                HIP_TRY(hipEventCreate(&e));
                }
        }
        /* Per-launch timing of the stepping kernel needs eager launches:
         * events recorded inside a graph cannot be timed. */
        const int tail_del = options ? options->tail_del : 0;
        int lanes_per_wave = options ? options->lanes_per_wave : 0;
        if (lanes_per_wave > 64)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "lanes_per_wave must be at most 64 (got %d)",
                    lanes_per_wave);
        /* Every history takes the closed form CSDA when the geometry has
         * no step limits (a uniform infinite medium), the energy loss is
         * CSDA and scattering is disabled (src/pumas.c:3200-3230). */
        const bool closed = (geometry->view.kind == GEO_INFINITE) &&
            (geometry->view.locals == nullptr) &&
            (context->mode.energy_loss == PUMAS_MODE_CSDA) &&
            (context->mode.scattering == PUMAS_MODE_DISABLED);
        const Launch L = { n_cu, block, per_cu, rounds_per_check, graph,
                tail_del, lanes_per_wave, lds_bytes, 0, 0, stream,
                (stats && !graph && !fixed) ? ev.marks.data() : nullptr,
                (stats && !fixed) ? ev.init : nullptr, closed };
        const size_t chunk = closed ? MAX_CLOSED_CHUNK : MAX_CHUNK;
        for (size_t base = 0; base < n; base += chunk) {
                const size_t m = (n - base < chunk) ? n - base : chunk;
                rc = transport_chunk(
                    a, base, m, f32, L, &total_steps, &error, stats, fixed);
                if (rc != PUMAS_RETURN_SUCCESS) break;
        }
        if (stats && !fixed) HIP_TRY(hipEventRecord(ev.e1, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (stats) {
                float ms = 0.f;
                if (!fixed) HIP_TRY(hipEventElapsedTime(&ms, ev.e0, ev.e1));
                stats->steps = total_steps;
                stats->histories = n;
                stats->kernel_ms = ms;
                stats->lds_bytes = lds_bytes;
        }
        if (rc != PUMAS_RETURN_SUCCESS) return rc;
        if (error != 0)
                return pg_fail((enum pumas_return)error,
                    "a history raised error code %llu", error);
        return PUMAS_RETURN_SUCCESS;
}

extern "C" enum pumas_return pumas_gpu_transport(
    const pumas_gpu_physics * physics, const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states, size_t n,
    uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream)
{
        pg_clear_error();
        return transport_impl(physics, context, geometry, states, n, seed,
            first_history, outputs, options, stats, stream, nullptr);
}

/* Public layout of the reference's struct pumas_context
 * (include/pumas.h:636-672): four host pointers (medium and random
 * callbacks, recorder, user data), then the value part, laid out as
 * struct pumas_gpu_context (mode, event, limit, accuracy). */
struct reference_context_layout {
        void * medium;
        void * random;
        void * recorder;
        void * user_data;
        struct pumas_gpu_context values;
};
static_assert(offsetof(reference_context_layout, values) == 32,
    "reference context: four pointers before its mode");
static_assert(offsetof(struct pumas_gpu_context, event) == 16 &&
        offsetof(struct pumas_gpu_context, limit) == 24 &&
        offsetof(struct pumas_gpu_context, accuracy) == 56,
    "value part of the reference context");

extern "C" void pumas_gpu_context_from_reference(
    const struct pumas_context * context, struct pumas_gpu_context * out)
{
        if ((context == NULL) || (out == NULL)) return;
        memcpy(out,
            &reinterpret_cast<const reference_context_layout *>(context)
                 ->values,
            sizeof(*out));
}

extern "C" enum pumas_return pumas_gpu_transport_context(
    const pumas_gpu_physics * physics, const struct pumas_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states, size_t n,
    uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream)
{
        pg_clear_error();
        if (context == NULL)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL context");
        struct pumas_gpu_context values;
        pumas_gpu_context_from_reference(context, &values);
        return transport_impl(physics, &values, geometry, states, n, seed,
            first_history, outputs, options, stats, stream, nullptr);
}

/* ------------------------------------------------------------------------
 * Transport without allocations.
 *
 * The per-chunk buffers of the transport (counters, in-flight records,
 * queues) live in one file-scope device array per GPU, so that callers
 * that cannot allocate device memory (code under stream capture, serving
 * loops, harnesses that forbid hipMalloc) can transport batches of up to
 * PUMAS_GPU_NOALLOC_MAX_HISTORIES histories. In host code the array's name
 * is not a device address: its address on each device is looked up once
 * with hipGetSymbolAddress and only that pointer is passed to kernels.
 * The counters are cleared by a kernel at the start of every call; the
 * records and queues are written before they are read. Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU race.
 * ------------------------------------------------------------------------ */
constexpr size_t NOALLOC_BYTES = chunk_bytes(PUMAS_GPU_NOALLOC_MAX_HISTORIES);
static_assert(PUMAS_GPU_NOALLOC_MAX_HISTORIES <= MAX_CHUNK, "one chunk");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
void * g_scratch[MAX_DEVICES];

/* The device address of `scratch` on the current device (looked up on the
 * first call on that device). */
enum pumas_return scratch_pointer(void ** out)
{
        int device = 0;
        HIP_TRY(hipGetDevice(&device));
        if ((device < 0) || (device >= MAX_DEVICES))
                return pg_fail(PUMAS_RETURN_INDEX_ERROR,
                    "device %d beyond the %d supported", device, MAX_DEVICES);
        std::lock_guard<std::mutex> guard(g_scratch_lock);
        if (g_scratch[device] == nullptr) {
                void * ptr = nullptr;
                const hipError_t e =
                    hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
                if ((e != hipSuccess) || (ptr == nullptr))
                        return pg_fail(PUMAS_RETURN_MEMORY_ERROR,
                            "hipGetSymbolAddress(scratch): %s",
                            hipGetErrorString(e));
                g_scratch[device] = ptr;
        }
        *out = g_scratch[device];
        return PUMAS_RETURN_SUCCESS;
}
} /* namespace */

extern "C" enum pumas_return pumas_gpu_transport_noalloc(
    const pumas_gpu_physics * physics, const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states, size_t n,
    uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats,
    void * stream)
{
        pg_clear_error();
        if (n > PUMAS_GPU_NOALLOC_MAX_HISTORIES)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "%zu histories exceed the %d of the transport without "
                    "allocations (use pumas_gpu_transport, or split the "
                    "batch)", n, PUMAS_GPU_NOALLOC_MAX_HISTORIES);
        void * fixed = nullptr;
        const enum pumas_return rc = scratch_pointer(&fixed);
        if (rc != PUMAS_RETURN_SUCCESS) return rc;
        return transport_impl(physics, context, geometry, states, n, seed,
            first_history, outputs, options, stats, stream, fixed);
}

/* ------------------------------------------------------------------------
 * Frame recorder.
 * ------------------------------------------------------------------------ */
static_assert(sizeof(Frame) == sizeof(struct pumas_gpu_frame),
    "Frame and struct pumas_gpu_frame must have the same layout");

extern "C" enum pumas_return pumas_gpu_recorder_create(size_t capacity,
    int period, uint64_t first_history, uint64_t n_histories,
    pumas_gpu_recorder ** recorder)
{
        pg_clear_error();
        if (recorder == NULL)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        *recorder = nullptr;
        auto * r = new pumas_gpu_recorder();
        r->capacity = capacity;
        r->period = period;
        r->first = first_history;
        r->n = n_histories;
        r->frames = nullptr;
        r->count = nullptr;
        hipError_t e = hipMalloc((void **)&r->count, sizeof(*r->count));
        if (e == hipSuccess)
                e = hipMalloc((void **)&r->frames,
                    (capacity > 0 ? capacity : 1) * sizeof(Frame));
        if (e == hipSuccess) e = hipMemset(r->count, 0x0, sizeof(*r->count));
        if (e != hipSuccess) {
                (void)hipFree(r->count);
                (void)hipFree(r->frames);
                delete r;
                return pg_fail(PUMAS_RETURN_MEMORY_ERROR,
                    "could not allocate the recorder (%s)",
                    hipGetErrorString(e));
        }
        *recorder = r;
        return PUMAS_RETURN_SUCCESS;
}

extern "C" void pumas_gpu_recorder_destroy(pumas_gpu_recorder ** recorder)
{
        if ((recorder == NULL) || (*recorder == NULL)) return;
        (void)hipFree((*recorder)->frames);
        (void)hipFree((*recorder)->count);
        delete *recorder;
        *recorder = nullptr;
}

extern "C" enum pumas_return pumas_gpu_recorder_clear(
    pumas_gpu_recorder * recorder)
{
        pg_clear_error();
        if (recorder == NULL)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL recorder");
        HIP_TRY(hipMemset(recorder->count, 0x0, sizeof(*recorder->count)));
        return PUMAS_RETURN_SUCCESS;
}

extern "C" enum pumas_return pumas_gpu_recorder_read(
    const pumas_gpu_recorder * recorder, struct pumas_gpu_frame * frames,
    size_t max_frames, uint64_t * length)
{
        pg_clear_error();
        if (recorder == NULL)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL recorder");
        unsigned long long count = 0;
        HIP_TRY(hipMemcpy(&count, recorder->count, sizeof(count),
            hipMemcpyDeviceToHost));
        if (length) *length = count;
        size_t k = (count < recorder->capacity) ? count : recorder->capacity;
        if (k > max_frames) k = max_frames;
        if ((k > 0) && frames)
                HIP_TRY(hipMemcpy(frames, recorder->frames, k * sizeof(Frame),
                    hipMemcpyDeviceToHost));
        return PUMAS_RETURN_SUCCESS;
}

extern "C" enum pumas_return pumas_gpu_transport_host(
    const pumas_gpu_physics * physics, const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, struct pumas_state * states, size_t n,
    uint64_t seed, uint64_t first_history,
    const struct pumas_gpu_outputs * outputs,
    const struct pumas_gpu_options * options, struct pumas_gpu_stats * stats)
{
        /* Device copies, freed on every return path. The outputs are
         * cleared first, so that a call that fails before its first launch
         * returns the input states and zero outputs. */
        struct Buffers {
                struct pumas_state * states = nullptr;
                struct pumas_gpu_outputs out = {};
                hipStream_t stream = nullptr;
                ~Buffers()
                {
                        if (stream) {
                                (void)hipStreamSynchronize(stream);
                                (void)hipStreamDestroy(stream);
                        }
                        (void)hipFree(states);
                        (void)hipFree(out.events);
                        (void)hipFree(out.media);
                        (void)hipFree(out.steps);
                        (void)hipFree(out.draws);
                        (void)hipFree(out.rc);
                }
        } b;
        const size_t ns = n > 0 ? n : 1;
        HIP_TRY(hipMalloc(&b.states, ns * sizeof(*states)));
        HIP_TRY(hipMemcpy(b.states, states, n * sizeof(*states),
            hipMemcpyHostToDevice));
#define PG_OUT(field, count)                                                  \
        if (outputs && outputs->field) {                                       \
                HIP_TRY(hipMalloc(&b.out.field,                                \
                    (count) * ns * sizeof(*outputs->field)));                  \
                HIP_TRY(hipMemset(b.out.field, 0x0,                            \
                    (count) * ns * sizeof(*outputs->field)));                  \
        }
        PG_OUT(events, 1)
        PG_OUT(media, 2)
        PG_OUT(steps, 1)
        PG_OUT(draws, 1)
        PG_OUT(rc, 1)
#undef PG_OUT
        /* A stream of its own: the legacy NULL stream cannot be captured
         * in a graph. */
        HIP_TRY(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
        const enum pumas_return rc = pumas_gpu_transport(physics, context,
            geometry, b.states, n, seed, first_history, &b.out, options, stats,
            b.stream);
        HIP_TRY(hipStreamSynchronize(b.stream));
        const std::string message = g_error;
        HIP_TRY(hipMemcpy(states, b.states, n * sizeof(*states),
            hipMemcpyDeviceToHost));
#define PG_BACK(field, count)                                                 \
        if (outputs && outputs->field)                                         \
                HIP_TRY(hipMemcpy(outputs->field, b.out.field,                 \
                    (count) * n * sizeof(*outputs->field),                     \
                    hipMemcpyDeviceToHost));
        PG_BACK(events, 1)
        PG_BACK(media, 2)
        PG_BACK(steps, 1)
        PG_BACK(draws, 1)
        PG_BACK(rc, 1)
#undef PG_BACK
        snprintf(g_error, sizeof(g_error), "%s", message.c_str());
        return rc;
}
