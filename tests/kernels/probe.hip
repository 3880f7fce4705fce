/*
 * Test infrastructure, not part of the library: device evaluations of
 * engine routines for function-level comparisons with the host C library
 * and the oracle (tests/test_gpu_libm.py, tests/test_gpu_functions.py).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <string.h>
#include <vector>

#include "../../include/pumas_gpu.h"
#include "engine_transport.h"
#include "ptab.h"

enum {
        F_EXP = 0, F_LOG, F_POW, F_SIN, F_COS, F_SINCOS_S, F_SINCOS_C, F_LOG10,
        F_LOGF, F_EXPF, F_SQRT, F_DIV, F_ATAN,
        F_OCML = 100 /* + the same index: the device library's function */
};

__global__ void k_math(int which, const double * x, const double * y,
    double * out, size_t n)
{
        const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n) return;
        const double a = x[i], b = y[i];
        double r = 0.;
        double s, c;
        switch (which) {
        case F_EXP: r = pg::lm::exp(a); break;
        case F_LOG: r = pg::lm::log(a); break;
        case F_POW: r = pg::lm::pow(a, b); break;
        case F_SIN: r = pg::lm::sin(a); break;
        case F_COS: r = pg::lm::cos(a); break;
        case F_SINCOS_S: pg::lm::sincos(a, &s, &c); r = s; break;
        case F_SINCOS_C: pg::lm::sincos(a, &s, &c); r = c; break;
        case F_LOG10: r = pg::lm::log10(a); break;
        case F_LOGF: r = (double)pg::lm::logf((float)a); break;
        case F_EXPF: r = (double)pg::lm::expf((float)a); break;
        case F_SQRT: r = sqrt(a); break;
        case F_DIV: r = a / b; break;
        case F_ATAN: r = atan(a); break;
        case F_OCML + F_EXP: r = exp(a); break;
        case F_OCML + F_LOG: r = log(a); break;
        case F_OCML + F_POW: r = pow(a, b); break;
        case F_OCML + F_SIN: r = sin(a); break;
        case F_OCML + F_COS: r = cos(a); break;
        case F_OCML + F_LOG10: r = log10(a); break;
        case F_OCML + F_LOGF: r = (double)logf((float)a); break;
        case F_OCML + F_EXPF: r = (double)expf((float)a); break;
        default: r = __builtin_nan(""); break;
        }
        out[i] = r;
}

/* Host C library values of the same functions. */
static double host_value(int which, double a, double b)
{
        double s, c;
        switch (which % F_OCML) {
        case F_EXP: return exp(a);
        case F_LOG: return log(a);
        case F_POW: return pow(a, b);
        case F_SIN: return sin(a);
        case F_COS: return cos(a);
        case F_SINCOS_S: sincos(a, &s, &c); return s;
        case F_SINCOS_C: sincos(a, &s, &c); return c;
        case F_LOG10: return log10(a);
        case F_LOGF: return (double)logf((float)a);
        case F_EXPF: return (double)expf((float)a);
        case F_SQRT: return sqrt(a);
        case F_DIV: return a / b;
        case F_ATAN: return atan(a);
        }
        return __builtin_nan("");
}

#define TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
        fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); return -1; } } while (0)

/* Device values of function `which` on n arguments (host arrays). */
extern "C" int probe_math(int which, const double * x, const double * y,
    double * out, size_t n)
{
        double *dx = nullptr, *dy = nullptr, *dout = nullptr;
        const size_t bytes = n * sizeof(double);
        TRY(hipMalloc(&dx, bytes));
        TRY(hipMalloc(&dy, bytes));
        TRY(hipMalloc(&dout, bytes));
        TRY(hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice));
        TRY(hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice));
        const unsigned block = 256;
        const unsigned grid = (unsigned)((n + block - 1) / block);
        hipLaunchKernelGGL(k_math, dim3(grid), dim3(block), 0, 0, which, dx,
            dy, dout, n);
        TRY(hipGetLastError());
        TRY(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost));
        TRY(hipFree(dx));
        TRY(hipFree(dy));
        TRY(hipFree(dout));
        return 0;
}

/* Host C library values, for the comparison. */
extern "C" void probe_host_math(int which, const double * x, const double * y,
    double * out, size_t n)
{
        for (size_t i = 0; i < n; i++) out[i] = host_value(which, x[i], y[i]);
}

/* ------------------------------------------------------------------------
 * Engine routines on a PTAB image copied to the device as is: the views
 * point at its sections (the layout the library uploads).
 * ------------------------------------------------------------------------ */
struct DeviceTables {
        std::vector<uint8_t> image;
        uint8_t * d_image = nullptr;
        int16_t * d_coarse = nullptr;
        pg::PhysicsView view;
        pg::ModelView mv;
};
static DeviceTables g_t;

static const void * section(const char * name, bool device)
{
        const ptab_header * h = (const ptab_header *)g_t.image.data();
        const ptab_entry * e =
            (const ptab_entry *)(g_t.image.data() + sizeof(ptab_header));
        for (uint32_t i = 0; i < h->n_entries; i++)
                if (!strncmp(e[i].name, name, PTAB_NAME_LEN))
                        return (device ? g_t.d_image : g_t.image.data()) +
                            e[i].offset;
        fprintf(stderr, "probe: missing section %s\n", name);
        return nullptr;
}

/* Loads the PTAB file at `path` for the probes below. */
extern "C" int probe_load(const char * path)
{
        FILE * f = fopen(path, "rb");
        if (!f) return -1;
        fseek(f, 0, SEEK_END);
        const long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        g_t.image.resize(n);
        if (fread(g_t.image.data(), 1, n, f) != (size_t)n) return -1;
        fclose(f);
        if (g_t.d_image) (void)hipFree(g_t.d_image);
        if (g_t.d_coarse) (void)hipFree(g_t.d_coarse);
        TRY(hipMalloc(&g_t.d_image, n));
        TRY(hipMemcpy(g_t.d_image, g_t.image.data(), n, hipMemcpyHostToDevice));
        const ptab_header * h = (const ptab_header *)g_t.image.data();
        pg::PhysicsView & v = g_t.view;
        memset(&v, 0, sizeof(v));
        v.n_energies = h->imeta[PTAB_N_ENERGIES];
        v.n_materials = h->imeta[PTAB_N_MATERIALS];
        v.n_elements = h->imeta[PTAB_N_ELEMENTS];
        v.n_components = h->imeta[PTAB_N_COMPONENTS];
        v.n_table_dcs = h->imeta[PTAB_N_TABLE_DCS];
        v.max_components = h->imeta[PTAB_MAX_COMPONENTS];
        v.mass = h->fmeta[PTAB_MASS];
        v.ctau = h->fmeta[PTAB_CTAU];
        v.cutoff = h->fmeta[PTAB_CUTOFF];
#define D(f, s) v.f = (const double *)section(s, true)
        D(K, "K"); D(K_dX, "K_dX"); D(K_dNI_in, "K_dNI_in");
        D(K_dNI_el, "K_dNI_el"); D(X, "X"); D(X_dK, "X_dK"); D(X_dT, "X_dT");
        D(T, "T"); D(T_dK, "T_dK"); D(dE, "dE"); D(dE_dK, "dE_dK");
        D(Omega, "Omega"); D(Omega_dK, "Omega_dK"); D(NI_el, "NI_el");
        D(NI_el_dK, "NI_el_dK"); D(NI_in, "NI_in"); D(NI_in_dK, "NI_in_dK");
        D(CS, "CS"); D(CS_dK, "CS_dK"); D(CSf, "CSf"); D(CSf_dK, "CSf_dK");
        D(CSn, "CSn"); D(CSn_dK, "CSn_dK"); D(Kt, "Kt"); D(a_max, "a_max");
        D(b_max, "b_max"); D(Mu0, "Mu0"); D(Mu0_dK, "Mu0_dK"); D(Lb, "Lb");
        D(Lb_dK, "Lb_dK"); D(Ms1, "Ms1"); D(Ms1_dK, "Ms1_dK");
        D(material_density, "material_density");
        D(component_fraction, "component_fraction");
        D(element_Z, "element_Z"); D(element_A, "element_A");
        D(element_I, "element_I"); D(element_nscr, "element_n_screening");
        D(element_amplitude, "element_amplitude");
        D(element_exponent, "element_exponent");
        D(element_rn, "element_nuclear_radius");
        D(element_lrad, "element_radiation_log");
#undef D
        v.DCS = (const float *)section("DCS", true);
        v.DCS_x = (const float *)section("DCS_x", true);
        v.DCS_envelope = (const float *)section("DCS_envelope", true);
        v.elements_in = (const int *)section("elements_in", true);
        v.component_element = (const int *)section("component_element", true);
        v.element_index = (const int *)section("element_index", true);
        v.material_component0 = nullptr;
        std::vector<int16_t> coarse(4096, 0);
        v.k_nbins = pg::k_coarse_build((const double *)section("K", false),
            v.n_energies, coarse.data(), (int)coarse.size(), &v.k_bin0);
        TRY(hipMalloc(&g_t.d_coarse, coarse.size() * sizeof(int16_t)));
        TRY(hipMemcpy(g_t.d_coarse, coarse.data(),
            coarse.size() * sizeof(int16_t), hipMemcpyHostToDevice));
        v.k_coarse = v.k_nbins ? g_t.d_coarse : nullptr;
        pg::ModelView & m = g_t.mv;
        m.z13 = (const double *)section("element_z13", true);
        m.zm13 = (const double *)section("element_zm13", true);
        m.z23 = (const double *)section("element_z23", true);
        m.dn = (const double *)section("element_dn", true);
        m.lrad = v.element_lrad;
        m.cst = (const double *)section("constants", true);
        m.gq9_x = (const double *)section("gq9_x", true);
        m.gq9_w = (const double *)section("gq9_w", true);
        m.gq12_x = (const double *)section("gq12_x", true);
        m.gq12_w = (const double *)section("gq12_w", true);
        return 0;
}

/* which: 0 exact DCS (dcs_exact), 1 tabulated DCS (dcs_evaluate), 2 dilog
 * of a[i], 3 polar angle of a DEL (random stream (seed, i)), draws in
 * draws[i]. */
__global__ void k_functions(pg::PhysicsView p, pg::ModelView mv, int which,
    int ip, int iel, const double * a, const double * b, double * out,
    uint32_t * draws, uint64_t seed, size_t n)
{
        const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n) return;
        if (which == 0) {
                out[i] = pg::dcs_exact(p, mv, ip, iel, a[i], b[i]);
        } else if (which == 1) {
                pg::Cache c;
                c.reset(p.n_energies);
                out[i] = pg::dcs_evaluate(p, mv, c, ip, iel, a[i], b[i]);
        } else if (which == 2) {
                out[i] = pg::dilog(a[i]);
        } else {
                pg::Rng rng;
                rng.init(seed, i, pg::STREAM_TRANSPORT, 0);
                out[i] = pg::polar_sample(p, rng, ip, iel, a[i], b[i]);
                draws[i] = rng.count;
        }
}

extern "C" int probe_functions(int which, int ip, int iel, const double * a,
    const double * b, double * out, uint32_t * draws, uint64_t seed,
    size_t n)
{
        double *da = nullptr, *db = nullptr, *dout = nullptr;
        uint32_t * ddr = nullptr;
        const size_t bytes = n * sizeof(double);
        TRY(hipMalloc(&da, bytes));
        TRY(hipMalloc(&db, bytes));
        TRY(hipMalloc(&dout, bytes));
        TRY(hipMalloc(&ddr, n * sizeof(uint32_t)));
        TRY(hipMemcpy(da, a, bytes, hipMemcpyHostToDevice));
        TRY(hipMemcpy(db, b, bytes, hipMemcpyHostToDevice));
        const unsigned block = 64;
        hipLaunchKernelGGL(k_functions, dim3((unsigned)((n + block - 1) / block)),
            dim3(block), 0, 0, g_t.view, g_t.mv, which, ip, iel, da, db, dout,
            ddr, seed, n);
        TRY(hipGetLastError());
        TRY(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost));
        if (draws)
                TRY(hipMemcpy(draws, ddr, n * sizeof(uint32_t),
                    hipMemcpyDeviceToHost));
        TRY(hipFree(da));
        TRY(hipFree(db));
        TRY(hipFree(dout));
        TRY(hipFree(ddr));
        return 0;
}
