/*
 * On-device tallies of transported histories: the flux estimator of a
 * backward (or forward) Monte Carlo run, sum over histories of
 * weight x surface flux model, with its sum of squares, and histograms of
 * the same estimator in kinetic energy (log bins) and cos(zenith) (linear
 * bins).
 *
 * The flux models are the sea level parameterisations of the reference's
 * examples (examples/pumas/flux.c): Gaisser's model with a constant charge
 * ratio (flux.c:39-65) and the GCCLY correction with Volkova's cos(theta*)
 * (flux.c:70-89). The estimator is the one of examples/pumas/straight.c
 * (a weighted mean over histories).
 *
 * Sums are deterministic: each thread accumulates its own histories (a
 * fixed set for a given batch size) in index order, a block adds its
 * threads' sums by a fixed pairwise tree, histogram bins are owned by one
 * thread each and scan the block's values in index order, blocks write
 * partial sums, and a second kernel adds the partials in block order.
 * Results therefore do not depend on scheduling, on other work running on
 * the device, or on the run.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pumas_gpu.h"
#include "status.h"

namespace {

constexpr int TALLY_BLOCK = 256;
constexpr int TALLY_GRID = 1024;

struct TallyArgs {
        const struct pumas_state * states;
        const enum pumas_event * events;
        unsigned long long n;
        int flux;
        int charge_fraction;
        double scale;
        int cos_from_direction;
        double cos_theta;
        unsigned int event_mask;
        int n_energy, n_cos;
        double log_emin, inv_dlog_e;
        double cos_min, inv_dcos;
        int n_out; /* 3 + 2 n_energy + 2 n_cos */
};

/* Fraction of the flux with the muon's charge, for the CMS charge ratio
 * 1.2766 (flux.c:39-52); 1 for a null charge. */
__device__ double charge_fraction(double charge)
{
        const double ratio = 1.2766;
        if (charge < 0.) return 1. / (1. + ratio);
        if (charge > 0.) return ratio / (1. + ratio);
        return 1.;
}

/* Gaisser's parameterisation, per m^2 s sr GeV (flux.c:57-65). */
__device__ double flux_gaisser(double cos_theta, double kinetic, double charge)
{
        const double e = kinetic + 0.10566;
        const double ec = 1.1 * e * cos_theta;
        const double r_pi = 1. + ec / 115.;
        const double r_k = 1. + ec / 850.;
        return 1.4E+03 * pow(e, -2.7) * (1. / r_pi + 0.054 / r_k) *
            charge_fraction(charge);
}

/* Volkova's cos(theta*) at the production height (flux.c:70-80). */
__device__ double cos_theta_star(double c)
{
        const double p0 = 0.102573, p1 = -0.068287, p2 = 0.958633,
                     p3 = 0.0407253, p4 = 0.817285;
        const double cs2 = (c * c + p0 * p0 + p1 * pow(c, p2) +
                               p3 * pow(c, p4)) /
            (1. + p0 * p0 + p1 + p3);
        return cs2 > 0. ? sqrt(cs2) : 0.;
}

/* Guan et al. (GCCLY) parameterisation (flux.c:83-89). */
__device__ double flux_gccly(double cos_theta, double kinetic, double charge)
{
        const double e = kinetic + 0.10566;
        const double cs = cos_theta_star(cos_theta);
        return pow(1. + 3.64 / (e * pow(cs, 1.29)), -2.7) *
            flux_gaisser(cs, kinetic, charge);
}

__global__ void __launch_bounds__(TALLY_BLOCK) k_tally(
    TallyArgs a, double * partial)
{
        __shared__ double s_v[TALLY_BLOCK];
        __shared__ int s_sel[TALLY_BLOCK], s_be[TALLY_BLOCK],
            s_bc[TALLY_BLOCK];
        const int j = threadIdx.x;
        double t_sum = 0., t_sum2 = 0., t_count = 0.;
        double e_sum = 0., e_sum2 = 0., c_sum = 0., c_sum2 = 0.;
        const unsigned long long chunks =
            (a.n + TALLY_BLOCK - 1) / TALLY_BLOCK;
        for (unsigned long long ch = blockIdx.x; ch < chunks;
             ch += gridDim.x) {
                const unsigned long long i = ch * TALLY_BLOCK + j;
                int sel = 0, be = -1, bc = -1;
                double v = 0.;
                if (i < a.n) {
                        const struct pumas_state * s = a.states + i;
                        sel = (a.event_mask == 0u) || (a.events == nullptr) ||
                            ((unsigned int)a.events[i] & a.event_mask);
                        if (sel) {
                                const double c = a.cos_from_direction ?
                                    -s->direction[2] :
                                    a.cos_theta;
                                const double q =
                                    a.charge_fraction ? s->charge : 0.;
                                double f = 1.;
                                if (a.flux == PUMAS_GPU_FLUX_GAISSER)
                                        f = flux_gaisser(c, s->energy, q);
                                else if (a.flux == PUMAS_GPU_FLUX_GCCLY)
                                        f = flux_gccly(c, s->energy, q);
                                v = s->weight * f * a.scale;
                                if ((a.n_energy > 0) && (s->energy > 0.)) {
                                        const double x = (log(s->energy) -
                                                             a.log_emin) *
                                            a.inv_dlog_e;
                                        if ((x >= 0.) && (x < a.n_energy))
                                                be = (int)x;
                                }
                                if (a.n_cos > 0) {
                                        const double x =
                                            (c - a.cos_min) * a.inv_dcos;
                                        if ((x >= 0.) && (x < a.n_cos))
                                                bc = (int)x;
                                }
                        }
                }
                if (sel) {
                        t_sum += v;
                        t_sum2 += v * v;
                        t_count += 1.;
                }
                if ((a.n_energy <= 0) && (a.n_cos <= 0)) continue;
                /* Histograms: thread j owns the j-th energy and cos bins
                 * and scans the block's values in index order. */
                s_v[j] = v;
                s_sel[j] = sel;
                s_be[j] = be;
                s_bc[j] = bc;
                __syncthreads();
                if (j < a.n_energy) {
                        for (int k = 0; k < TALLY_BLOCK; k++) {
                                if (s_be[k] != j) continue;
                                e_sum += s_v[k];
                                e_sum2 += s_v[k] * s_v[k];
                        }
                }
                if (j < a.n_cos) {
                        for (int k = 0; k < TALLY_BLOCK; k++) {
                                if (s_bc[k] != j) continue;
                                c_sum += s_v[k];
                                c_sum2 += s_v[k] * s_v[k];
                        }
                }
                __syncthreads();
        }
        /* Totals: the threads' sums added by a fixed pairwise tree. */
        __shared__ double s_t[3][TALLY_BLOCK];
        s_t[0][j] = t_sum;
        s_t[1][j] = t_sum2;
        s_t[2][j] = t_count;
        __syncthreads();
        for (int half = TALLY_BLOCK / 2; half > 0; half >>= 1) {
                if (j < half) {
                        s_t[0][j] += s_t[0][j + half];
                        s_t[1][j] += s_t[1][j + half];
                        s_t[2][j] += s_t[2][j + half];
                }
                __syncthreads();
        }
        double * out = partial + (size_t)blockIdx.x * a.n_out;
        if (j == 0) {
                out[0] = s_t[0][0];
                out[1] = s_t[1][0];
                out[2] = s_t[2][0];
        }
        if (j < a.n_energy) {
                out[3 + 2 * j] = e_sum;
                out[4 + 2 * j] = e_sum2;
        }
        if (j < a.n_cos) {
                out[3 + 2 * a.n_energy + 2 * j] = c_sum;
                out[4 + 2 * a.n_energy + 2 * j] = c_sum2;
        }
}

/* Sum of the block partials, in block order: one thread per output. */
__global__ void k_tally_reduce(const double * partial, int n_blocks,
    int n_out, double * out)
{
        const int o = blockIdx.x * blockDim.x + threadIdx.x;
        if (o >= n_out) return;
        double acc = 0.;
        for (int b = 0; b < n_blocks; b++) acc += partial[(size_t)b * n_out + o];
        out[o] = acc;
}

} /* namespace */

extern "C" enum pumas_return pumas_gpu_tally(const struct pumas_state * states,
    const enum pumas_event * events, size_t n,
    const struct pumas_gpu_tally_spec * spec,
    struct pumas_gpu_tally_result * result, void * stream_)
{
        pg_clear_error();
        if ((spec == NULL) || (result == NULL))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        if ((states == NULL) && (n > 0))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "no states (null)");
        if ((spec->flux < PUMAS_GPU_FLUX_NONE) ||
            (spec->flux > PUMAS_GPU_FLUX_GCCLY))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "unknown flux model %d", spec->flux);
        if ((spec->n_energy_bins < 0) ||
            (spec->n_energy_bins > PUMAS_GPU_TALLY_MAX_BINS) ||
            (spec->n_cos_bins < 0) ||
            (spec->n_cos_bins > PUMAS_GPU_TALLY_MAX_BINS))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "histograms hold at most %d bins",
                    PUMAS_GPU_TALLY_MAX_BINS);
        if ((spec->n_energy_bins > 0) &&
            !((spec->energy_min > 0.) &&
                (spec->energy_max > spec->energy_min)))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "bad energy range [%g, %g]", spec->energy_min,
                    spec->energy_max);
        if ((spec->n_cos_bins > 0) && !(spec->cos_max > spec->cos_min))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "bad cos(theta) range [%g, %g]", spec->cos_min,
                    spec->cos_max);

        TallyArgs a;
        memset(&a, 0x0, sizeof(a));
        a.states = states;
        a.events = events;
        a.n = n;
        a.flux = spec->flux;
        a.charge_fraction = spec->charge_fraction;
        a.scale = (spec->scale != 0.) ? spec->scale : 1.;
        a.cos_from_direction = spec->cos_from_direction;
        a.cos_theta = spec->cos_theta;
        a.event_mask = spec->event_mask;
        a.n_energy = spec->n_energy_bins;
        a.n_cos = spec->n_cos_bins;
        if (a.n_energy > 0) {
                a.log_emin = log(spec->energy_min);
                a.inv_dlog_e = a.n_energy /
                    (log(spec->energy_max) - a.log_emin);
        }
        if (a.n_cos > 0) {
                a.cos_min = spec->cos_min;
                a.inv_dcos = a.n_cos / (spec->cos_max - spec->cos_min);
        }
        a.n_out = 3 + 2 * a.n_energy + 2 * a.n_cos;

        hipStream_t stream = static_cast<hipStream_t>(stream_);
        const unsigned long long chunks = (n + TALLY_BLOCK - 1) / TALLY_BLOCK;
        int grid = (chunks < (unsigned long long)TALLY_GRID) ? (int)chunks :
                                                               TALLY_GRID;
        if (grid < 1) grid = 1;
        double * partial = nullptr;
        const size_t bytes = ((size_t)grid + 1) * a.n_out * sizeof(double);
        HIP_TRY(hipMallocAsync((void **)&partial, bytes, stream));
        double * out = partial + (size_t)grid * a.n_out;
        hipLaunchKernelGGL(k_tally, dim3(grid), dim3(TALLY_BLOCK), 0, stream,
            a, partial);
        hipLaunchKernelGGL(k_tally_reduce, dim3((a.n_out + 255) / 256),
            dim3(256), 0, stream, partial, grid, a.n_out, out);
        HIP_TRY(hipGetLastError());
        std::vector<double> host(a.n_out);
        HIP_TRY(hipMemcpyAsync(host.data(), out, a.n_out * sizeof(double),
            hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipFreeAsync(partial, stream));
        HIP_TRY(hipStreamSynchronize(stream));

        result->sum = host[0];
        result->sum2 = host[1];
        result->count = (uint64_t)host[2];
        if (result->energy)
                memcpy(result->energy, host.data() + 3,
                    2 * a.n_energy * sizeof(double));
        if (result->cos_theta)
                memcpy(result->cos_theta, host.data() + 3 + 2 * a.n_energy,
                    2 * a.n_cos * sizeof(double));
        return PUMAS_RETURN_SUCCESS;
}

extern "C" enum pumas_return pumas_gpu_tally_host(
    const struct pumas_state * states, const enum pumas_event * events,
    size_t n, const struct pumas_gpu_tally_spec * spec,
    struct pumas_gpu_tally_result * result)
{
        struct pumas_state * d_states = nullptr;
        enum pumas_event * d_events = nullptr;
        const size_t ns = n > 0 ? n : 1;
        HIP_TRY(hipMalloc(&d_states, ns * sizeof(*states)));
        if (n > 0)
                HIP_TRY(hipMemcpy(d_states, states, n * sizeof(*states),
                    hipMemcpyHostToDevice));
        if (events) {
                HIP_TRY(hipMalloc(&d_events, ns * sizeof(*events)));
                if (n > 0)
                        HIP_TRY(hipMemcpy(d_events, events,
                            n * sizeof(*events), hipMemcpyHostToDevice));
        }
        const enum pumas_return rc =
            pumas_gpu_tally(d_states, d_events, n, spec, result, nullptr);
        (void)hipFree(d_states);
        if (d_events) (void)hipFree(d_events);
        return rc;
}
