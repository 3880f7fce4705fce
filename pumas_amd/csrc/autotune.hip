/*
 * Autotuner of the transport launch: block size, resident blocks per CU and
 * the EHS batch size of the stepping wavefronts, timed on a sample of the
 * caller's workload. Results of a history do not depend on these settings
 * (they change the interleaving of histories, not their operations), so
 * the tuner only measures time.
 *
 * No counterpart in the reference library (one history per call).
 */
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/pumas_gpu.h"
#include "status.h"

namespace {

struct Shape {
        int block, per_cu;
};

/* Launch shapes at the stepping kernel's full residency (8 wavefronts per
 * CU, the limit of its 256 registers per lane). Fewer resident wavefronts
 * only look faster on samples so small that the tail of the launch (the
 * longest histories) dominates. One block per CU holds one LDS copy of the
 * step tables, which the FP64 ones need; two blocks hold two copies of the
 * FP32 ones; smaller blocks read the tables from HBM. */
const Shape SHAPES[] = { { 512, 1 }, { 256, 2 }, { 128, 4 }, { 64, 8 } };
const int EHS_BATCHES[] = { 8, 16, 24, 32 };

/* Timed runs per candidate; the fastest counts (run to run noise of a
 * sample of ~10^6 histories is a few per cent, the size of the differences
 * between neighbouring candidates). */
constexpr int REPEATS = 2;

} /* namespace */

extern "C" enum pumas_return pumas_gpu_autotune(
    const pumas_gpu_physics * physics, const struct pumas_gpu_context * context,
    const pumas_gpu_geometry * geometry, const struct pumas_state * states,
    size_t n, uint64_t seed, enum pumas_gpu_precision precision,
    struct pumas_gpu_options * best, double * steps_per_s, void * stream_)
{
        pg_clear_error();
        if ((physics == NULL) || (context == NULL) || (geometry == NULL) ||
            (states == NULL) || (best == NULL))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        if (n == 0)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "empty sample");
        hipStream_t stream = static_cast<hipStream_t>(stream_);
        struct pumas_state * work = nullptr;
        HIP_TRY(hipMallocAsync((void **)&work, n * sizeof(*states), stream));

        double best_rate = -1.;
        struct pumas_gpu_options o;
        enum pumas_return rc = PUMAS_RETURN_SUCCESS;
        for (const Shape & sh : SHAPES) {
                for (int eb : EHS_BATCHES) {
                        memset(&o, 0x0, sizeof(o));
                        o.precision = precision;
                        o.block_size = sh.block;
                        o.blocks_per_cu = sh.per_cu;
                        o.ehs_batch = eb;
                        for (int k = 0; k < REPEATS; k++) {
                                HIP_TRY(hipMemcpyAsync(work, states,
                                    n * sizeof(*states),
                                    hipMemcpyDeviceToDevice, stream));
                                struct pumas_gpu_stats st;
                                rc = pumas_gpu_transport(physics, context,
                                    geometry, work, n, seed, 0, nullptr, &o,
                                    &st, stream);
                                if (rc != PUMAS_RETURN_SUCCESS) break;
                                const double rate = (st.kernel_ms > 0.) ?
                                    st.steps / (st.kernel_ms * 1e-3) :
                                    0.;
                                if (rate > best_rate) {
                                        best_rate = rate;
                                        *best = o;
                                }
                        }
                        if (rc != PUMAS_RETURN_SUCCESS) break;
                }
                if (rc != PUMAS_RETURN_SUCCESS) break;
        }
        const hipError_t e = hipFreeAsync(work, stream);
        if (rc != PUMAS_RETURN_SUCCESS) return rc;
        HIP_TRY(e);
        HIP_TRY(hipStreamSynchronize(stream));
        if (steps_per_s) *steps_per_s = best_rate;
        return PUMAS_RETURN_SUCCESS;
}
