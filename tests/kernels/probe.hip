/*
 * Test infrastructure, not part of the library: device evaluations of
 * engine routines for function-level comparisons with the host C library
 * and the oracle (tests/test_gpu_libm.py, tests/test_gpu_functions.py).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "libm.h"

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
