// math.hip — device evaluation of the library's FP64 log/log10/exp/pow
// (clm_math.h), of the kernels' wavefront-dispatched fast paths (common.hpp) and
// of its Julia max/min over arrays, so that tests can compare the device build of
// those functions with the host build the CPU restatement links.
#include "common.hpp"
#include "../../include/clmsoil_ext.h"

namespace clmsoil {
namespace {

__global__ void math_eval_kernel(int op, const double *x, const double *y, double *out,
                                 int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v;
    switch (op) {
        case CLMSOIL_MATH_POW: v = clm_pow(x[i], y[i]); break;
        case CLMSOIL_MATH_LOG: v = clm_log(x[i]); break;
        case CLMSOIL_MATH_LOG10: v = clm_log10(x[i]); break;
        case CLMSOIL_MATH_EXP: v = clm_exp(x[i]); break;
        case CLMSOIL_MATH_JL_MAX: v = jl_max(x[i], y[i]); break;
        case CLMSOIL_MATH_JL_MIN: v = jl_min(x[i], y[i]); break;
        case CLMSOIL_MATH_POW_W: v = pow_w(x[i], y[i]); break;
        case CLMSOIL_MATH_POW_BASE_W: {
            const double b = x[i];
            const bool ok = b >= 0x1p-1022 && b <= 0x1.fffffffffffffp+1023 && b != 1.0;
            v = ok ? pow_base_w(clm_log_ext(b), b, y[i]) : clm_pow(b, y[i]);
            break;
        }
        default: v = log10_w(x[i]); break;
    }
    out[i] = v;
}

}  // namespace
}  // namespace clmsoil

extern "C" int clmsoil_math_eval(int op, const double *x, const double *y, double *out,
                                 int64_t n, void *stream) {
    if (op < CLMSOIL_MATH_POW || op > CLMSOIL_MATH_LOG10_W) return CLMSOIL_ERR_ARG;
    if (n < 0) return CLMSOIL_ERR_SIZE;
    if (n == 0) return CLMSOIL_OK;
    const bool binary = op == CLMSOIL_MATH_POW || (op >= CLMSOIL_MATH_JL_MAX && op != CLMSOIL_MATH_LOG10_W);
    if (!x || !out || (binary && !y)) return CLMSOIL_ERR_NULL;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(clmsoil::math_eval_kernel, dim3((unsigned)grid), dim3(block), 0,
                       (hipStream_t)stream, op, x, y, out, n);
    return hipGetLastError() == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP;
}
