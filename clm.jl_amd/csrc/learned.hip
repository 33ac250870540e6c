// learned.hip — first-substep predictor for the moisture-form Richards step:
// column features, a 24-64-64-1 GELU MLP (FP32) for inference, its loss gradient
// and the Adam update. No counterpart in the reference (CLM.jl starts every
// column at dtsub = dtime, src/biogeophys/soil_water_movement.jl:1259).
//
// Feature extraction and the plain forward pass run one lane per column; their
// weights are read at wavefront-uniform addresses (the loops are fully unrolled
// over constant indices), so they come through scalar loads and the scalar
// cache. The dtsub0 predictor runs the network on the matrix cores for 16
// columns per wavefront (mlp_wave.hpp), exactly as the Richards kernel does when
// the inference is fused into its column load.
//
// Training: clmsoil_mlp_grad gives one wavefront a tile of 64 samples. Each lane
// runs forward and backward for its sample and leaves its activations and deltas
// in LDS; the wavefront then reduces the tile's parameter gradients (each lane
// owns one hidden unit's rows) and writes one row of partial sums. The Adam
// kernel sums the rows in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/clmsoil.h"
#include "../../include/clmsoil_learned.h"
#include "common.hpp"
#include "mlp_wave.hpp"
#include "state.hpp"

namespace clmsoil {
namespace {

constexpr int NIN = CLMSOIL_MLP_NIN, NH = CLMSOIL_MLP_NHID, NP = CLMSOIL_MLP_NPARAM;
constexpr int OW1 = 0, OB1 = NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH;
static_assert(OB3 + 1 == NP, "parameter packing");
constexpr int TILE = CLMSOIL_MLP_GRAD_TILE;
constexpr int LDP = TILE + 1;   // padded LDS row: lane s and row i hit different banks

__device__ __forceinline__ float gelu(float x) {
    const float k = 0.7978845608028654f;   // sqrt(2/pi)
    const float u = k * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_d(float x) {
    const float k = 0.7978845608028654f;
    const float x2 = x * x;
    const float th = tanhf(k * (x + 0.044715f * x * x2));
    const float du = k * (1.0f + 3.0f * 0.044715f * x2);
    return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * du;
}

struct FeatParams {
    clmsoil_mf_args a;
    int nlevsno, nlevsoi;
    double dtime;
    const float *norm;
};

// Raw (or standardised) features of column c.
__device__ __forceinline__ void column_features(const FeatParams &F, int64_t c, float f[NIN]) {
    const clmsoil_mf_args &a = F.a;
    const int64_t ld = a.ld;
    int n = (int)a.nbedrock[c];
    n = n < 0 ? 0 : (n > F.nlevsoi ? F.nlevsoi : n);
#pragma unroll
    for (int j = 1; j <= MAX_NLEVSOI; ++j) {
        float s = 0.0f;
        if (j <= n) {
            const int64_t os = c + (int64_t)(F.nlevsno + j - 1) * ld;
            const double vwc = jl_max(a.h2osoi_liq[os], 1.0e-6) / (a.dz[os] * DENH2O);
            double v = vwc / a.watsat[c + (int64_t)(j - 1) * ld];
            v = jl_min(v, 1.0);
            v = jl_max(0.01, v);
            s = (float)v;
        }
        f[j - 1] = s;
    }
    f[20] = (float)log(a.hksat[c]);
    f[21] = (float)a.bsw[c];
    f[22] = (float)a.qflx_infl[c];
    f[23] = (float)F.dtime;
    if (F.norm) {
#pragma unroll
        for (int i = 0; i < NIN; ++i) f[i] = (f[i] - F.norm[i]) / F.norm[NIN + i];
    }
}

__device__ __forceinline__ float mlp_eval(const float *__restrict__ P, const float x[NIN]) {
    float h1[NH];
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        float z = P[OB1 + i];
#pragma unroll
        for (int f = 0; f < NIN; ++f) z += P[OW1 + i * NIN + f] * x[f];
        h1[i] = gelu(z);
    }
    float y = P[OB3];
#pragma unroll 4
    for (int i = 0; i < NH; ++i) {
        float z = P[OB2 + i];
#pragma unroll
        for (int j = 0; j < NH; ++j) z += P[OW2 + i * NH + j] * h1[j];
        y += P[OW3 + i] * gelu(z);
    }
    return y;
}

__global__ __launch_bounds__(256) void features_kernel(FeatParams F, float *x, int64_t ldx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= F.a.nc) return;
    float f[NIN];
    if (F.a.mask && !F.a.mask[c]) {
#pragma unroll
        for (int i = 0; i < NIN; ++i) f[i] = 0.0f;
    } else {
        column_features(F, c, f);
    }
#pragma unroll
    for (int i = 0; i < NIN; ++i) x[(int64_t)i * ldx + c] = f[i];
}

// Prediction for 16 columns per wavefront, 4 lanes per column (the Richards
// kernel's layout): the features of mlp_column_features, then mlp_wave16, as the
// Richards kernel's fused inference does, so dtsub0 here equals the fused first
// substep bit for bit (before its clamp).
__global__ __launch_bounds__(256) void predict_kernel(FeatParams F, const float *__restrict__ P,
                                                      double *dtsub0) {
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    const int k = (int)(threadIdx.x & 3);
    const clmsoil_mf_args &a = F.a;
    const bool live = c < a.nc && (!a.mask || a.mask[c]);
    int n = live ? (int)a.nbedrock[c] : 0;
    n = n < 0 ? 0 : (n > F.nlevsoi ? F.nlevsoi : n);
    float fe[6];
    mlp_column_features(a.h2osoi_liq, a.dz, a.watsat, a.hksat, a.bsw, a.qflx_infl, a.ld, F.nlevsno,
                        c, n, k, live, F.dtime, F.norm, fe);
    const float y = mlp_wave16(P, fe[0], fe[1], fe[2], fe[3], fe[4], fe[5]);   // every lane of the wavefront
    if (live && k == 0) {
        const double lr = (double)F.norm[2 * NIN] + (double)F.norm[2 * NIN + 1] * (double)y;
        dtsub0[c] = F.dtime * exp(lr);
    }
}

__global__ __launch_bounds__(256) void forward_kernel(const float *__restrict__ P,
                                                      const float *x, int64_t n, int64_t ldx,
                                                      float *y) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float f[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) f[i] = x[(int64_t)i * ldx + s];
    y[s] = mlp_eval(P, f);
}

__global__ __launch_bounds__(TILE) void grad_kernel(const float *__restrict__ P, const float *x,
                                                    const float *t, int64_t n, int64_t ldx,
                                                    float inv_total, float *partial,
                                                    float *loss_partial) {
    __shared__ float A[NH][LDP];     // h2, then d2
    __shared__ float B[NH][LDP];     // h1
    __shared__ float D1[NH][LDP];    // layer-1 GELU slope, then d1
    __shared__ float X[NIN][LDP];    // x
    __shared__ float DY[TILE], LS[TILE];
    const int s = (int)threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * TILE + s;
    const bool ok = g < n;

    float xv[NIN];
#pragma unroll
    for (int f = 0; f < NIN; ++f) xv[f] = ok ? x[(int64_t)f * ldx + g] : 0.0f;
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        float z = P[OB1 + i];
#pragma unroll
        for (int f = 0; f < NIN; ++f) z += P[OW1 + i * NIN + f] * xv[f];
        B[i][s] = ok ? gelu(z) : 0.0f;
        D1[i][s] = gelu_d(z);   // layer-1 GELU slope, multiplied in below
    }
    // lane s reads only its own column of B, so no barrier is needed yet
    float z2[NH];
    float y = P[OB3];
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        float z = P[OB2 + i];
#pragma unroll
        for (int j = 0; j < NH; ++j) z += P[OW2 + i * NH + j] * B[j][s];
        z2[i] = z;
        const float h2 = gelu(z);
        A[i][s] = ok ? h2 : 0.0f;
        y += P[OW3 + i] * h2;
    }
    const float e = ok ? (y - t[g]) : 0.0f;
    const float dy = 2.0f * e * inv_total;
    DY[s] = dy;
    LS[s] = e * e;
    // backward through the two hidden layers (per sample)
#pragma unroll
    for (int i = 0; i < NH; ++i) z2[i] = dy * P[OW3 + i] * gelu_d(z2[i]);   // d2
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < NH; ++i) acc += P[OW2 + i * NH + j] * z2[i];
        D1[j][s] = ok ? acc * D1[j][s] : 0.0f;
    }
#pragma unroll
    for (int f = 0; f < NIN; ++f) X[f][s] = xv[f];
    __syncthreads();

    // ---- tile reduction: lane s owns hidden unit s ----
    float *out = partial + (int64_t)blockIdx.x * NP;
    {
        float w3 = 0.0f;
        for (int k = 0; k < TILE; ++k) w3 += DY[k] * A[s][k];
        out[OW3 + s] = w3;
        if (s == 0) {
            float b3 = 0.0f, l = 0.0f;
            for (int k = 0; k < TILE; ++k) { b3 += DY[k]; l += LS[k]; }
            out[OB3] = b3;
            loss_partial[blockIdx.x] = l * inv_total;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NH; ++i) A[i][s] = z2[i];   // d2 of sample s
    __syncthreads();
    {
        float acc[NH];
#pragma unroll
        for (int j = 0; j < NH; ++j) acc[j] = 0.0f;
        float b2 = 0.0f;
        for (int k = 0; k < TILE; ++k) {
            const float d = A[s][k];
            b2 += d;
#pragma unroll
            for (int j = 0; j < NH; ++j) acc[j] += d * B[j][k];
        }
#pragma unroll
        for (int j = 0; j < NH; ++j) out[OW2 + s * NH + j] = acc[j];
        out[OB2 + s] = b2;
    }
    {
        float acc[NIN];
#pragma unroll
        for (int f = 0; f < NIN; ++f) acc[f] = 0.0f;
        float b1 = 0.0f;
        for (int k = 0; k < TILE; ++k) {
            const float d = D1[s][k];
            b1 += d;
#pragma unroll
            for (int f = 0; f < NIN; ++f) acc[f] += d * X[f][k];
        }
#pragma unroll
        for (int f = 0; f < NIN; ++f) out[OW1 + s * NIN + f] = acc[f];
        out[OB1 + s] = b1;
    }
}

__global__ __launch_bounds__(256) void adam_kernel(float *P, float *m, float *v,
                                                   const float *partial, int64_t nparts,
                                                   float lr, float b1, float b2, float eps,
                                                   float bc1, float bc2, float *gout) {
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= NP) return;
    float g = 0.0f;
    for (int64_t b = 0; b < nparts; ++b) g += partial[b * NP + p];
    if (gout) gout[p] = g;
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / bc1, vh = vv / bc2;
    P[p] = P[p] - lr * mh / (sqrtf(vh) + eps);
}

int feat_params(const clmsoil_mf_args *a, const clmsoil_levels *lev, double dtime,
                const float *norm, FeatParams *F) {
    if (!a || !lev) return CLMSOIL_ERR_NULL;
    int st = validate_levels(lev);
    if (st) return st;
    if (a->nc < 0 || a->ld < a->nc) return CLMSOIL_ERR_SIZE;
    if (!(dtime > 0.0) || !std::isfinite(dtime)) return CLMSOIL_ERR_ARG;
    if (a->nc > 0 && (!a->nbedrock || !a->h2osoi_liq || !a->dz || !a->watsat || !a->hksat ||
                      !a->bsw || !a->qflx_infl))
        return CLMSOIL_ERR_NULL;
    F->a = *a;
    F->nlevsno = lev->nlevsno;
    F->nlevsoi = lev->nlevsoi;
    F->dtime = dtime;
    F->norm = norm;
    return CLMSOIL_OK;
}

int launched() { return hipGetLastError() == hipSuccess ? CLMSOIL_OK : CLMSOIL_ERR_HIP; }

}  // namespace
}  // namespace clmsoil

using namespace clmsoil;

extern "C" {

int clmsoil_richards_features(const clmsoil_mf_args *a, const clmsoil_levels *lev, double dtime,
                              const float *norm, float *x, int64_t ldx, void *stream) {
    FeatParams F;
    int st = feat_params(a, lev, dtime, norm, &F);
    if (st) return st;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!x) return CLMSOIL_ERR_NULL;
    if (ldx < a->nc) return CLMSOIL_ERR_SIZE;
    const int64_t grid = (a->nc + 255) / 256;
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                       F, x, ldx);
    return launched();
}

int clmsoil_mlp_predict_dtsub0(const clmsoil_mf_args *a, const clmsoil_levels *lev, double dtime,
                               const float *params, const float *norm, double *dtsub0,
                               void *stream) {
    FeatParams F;
    int st = feat_params(a, lev, dtime, norm, &F);
    if (st) return st;
    if (a->nc == 0) return CLMSOIL_OK;
    if (!params || !norm) return CLMSOIL_ERR_NOT_TRAINED;
    if (!dtsub0) return CLMSOIL_ERR_NULL;
    const int64_t grid = (a->nc * 4 + 255) / 256;
    hipLaunchKernelGGL(predict_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                       F, params, dtsub0);
    return launched();
}

int clmsoil_mlp_forward(const float *params, const float *x, int64_t n, int64_t ldx, float *y,
                        void *stream) {
    if (n < 0 || ldx < n) return CLMSOIL_ERR_SIZE;
    if (n == 0) return CLMSOIL_OK;
    if (!params || !x || !y) return CLMSOIL_ERR_NULL;
    const int64_t grid = (n + 255) / 256;
    hipLaunchKernelGGL(forward_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                       params, x, n, ldx, y);
    return launched();
}

int clmsoil_mlp_grad(const float *params, const float *x, const float *t, int64_t n, int64_t ldx,
                     int64_t n_total, float *partial, float *loss_partial, void *stream) {
    if (n < 0 || ldx < n || n_total < n || n_total < 1) return CLMSOIL_ERR_SIZE;
    if (n == 0) return CLMSOIL_OK;
    if (!params || !x || !t || !partial || !loss_partial) return CLMSOIL_ERR_NULL;
    const int64_t grid = (n + TILE - 1) / TILE;
    hipLaunchKernelGGL(grad_kernel, dim3((unsigned)grid), dim3(TILE), 0, (hipStream_t)stream,
                       params, x, t, n, ldx, (float)(1.0 / (double)n_total), partial,
                       loss_partial);
    return launched();
}

int clmsoil_mlp_adam(float *params, float *m, float *v, const float *partial, int64_t nparts,
                     int64_t t, float lr, float beta1, float beta2, float eps, float *grad_out,
                     void *stream) {
    if (!params || !m || !v || !partial) return CLMSOIL_ERR_NULL;
    if (nparts < 1 || t < 1) return CLMSOIL_ERR_ARG;
    if (!(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return CLMSOIL_ERR_ARG;
    const float bc1 = (float)(1.0 - std::pow((double)beta1, (double)t));
    const float bc2 = (float)(1.0 - std::pow((double)beta2, (double)t));
    hipLaunchKernelGGL(adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       params, m, v, partial, nparts, lr, beta1, beta2, eps, bc1, bc2, grad_out);
    return launched();
}

}  // extern "C"
