// mlp_wave.hpp — the first-substep network (24 -> 64 GELU -> 64 GELU -> 1, FP32
// weights and accumulation) evaluated for the 16 columns of a wavefront on the
// matrix cores, with each column held by a group of 4 lanes (layer j of the
// column in lane (j-1) mod 4, slot (j-1) div 4, as in the Richards kernel).
//
// Both layers are v_mfma_f32_16x16x4f32 products with the 16 columns as the N
// dimension: layer 1 is 4 unit tiles x 6 K-steps over the 24 features, layer 2
// 4 unit tiles x 16 K-steps over the 64 hidden units; the output layer is a
// per-lane dot product of 16 terms and a sum over the four lane rows. The
// K-order of layer 2 follows the accumulator layout of layer 1, so its B
// operand is read straight from layer 1's output registers. Weights are read
// from memory at lane-dependent addresses (23 KB, cache resident).
//
// The standalone predictor and the Richards kernel's fused inference both call
// mlp_wave16 with the same feature layout, so they produce the same bits. The
// parameter block must be 16-byte aligned (the vector loads of biases and
// layer-2 weights; every section offset of the packing is a multiple of 4).
#pragma once
#include "common.hpp"
#include "../../include/clmsoil_learned.h"

namespace clmsoil {

namespace mlp {
constexpr int NIN = CLMSOIL_MLP_NIN, NH = CLMSOIL_MLP_NHID;
constexpr int OW1 = 0, OB1 = NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH;
}  // namespace mlp

typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// GELU, tanh form 0.5 x (1 + tanh(u)) = x / (1 + exp(-2u)),
// u = sqrt(2/pi) (x + 0.044715 x^3), with the hardware exp2 and reciprocal
// (within a few ulp of the training kernels' tanhf form; an overflowing exp
// gives x * 0 = -0 for large negative x, as the limit says).
__device__ __forceinline__ float mlp_gelu(float x) {
    const float k2 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;   // -2 sqrt(2/pi) log2(e)
    const float e = __builtin_amdgcn_exp2f(k2 * (x + 0.044715f * x * x * x));
    return x * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float lane_read_f(float v, int src_lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
}

// Standardised feature f = 4*i + k of a column from its raw value (lane k, slot i),
// with the hardware reciprocal of the scale.
__device__ __forceinline__ float mlp_standardise(const float *__restrict__ norm, int f, float raw) {
    return (raw - norm[f]) * __builtin_amdgcn_rcpf(norm[mlp::NIN + f]);
}

// Standardised features of column c for lane k of its group of 4, read from the
// column's state in memory: node saturations of soil layers 4i+k+1 (the
// reference's s2, soil_water_movement.jl:1273-1286; 0 below the bedrock layer n
// and for columns that are not `live`), and in slot 5 log ksat (k = 0) and b
// (k = 1) of layer 1, the infiltration flux (k = 2) and dtime (k = 3).
__device__ __forceinline__ void mlp_column_features(const double *liq, const double *dz,
                                                    const double *watsat, const double *hksat,
                                                    const double *bsw, const double *qinfl,
                                                    int64_t ld, int nlevsno, int64_t c, int n,
                                                    int k, bool live, double dtime,
                                                    const float *__restrict__ norm, float fe[6]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const int j = 4 * i + k + 1;
        float raw = 0.0f;
        if (live && j <= n) {
            const int64_t os = c + (int64_t)(nlevsno + j - 1) * ld;
            double v = jl_max(liq[os], 1.0e-6) / (1.0e3 * dz[os]);
            v = v / watsat[c + (int64_t)(j - 1) * ld];
            v = jl_min(v, 1.0);
            v = jl_max(0.01, v);
            raw = (float)v;
        }
        fe[i] = mlp_standardise(norm, 4 * i + k, raw);
    }
    float r5 = 0.0f;
    if (live) {
        if (k == 0) r5 = (float)log(hksat[c]);
        else if (k == 1) r5 = (float)bsw[c];
        else if (k == 2) r5 = (float)qinfl[c];
        else r5 = (float)dtime;
    }
    fe[5] = mlp_standardise(norm, 20 + k, r5);
}

// Network output of column m of the wavefront, returned in every lane of group m
// (lanes 4m .. 4m+3). Lane 4m + k passes fe[i] = standardised feature 4i + k of
// column m (i = 0..5). Every lane of the wavefront must call this (the matrix
// instructions combine the lanes); columns whose lanes hold garbage only spoil
// their own outputs.
__device__ __noinline__ float mlp_wave16(const float *__restrict__ P, float fe0, float fe1, float fe2,
                                        float fe3, float fe4, float fe5) {
    using namespace mlp;
    const float fe[6] = {fe0, fe1, fe2, fe3, fe4, fe5};
    const int l = (int)__lane_id();
    const int col = l & 15, q = l >> 4;
    // B operand of layer 1, K-step s: row q is feature 4s + q of column col, held
    // by lane 4*col + q in its slot s
    float xb[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) xb[s] = lane_read_f(fe[s], 4 * col + q);
    // Weights are loaded one tile at a time (scheduling barriers between the
    // tiles), so that few registers are in flight beside the caller's state.
    // layer 1: tile t holds units 16t + 4q + r (register r) of column col
    mlp_f4 h[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mlp_f4 acc = *reinterpret_cast<const mlp_f4 *>(P + OB1 + 16 * t + 4 * q);
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const float a = P[OW1 + (16 * t + col) * NIN + 4 * s + q];
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xb[s], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) h[t][r] = mlp_gelu(acc[r]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // layer 2: K-step (t, r) takes hidden unit 16t + 4q + r in row q, which is
    // register r of layer-1 tile t in this very lane; lane (col, q) reads the
    // four weights W2[16 t2 + col][16t + 4q .. 16t + 4q + 3] as one vector
    float y = 0.0f;
#pragma unroll
    for (int t2 = 0; t2 < 4; ++t2) {
        mlp_f4 acc = *reinterpret_cast<const mlp_f4 *>(P + OB2 + 16 * t2 + 4 * q);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const mlp_f4 a = *reinterpret_cast<const mlp_f4 *>(P + OW2 + (16 * t2 + col) * NH +
                                                               16 * t + 4 * q);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], h[t][r], acc, 0, 0, 0);
        }
        const mlp_f4 w3 = *reinterpret_cast<const mlp_f4 *>(P + OW3 + 16 * t2 + 4 * q);
#pragma unroll
        for (int r = 0; r < 4; ++r) y += w3[r] * mlp_gelu(acc[r]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // sum the four rows q of column col (lanes col, col+16, col+32, col+48)
    y += lane_read_f(y, l ^ 16);
    y += lane_read_f(y, l ^ 32);
    y += P[OB3];
    // column m's value from lane m to the lanes of group m
    return lane_read_f(y, l >> 2);
}

}  // namespace clmsoil
