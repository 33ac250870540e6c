// common.hpp — shared device helpers for the clmsoil HIP kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clmsoil.h"

// The project's FP64 log/log10/exp/pow (clm_math.h) as device functions with
// device-resident tables; the CPU restatement compiles the same source.
#define CLM_MATH_FN __device__ static __forceinline__
#define CLM_MATH_TAB __device__ static const
#include "clm_math.h"

namespace clmsoil {

// Physical constants, src/constants/varcon.jl:15-70.
constexpr double GRAV = 9.80616, CPLIQ = 4188.0, CPICE = 2117.27, HFUS = 3.337e5,
                 DENH2O = 1000.0, DENICE = 917.0, TFRZ = 273.15, TKAIR = 0.023,
                 TKICE = 2.290, TKWAT = 0.57, CAPR = 0.34, CNFAC = 0.5,
                 THK_BEDROCK = 3.0, CSOL_BEDROCK = 2.0e6;
constexpr double M_TO_MM = 1.0e3;          // soil_water_movement.jl:37
constexpr double THIN_SFCLAYER = 1.0e-6;   // soil_temperature.jl:26

// Minimum waves per SIMD the fused kernels are compiled for (register cap).
#ifndef CLMSOIL_RICHARDS_WAVES
#define CLMSOIL_RICHARDS_WAVES 2
#endif
#ifndef CLMSOIL_HEAT_WAVES
#define CLMSOIL_HEAT_WAVES 2
#endif

// MI355X (gfx950) has 8 accelerator complex dies (XCDs); workgroups are
// dispatched to them round-robin by workgroup id.
constexpr int NUM_XCD = 8;

// Compile-time layout limits (CLM5 20SL_8.5m, varpar.jl:112-149).
constexpr int MAX_NLEVSOI = 20;
constexpr int MAX_NLEVGRND = 25;
constexpr int MAX_NLEVSNO = 12;

// Julia Base.max / Base.min on Float64: the result of a diff-based select, so a
// NaN operand propagates, max(-0.0, 0.0) = 0.0 and min(0.0, -0.0) = -0.0.
__host__ __device__ __forceinline__ double jl_max(double x, double y) {
    double d = x - y;
    double am = __builtin_signbit(d) ? y : x;
    return (x != x || y != y) ? d : am;
}
__host__ __device__ __forceinline__ double jl_min(double x, double y) {
    double d = x - y;
    double am = __builtin_signbit(d) ? x : y;
    return (x != x || y != y) ? d : am;
}

// Transcendentals through the fast paths of clm_math.h when every active lane
// of the wavefront has an argument in their domain (wavefront-uniform branch),
// through the general functions otherwise: the same bits either way.
__device__ __forceinline__ bool wave_all(bool p) { return __ballot(!p) == 0ull; }
__device__ __forceinline__ double pow_w(double x, double y) {
    if (__builtin_expect(wave_all(clm_pow_pos_ok(x, y)), 1)) return clm_pow_pos(x, y);
    return clm_pow(x, y);
}
// b^y for a fixed base b (positive, normal, finite, b != 1) with Lb = clm_log_ext(b)
__device__ __forceinline__ double pow_base_w(clm_dd Lb, double b, double y) {
    if (__builtin_expect(wave_all(fabs(y) <= 0x1.fffffffffffffp+1023), 1))
        return clm_pow_base(Lb, y);
    return clm_pow(b, y);
}
__device__ __forceinline__ double log10_w(double x) {
    if (__builtin_expect(wave_all(clm_pow_pos_ok(x, 0.0)), 1)) return clm_log10_pos(x);
    return clm_log10(x);
}

// A pointer field of a kernel's by-value argument block, loaded from the kernarg
// segment where it is needed (one scalar load). A kernel with dozens of pointer
// arguments otherwise holds all of them in SGPRs from its entry on; past ~100
// SGPRs the compiler spills them to VGPR lanes and reads every use back with two
// v_readlane. OFF is the field's byte offset in the block, which must be the
// kernel's first argument.
// The value is cast through a global-address-space pointer: an asm result has no
// known address space, and accesses through a generic pointer compile to flat
// instructions, which also count against lgkmcnt (so every LDS, DPP or scalar
// wait waits for them too) and take a full 64-bit VGPR address.
template <typename T, unsigned OFF>
__device__ __forceinline__ T karg_ptr() {
    static_assert(sizeof(T) == 8 && OFF % 4 == 0, "64-bit field at a dword offset");
    uint64_t v;
    asm volatile("s_load_dwordx2 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)"
                 : "=s"(v)
                 : "s"(__builtin_amdgcn_kernarg_segment_ptr()), "i"(OFF));
    using E = __attribute__((address_space(1))) char;
    return (T)(char *)(E *)v;
}

// Column-major [c, j] access with 1-based level j (the reference's index).
struct Mat {
    double *p;
    int64_t ld;
    __device__ __forceinline__ double &operator()(int64_t c, int j) const {
        return p[c + (int64_t)(j - 1) * ld];
    }
};
struct CMat {
    const double *p;
    int64_t ld;
    __device__ __forceinline__ double operator()(int64_t c, int j) const {
        return __builtin_nontemporal_load(&p[c + (int64_t)(j - 1) * ld]);
    }
};

// Broadcast lane `k` of each group of G consecutive lanes (G in {2, 4}) to the
// whole group with a DPP quad permutation; G == 1 is the identity.
template <int G>
__device__ __forceinline__ double group_bcast(double v, int k);

template <>
__device__ __forceinline__ double group_bcast<1>(double v, int) { return v; }

// Quad permutations read a valid lane for every lane, so the DPP move needs no
// "old" value (mov_dpp: none is materialised, unlike update_dpp's).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// quad_perm encoding: sel0 | sel1<<2 | sel2<<4 | sel3<<6
template <>
__device__ __forceinline__ double group_bcast<2>(double v, int k) {
    // pairs (0,1) and (2,3) of each quad
    return k == 0 ? dpp_f64<(0 | 0 << 2 | 2 << 4 | 2 << 6)>(v)
                  : dpp_f64<(1 | 1 << 2 | 3 << 4 | 3 << 6)>(v);
}
template <>
__device__ __forceinline__ double group_bcast<4>(double v, int k) {
    switch (k) {
        case 0: return dpp_f64<0x00>(v);
        case 1: return dpp_f64<0x55>(v);
        case 2: return dpp_f64<0xAA>(v);
        default: return dpp_f64<0xFF>(v);
    }
}

// Value held by the next lane of the group ((k+1) mod G) / previous lane.
template <int G>
__device__ __forceinline__ double group_next(double v);
template <int G>
__device__ __forceinline__ double group_prev(double v);
template <>
__device__ __forceinline__ double group_next<1>(double v) { return v; }
template <>
__device__ __forceinline__ double group_prev<1>(double v) { return v; }
template <>
__device__ __forceinline__ double group_next<2>(double v) {
    return dpp_f64<(1 | 0 << 2 | 3 << 4 | 2 << 6)>(v);
}
template <>
__device__ __forceinline__ double group_prev<2>(double v) {
    return dpp_f64<(1 | 0 << 2 | 3 << 4 | 2 << 6)>(v);
}
template <>
__device__ __forceinline__ double group_next<4>(double v) {
    return dpp_f64<(1 | 2 << 2 | 3 << 4 | 0 << 6)>(v);
}
template <>
__device__ __forceinline__ double group_prev<4>(double v) {
    return dpp_f64<(3 | 0 << 2 | 1 << 4 | 2 << 6)>(v);
}

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
// Min of an int over the G lanes of a group.
template <int G>
__device__ __forceinline__ int group_min_i(int v) {
    if constexpr (G >= 2) v = min(v, dpp_i32<(1 | 0 << 2 | 3 << 4 | 2 << 6)>(v));
    if constexpr (G >= 4) v = min(v, dpp_i32<(2 | 3 << 2 | 0 << 4 | 1 << 6)>(v));
    return v;
}

// Max over the G lanes of a group.
template <int G>
__device__ __forceinline__ double group_max(double v) {
    if constexpr (G >= 2) v = jl_max(v, dpp_f64<(1 | 0 << 2 | 3 << 4 | 2 << 6)>(v));
    if constexpr (G >= 4) v = jl_max(v, dpp_f64<(2 | 3 << 2 | 0 << 4 | 1 << 6)>(v));
    return v;
}

// Device error word: kernels OR a bit in when a column hits an input the
// reference would reject (e.g. nbedrock out of range); the host reads it back.
enum DeviceError : unsigned {
    DEV_ERR_NBEDROCK = 1u,
    DEV_ERR_SNL = 2u,
    DEV_ERR_ITER = 4u,
    DEV_ERR_TAPE = 8u,
    DEV_ERR_SUBGRID = 16u,
    DEV_ERR_JWT = 32u,
    DEV_ERR_ADJ = 64u,
};

}  // namespace clmsoil
