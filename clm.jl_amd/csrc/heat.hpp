// heat.hpp — launch interface of the fused snow+soil heat-step kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clmsoil.h"

namespace clmsoil {

struct HeatParams {
    clmsoil_heat_args a;     // caller's arrays (device pointers)
    int nlevsno, nlevgrnd;
    int snow_method;         // 1 Jordan1991, 2 Sturm1997
    double dtime;
    unsigned *err;           // device error word
    // Columns whose elimination needs a row interchange are finished by the
    // pivoting pass; the fast pass records them here (NULL = solve in place).
    int64_t *pivot_list;
    unsigned *pivot_count;
    double *dense_ws;        // PIVOT_GROUPS dense systems for the pivoting pass
};

// Grid of the pivoting pass and the dense workspace it needs.
constexpr int PIVOT_BLOCK = 64;
constexpr int PIVOT_GRID = 64;
constexpr int PIVOT_GROUPS = PIVOT_GRID * PIVOT_BLOCK / 4;
constexpr size_t DENSE_WS_BYTES = (size_t)PIVOT_GROUPS * (38 * 38 + 38) * sizeof(double);

hipError_t launch_heat(const HeatParams &P, int G, int block, hipStream_t stream);
hipError_t launch_heat_pivot(const HeatParams &P, hipStream_t stream);

}  // namespace clmsoil
