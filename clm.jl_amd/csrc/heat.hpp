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
};

hipError_t launch_heat(const HeatParams &P, int G, int block, hipStream_t stream);

}  // namespace clmsoil
