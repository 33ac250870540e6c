// lake.hpp — launch interface of the lake temperature solve.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clmsoil_lake.h"

namespace clmsoil {

constexpr int LK_NMAX = 64;   // nlevsno + nlevlak + nlevgrnd at most

hipError_t launch_lake_temperature_solve(const clmsoil_lake_solve_args &a, int nlevsno,
                                         int nlevgrnd, double dtime, unsigned *err,
                                         hipStream_t stream);

}  // namespace clmsoil
