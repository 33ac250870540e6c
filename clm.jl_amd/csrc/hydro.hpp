// hydro.hpp — launch interface of the soil water fractions and water table kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clmsoil_hydro.h"

namespace clmsoil {

hipError_t launch_soil_water_fractions(const clmsoil_fractions_args &a, int nlevsno, int nlevsoi,
                                       hipStream_t stream);
hipError_t launch_water_table(const clmsoil_water_table_args &a, int nlevsno, int nlevsoi,
                              double dtime, double aq_sp_yield_min, int recompute_frost_table,
                              hipStream_t stream);

// ops: 1 perched, 2 theta-based, 4 clamp to bedrock, 7 all three in that order.
hipError_t launch_theta_water_table(const clmsoil_theta_wt_args &a, int nlevsno, int nlevsoi,
                                    int ops, unsigned *err, hipStream_t stream);

}  // namespace clmsoil
