// ghf.hpp — launch interface of the ground heat flux kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/clmsoil_ghf.h"

namespace clmsoil {

struct GhfParams {
    clmsoil_ghf_args a;   // device pointers as passed by the caller
    int nlevsno;
    unsigned *err;
};

hipError_t launch_ground_heat_flux(const GhfParams &P, int block, hipStream_t stream);
hipError_t launch_soil_energy_balance(const clmsoil_seb_args &a, int nlevsno, int nlevgrnd,
                                      double dtime, unsigned *err, hipStream_t stream);

}  // namespace clmsoil
