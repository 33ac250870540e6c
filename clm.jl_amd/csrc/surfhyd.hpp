// surfhyd.hpp — launch interface of the surface hydrology kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clmsoil_surfhyd.h"

namespace clmsoil {

hipError_t launch_surface_hydrology(const clmsoil_surfhyd_args &a,
                                    const clmsoil_surfhyd_params &p, int nlevsno, double dtime,
                                    hipStream_t stream);

hipError_t launch_root_sink(const clmsoil_rootsink_args &a, int nlevsoi, unsigned *err,
                            hipStream_t stream);

}  // namespace clmsoil
