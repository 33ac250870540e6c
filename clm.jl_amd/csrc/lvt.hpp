// lvt.hpp — launch interface of the soil organic matter vertical transport.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clmsoil_lvt.h"

namespace clmsoil {

hipError_t launch_litter_vert_transp(const clmsoil_lvt_args &a, const clmsoil_lvt_params &p,
                                     double dtime, unsigned *err, hipStream_t stream);

}  // namespace clmsoil
