// ch4.hpp — launch interface of the methane and oxygen transport step.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clmsoil_ch4.h"

namespace clmsoil {

hipError_t launch_ch4_tran(const clmsoil_ch4_tran_args &a, const clmsoil_ch4_tran_params &p, double dtime,
                           unsigned *err, hipStream_t stream);

}  // namespace clmsoil
