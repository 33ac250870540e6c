// zd09.hpp — launch interface of the Zeng-Decker 2009 soil water step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clmsoil {

struct ZdParams {
    int64_t nc, ld;
    const uint8_t *mask;
    const double *z, *dz, *zi;
    const double *watsat, *hksat, *bsw, *sucsat, *smpmin;
    double *h2osoi_liq;
    const double *h2osoi_ice, *h2osoi_vol, *t_soisno;
    const double *zwt, *qflx_infl, *qflx_rootsoi;
    double *icefrac, *smp_l, *hk_l, *qcharge, *qflx_deficit;
    int nlevsno, nlevsoi, nlevgrnd;
    double dtime, e_ice;
};

hipError_t launch_zengdecker2009(const ZdParams &P, int block, hipStream_t stream);

}  // namespace clmsoil
