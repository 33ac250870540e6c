// richards.hpp — launch interface of the fused moisture-form Richards kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clmsoil {

struct RichardsParams {
    int64_t nc, ld;
    const uint8_t *mask;
    const int64_t *nbedrock;
    double *h2osoi_liq, *smp_l, *hk_l, *qcharge, *nsubsteps, *nrejected;
    const double *z, *zi, *dz, *qflx_rootsoi, *watsat, *hksat, *bsw, *sucsat, *icefrac;
    const double *zwt, *qflx_infl, *smpmin, *dtsub0;
    int nlevsno, nlevsoi;
    int lower_bc;
    int expensive;
    int aquifer_pass;       // run compute_qcharge!'s pass after the solve
    int max_iter;           // substep-iteration cap (hang guard)
    double dtime, e_ice, dtmin, xtu, xtl, verysmall;
    unsigned *err;          // device error word
    unsigned *call_err;     // the call's own error word (workspace), or NULL
    unsigned *work_counters;  // NUM_XCD zeroed counters (persistent launch), or NULL
    // record an input the reference would reject (bits OR-ed, never cleared here)
    __device__ __forceinline__ void flag(unsigned f) const {
        atomicOr(err, f);
        if (call_err) atomicOr(call_err, f);
    }
};

// G = lanes per column (1, 2 or 4); block = threads per workgroup; persistent
// selects the resident-grid launch whose groups refill from work counters.
// blocks_per_cu sets the persistent grid (0 = as many blocks as the occupancy
// calculator says fit on a CU).
hipError_t launch_richards(const RichardsParams &P, int G, int block, bool persistent,
                           int blocks_per_cu, hipStream_t stream);

}  // namespace clmsoil
