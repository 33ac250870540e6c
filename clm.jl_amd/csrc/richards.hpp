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
    // first-substep network (packed weights and normalisation block of
    // clmsoil_learned.h), evaluated for each column as it is loaded; NULL = none
    const float *mlp, *mlp_norm;
    int nlevsno, nlevsoi;
    int lower_bc;
    int expensive;
    int aquifer_pass;       // run compute_qcharge!'s pass after the solve
    int max_iter;           // substep-iteration cap (hang guard)
    int refill_sixteenths;  // persistent launch: refill when this many 16ths of a
                            // wavefront's groups are idle (1..16)
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
// lds_fields: per-layer fields the G = 4 / 256-thread kernels keep in a
// lane-private LDS cache of the column (7: bsw, sucsat, root sink, smp, hk, khat,
// node distance; 5: the first five; 0: none, all in registers or memory).
hipError_t launch_richards(const RichardsParams &P, int G, int block, bool persistent,
                           int blocks_per_cu, int lds_fields, hipStream_t stream);
// The same step with the first-substep network fused into the column load
// (P.mlp and P.mlp_norm set): G = 4 lanes per column, 256-thread blocks.
hipError_t launch_richards_learned(const RichardsParams &P, bool persistent, int blocks_per_cu,
                                   hipStream_t stream);

}  // namespace clmsoil
