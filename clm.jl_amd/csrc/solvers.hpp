// solvers.hpp — launch interface of the batched tridiagonal / band solvers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clmsoil {

struct TridiagParams {
    double *u;
    const double *a, *b, *c, *r;
    const int64_t *jtop;
    const uint8_t *mask;
    int64_t ncols, nlevs, ld;
    unsigned *err;
};

struct BandParams {
    double *tvector;
    const double *bmatrix, *rvector;
    const int64_t *jtop, *jbot;
    const uint8_t *mask;
    int kl, ku, nlevsno_off;
    int64_t nc, nlev, ld;
    unsigned *err;
};

hipError_t launch_tridiag_multi(const TridiagParams &P, hipStream_t stream);
hipError_t launch_band_solve(const BandParams &P, hipStream_t stream);

}  // namespace clmsoil
