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

// Reverse-mode operands (see clmsoil_adjoint.h).
struct TridiagVjpParams {
    double *ubar;
    const double *u, *a, *b, *c;
    const int64_t *jtop;
    const uint8_t *mask;
    int64_t ncols, nlevs, ld;
    double *abar, *bbar, *cbar, *rbar;
    unsigned *err;
};

struct BandVjpParams {
    double *tbar;
    const double *tvector, *bmatrix;
    const int64_t *jtop, *jbot;
    const uint8_t *mask;
    int kl, ku, nlevsno_off;
    int64_t nc, nlev, ld;
    double *bmatrix_bar, *rvector_bar;
    unsigned *err;
};

hipError_t launch_tridiag_multi(const TridiagParams &P, hipStream_t stream);
hipError_t launch_tridiag_vjp(const TridiagVjpParams &P, hipStream_t stream);
hipError_t launch_band_vjp(const BandVjpParams &P, hipStream_t stream);
hipError_t launch_band_solve(const BandParams &P, hipStream_t stream);

}  // namespace clmsoil
