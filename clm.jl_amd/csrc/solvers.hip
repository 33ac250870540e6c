// solvers.hip — batched per-column tridiagonal and banded solves (gfx950).
//
// tridiagonal_multi!  (src/infrastructure/tridiagonal.jl:86-142): one lane per
//   column, forward sweep with the 1e-30 pivot guard and register-carried
//   recurrences; cp is kept in lane-private storage for the back substitution and
//   dp in the output array.
// batched_band_solve! (src/infrastructure/band_diagonal.jl:87-189): the
//   reference densifies every column and eliminates with partial pivoting. A band
//   matrix keeps its lower bandwidth kl under partial pivoting and its upper
//   bandwidth grows to kl+ku, so the elimination here works on a (2kl+ku+1)-wide
//   window per row; every entry it skips is one the dense code updates as
//   A - f*0. Pivot choice, update order and back substitution are the
//   reference's, so results are the same.
#include "common.hpp"
#include "solvers.hpp"

namespace clmsoil {

constexpr int TRI_MAX = 64;

__global__ __launch_bounds__(256) void tridiag_multi_kernel(TridiagParams P) {
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= P.ncols) return;
    if (P.mask && !P.mask[col]) return;
    const int64_t ld = P.ld;
    const int64_t nlevs = P.nlevs;
    const int64_t j = P.jtop[col];
    if (j < 1 || j > nlevs) {
        atomicOr(P.err, (unsigned)DEV_ERR_NBEDROCK);
        return;
    }
    auto at = [&](const double *X, int64_t jj) { return X[col + (jj - 1) * ld]; };
    double cp[TRI_MAX];
    const double tiny = 1e-30;
    double bet = at(P.b, j);
    if (fabs(bet) < tiny) bet = tiny;
    double cp_prev = at(P.c, j) / bet;
    double dp_prev = at(P.r, j) / bet;
    cp[j - 1] = cp_prev;
    P.u[col + (j - 1) * ld] = dp_prev;
    for (int64_t jj = j + 1; jj <= nlevs; ++jj) {
        const double ajj = at(P.a, jj);
        double denom = at(P.b, jj) - ajj * cp_prev;
        if (fabs(denom) < tiny) denom = copysign(tiny, denom == 0 ? 1.0 : denom);
        cp_prev = at(P.c, jj) / denom;
        dp_prev = (at(P.r, jj) - ajj * dp_prev) / denom;
        cp[jj - 1] = cp_prev;
        P.u[col + (jj - 1) * ld] = dp_prev;
    }
    double u_next = dp_prev;
    for (int64_t jj = nlevs - 1; jj >= j; --jj) {
        u_next = P.u[col + (jj - 1) * ld] - cp[jj - 1] * u_next;
        P.u[col + (jj - 1) * ld] = u_next;
    }
}

constexpr int BAND_NMAX = 64;
constexpr int BAND_W = 7;   // 2*kl + ku + 1 for kl = ku = 2

__global__ __launch_bounds__(256) void band_solve_kernel(BandParams P) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.nc) return;
    if (P.mask && !P.mask[c]) return;
    const int kl = P.kl, ku = P.ku;
    const int64_t ld = P.ld;
    const int64_t nband = kl + ku + 1;
    const int64_t jt = P.jtop[c] + P.nlevsno_off + 1;
    const int64_t n = P.jbot[c] - P.jtop[c] + 1;
    if (n <= 0) return;
    if (jt < 1 || jt + n - 1 > P.nlev || n > BAND_NMAX) {
        atomicOr(P.err, (unsigned)DEV_ERR_SNL);
        return;
    }
    // W[i][q] holds A[i, i - kl + q], q = 0 .. 2kl+ku
    double W[BAND_NMAX][BAND_W];
    double x[BAND_NMAX];
    const int wid = 2 * kl + ku + 1;
    for (int i = 1; i <= n; ++i) {
        for (int q = 0; q < wid; ++q) W[i - 1][q] = 0.0;
        for (int d = -kl; d <= ku; ++d) {           // A[i, i+d] = bmatrix[c, kl+1-d, jt+i-1]
            const int jcol = i + d;
            if (jcol < 1 || jcol > n) continue;
            const int bb = kl + 1 - d;               // (i - j) + kl + 1 with j = i + d
            W[i - 1][d + kl] = P.bmatrix[c + (int64_t)(bb - 1) * ld + (jt + i - 2) * ld * nband];
        }
        x[i - 1] = P.rvector[c + (jt + i - 2) * ld];
    }
#define WA(i, jc) W[(i) - 1][(jc) - (i) + kl]
    bool dead = false;
    for (int k = 1; k <= n - 1; ++k) {
        int piv = k;
        double maxv = fabs(WA(k, k));
        const int ilast = (k + kl < n) ? k + kl : (int)n;
        for (int i = k + 1; i <= ilast; ++i) {
            const double v = fabs(WA(i, k));
            if (v > maxv) { maxv = v; piv = i; }
        }
        const double akk = WA(piv, k);
        if (akk == 0.0 || akk != akk) dead = true;   // the dense code turns every row below into NaN
        const int jlast = (k + kl + ku < n) ? k + kl + ku : (int)n;
        if (piv != k) {
            for (int jc = k; jc <= jlast; ++jc) {
                const double akj = WA(k, jc), apj = WA(piv, jc);
                WA(k, jc) = apj;
                WA(piv, jc) = akj;
            }
            const double xk = x[k - 1], xp = x[piv - 1];
            x[k - 1] = xp;
            x[piv - 1] = xk;
        }
        for (int i = k + 1; i <= ilast; ++i) {
            const double f = WA(i, k) / akk;
            for (int jc = k; jc <= jlast; ++jc) WA(i, jc) = WA(i, jc) - f * WA(k, jc);
            x[i - 1] = x[i - 1] - f * x[k - 1];
        }
    }
    for (int i = (int)n; i >= 1; --i) {
        double s = x[i - 1];
        const int jlast = (i + kl + ku < n) ? i + kl + ku : (int)n;
        for (int jc = i + 1; jc <= jlast; ++jc) s = s - WA(i, jc) * x[jc - 1];
        x[i - 1] = s / WA(i, i);
    }
#undef WA
    for (int i = 1; i <= n; ++i)
        P.tvector[c + (jt + i - 2) * ld] = dead ? __builtin_nan("") : x[i - 1];
}

hipError_t launch_tridiag_multi(const TridiagParams &P, hipStream_t stream) {
    if (P.ncols == 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(tridiag_multi_kernel, dim3((unsigned)((P.ncols + block - 1) / block)),
                       dim3(block), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_band_solve(const BandParams &P, hipStream_t stream) {
    if (P.nc == 0) return hipSuccess;
    const int block = 128;
    hipLaunchKernelGGL(band_solve_kernel, dim3((unsigned)((P.nc + block - 1) / block)),
                       dim3(block), 0, stream, P);
    return hipGetLastError();
}

}  // namespace clmsoil
