// solvers.hip — batched per-column tridiagonal and banded solves (gfx950).
//
// tridiagonal_multi!  (src/infrastructure/tridiagonal.jl:86-142): one lane per
//   column, forward sweep with the 1e-30 pivot guard and register-carried
//   recurrences; cp is kept in lane-private storage for the back substitution and
//   dp in the output array.
// batched_band_solve! (src/infrastructure/band_diagonal.jl:87-189): the
//   reference densifies every column and eliminates with partial pivoting. A band
//   matrix with lower bandwidth p and upper bandwidth q keeps p under partial
//   pivoting while q grows to p+q, so the elimination here works on a
//   (2p+q+1)-wide window per row; every entry it skips is one the dense code
//   updates as A - f*0. Pivot choice, update order and back substitution are the
//   reference's, so results are the same. (The reference's own argument names
//   are the other way round: its kl counts super-diagonals and ku sub-diagonals.)
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
constexpr int BAND_W = 7;   // 2*lower + upper + 1 for bandwidths <= 2

// Partially pivoted elimination and back substitution of one banded system with
// lower bandwidth kl and upper bandwidth ku (the usual meaning here, not the
// reference's argument names) held in a window W[i][q] = A[i, i - kl + q]
// (q = 0 .. 2kl+ku), right-hand side x
// (overwritten by the solution). Returns true when a pivot is zero or NaN (the
// reference's dense elimination then fills the rows below with NaN).
__device__ bool band_eliminate(double (*W)[BAND_W], double *x, int n, int kl, int ku) {
#define WA(i, jc) W[(i) - 1][(jc) - (i) + kl]
    bool dead = false;
    for (int k = 1; k <= n - 1; ++k) {
        int piv = k;
        double maxv = fabs(WA(k, k));
        const int ilast = (k + kl < n) ? k + kl : n;
        for (int i = k + 1; i <= ilast; ++i) {
            const double v = fabs(WA(i, k));
            if (v > maxv) { maxv = v; piv = i; }
        }
        const double akk = WA(piv, k);
        if (akk == 0.0 || akk != akk) dead = true;   // the dense code turns every row below into NaN
        const int jlast = (k + kl + ku < n) ? k + kl + ku : n;
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
    for (int i = n; i >= 1; --i) {
        double s = x[i - 1];
        const int jlast = (i + kl + ku < n) ? i + kl + ku : n;
        for (int jc = i + 1; jc <= jlast; ++jc) s = s - WA(i, jc) * x[jc - 1];
        x[i - 1] = s / WA(i, i);
    }
#undef WA
    return dead;
}

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
    // The reference's A[i, j] = bmatrix[c, (i-j)+kl+1, jt+i-1] for j-kl <= i <= j+ku
    // (band_diagonal.jl:96-100): ku sub-diagonals and kl super-diagonals.
    // W[i][q] holds A[i, i - ku + q], q = 0 .. 2ku+kl
    const int lo = ku, up = kl;
    double W[BAND_NMAX][BAND_W];
    double x[BAND_NMAX];
    const int wid = 2 * lo + up + 1;
    for (int i = 1; i <= n; ++i) {
        for (int q = 0; q < wid; ++q) W[i - 1][q] = 0.0;
        for (int d = -lo; d <= up; ++d) {           // A[i, i+d] = bmatrix[c, kl+1-d, jt+i-1]
            const int jcol = i + d;
            if (jcol < 1 || jcol > n) continue;
            const int bb = kl + 1 - d;               // (i - j) + kl + 1 with j = i + d
            W[i - 1][d + lo] = P.bmatrix[c + (int64_t)(bb - 1) * ld + (jt + i - 2) * ld * nband];
        }
        x[i - 1] = P.rvector[c + (jt + i - 2) * ld];
    }
    const bool dead = band_eliminate(W, x, (int)n, lo, up);
    for (int i = 1; i <= n; ++i)
        P.tvector[c + (jt + i - 2) * ld] = dead ? __builtin_nan("") : x[i - 1];
}

// Reverse of tridiag_multi_kernel (one lane per column): lambda = A^-T ubar with
// the same 1e-30 pivot guard, then the band adjoints of enzyme_rules.jl:163-202.
__global__ __launch_bounds__(256) void tridiag_vjp_kernel(TridiagVjpParams P) {
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= P.ncols) return;
    if (P.mask && !P.mask[col]) return;
    const int64_t ld = P.ld, nlevs = P.nlevs;
    const int64_t j0 = P.jtop[col];
    if (j0 < 1 || j0 > nlevs) {
        atomicOr(P.err, (unsigned)DEV_ERR_NBEDROCK);
        return;
    }
    auto at = [&](const double *X, int64_t jj) { return X[col + (jj - 1) * ld]; };
    double gam[TRI_MAX], lam[TRI_MAX];
    const double tiny = 1e-30;
    // A^T: sub-diagonal c[j-1], diagonal b[j], super-diagonal a[j+1]
    double bet = at(P.b, j0);
    if (fabs(bet) < tiny) bet = tiny;
    lam[j0 - 1] = at(P.ubar, j0) / bet;
    for (int64_t jj = j0 + 1; jj <= nlevs; ++jj) {
        const double sub = at(P.c, jj - 1);
        gam[jj - 1] = at(P.a, jj) / bet;
        bet = at(P.b, jj) - sub * gam[jj - 1];
        if (fabs(bet) < tiny) bet = copysign(tiny, bet == 0 ? 1.0 : bet);
        lam[jj - 1] = (at(P.ubar, jj) - sub * lam[jj - 2]) / bet;
    }
    for (int64_t jj = nlevs - 1; jj >= j0; --jj) lam[jj - 1] = lam[jj - 1] - gam[jj] * lam[jj];
    for (int64_t jj = j0; jj <= nlevs; ++jj) {
        const int64_t o = col + (jj - 1) * ld;
        const double l = lam[jj - 1];
        if (P.rbar) P.rbar[o] += l;
        if (P.bbar) P.bbar[o] -= l * P.u[o];
        if (P.abar && jj > j0) P.abar[o] -= l * P.u[o - ld];
        if (P.cbar && jj < nlevs) P.cbar[o] -= l * P.u[o + ld];
        P.ubar[o] = 0.0;
    }
}

// Reverse of band_solve_kernel: the transposed system A^T (lower bandwidth ku,
// upper kl) is loaded into the same window and eliminated with partial pivoting.
__global__ __launch_bounds__(128) void band_vjp_kernel(BandVjpParams P) {
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
    auto bm = [&](int row, int band) -> int64_t {   // offset of bmatrix[c, band, jt+row-1]
        return c + (int64_t)(band - 1) * ld + (jt + row - 2) * ld * nband;
    };
    // A has ku sub-diagonals and kl super-diagonals, so A^T has kl and ku:
    // W[i][q] holds A^T[i, i - kl + q] = A[i - kl + q, i], q = 0 .. 2kl+ku
    double W[BAND_NMAX][BAND_W];
    double lam[BAND_NMAX];
    const int wid = 2 * kl + ku + 1;
    for (int i = 1; i <= n; ++i) {
        for (int q = 0; q < wid; ++q) W[i - 1][q] = 0.0;
        for (int d = -kl; d <= ku; ++d) {            // A^T[i, i+d] = A[i+d, i]
            const int r = i + d;
            if (r < 1 || r > n) continue;
            W[i - 1][d + kl] = P.bmatrix[bm(r, d + kl + 1)];   // band (r - i) + kl + 1
        }
        lam[i - 1] = P.tbar[c + (jt + i - 2) * ld];
    }
    const bool dead = band_eliminate(W, lam, (int)n, kl, ku);
    for (int i = 1; i <= n; ++i) {
        const double l = dead ? __builtin_nan("") : lam[i - 1];
        const int64_t o = c + (jt + i - 2) * ld;
        if (P.rvector_bar) P.rvector_bar[o] += l;
        if (P.bmatrix_bar) {
            for (int d = -ku; d <= kl; ++d) {        // A[i, i+d], band kl + 1 - d
                const int jc = i + d;
                if (jc < 1 || jc > n) continue;
                P.bmatrix_bar[bm(i, kl + 1 - d)] -= l * P.tvector[c + (jt + jc - 2) * ld];
            }
        }
    }
    for (int i = 1; i <= n; ++i) P.tbar[c + (jt + i - 2) * ld] = 0.0;
}

hipError_t launch_tridiag_vjp(const TridiagVjpParams &P, hipStream_t stream) {
    if (P.ncols == 0) return hipSuccess;
    hipLaunchKernelGGL(tridiag_vjp_kernel, dim3((unsigned)((P.ncols + 255) / 256)), dim3(256), 0,
                       stream, P);
    return hipGetLastError();
}

hipError_t launch_band_vjp(const BandVjpParams &P, hipStream_t stream) {
    if (P.nc == 0) return hipSuccess;
    hipLaunchKernelGGL(band_vjp_kernel, dim3((unsigned)((P.nc + 127) / 128)), dim3(128), 0, stream,
                       P);
    return hipGetLastError();
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
