"""HIP batched tridiagonal and banded operators vs the oracle and LAPACK."""
import numpy as np
import pytest

import oracle
from _util import rel_err
from clmsoil import api
from test_oracle_kat import _tri_case, band_solve, band_system


@pytest.mark.gpu
def test_tridiagonal_multi(gpu):
    import torch
    rng = np.random.default_rng(3)
    for ncols, nlevs, jt in ((4096, 21, 1), (1000, 5, 2), (513, 25, 4)):
        a, b, c, r = _tri_case(rng, ncols, nlevs)
        b[:, ::7] = 1e-31          # exercise the 1e-30 guard
        jtop = np.full(ncols, jt, dtype=np.int64)
        mask = (rng.random(ncols) < 0.9).astype(np.uint8)
        u = np.full((nlevs, ncols), np.nan)
        oracle.lib().orc_tridiagonal_multi(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                           r.ctypes.data, jtop.ctypes.data, mask.ctypes.data,
                                           ncols, nlevs, ncols)
        t = api.to_device(dict(a=a, b=b, c=c, r=r, jtop=jtop, mask=mask,
                               u=np.full((nlevs, ncols), np.nan)))
        api.tridiagonal_multi(t["u"], t["a"], t["b"], t["c"], t["r"], t["jtop"], t["mask"], ncols, nlevs)
        torch.cuda.synchronize()
        assert rel_err(t["u"].cpu().numpy(), u) <= 1e-12


def random_band_columns(rng, nc, n, kl, ku):
    """Band matrices in the reference's convention: A[i, j] != 0 only for
    j - kl <= i <= j + ku (ku sub-diagonals, kl super-diagonals)."""
    As, xs = [], []
    for col in range(nc):
        A = np.zeros((n, n))
        for d in range(-ku, kl + 1):           # np.diag offset d: A[i, i + d]
            A += np.diag(rng.normal(size=n - abs(d)), d)
        if col % 2:
            A += np.diag(np.full(n, 10.0))     # half the columns diagonally dominant
        As.append(A); xs.append(rng.normal(size=n))
    return As, xs


@pytest.mark.gpu
@pytest.mark.parametrize("kl,ku", [(1, 1), (2, 2), (1, 2), (2, 1)])
def test_batched_band_solve(gpu, kl, ku):
    import torch
    rng = np.random.default_rng(4 + kl + 3 * ku)
    nc, n = 777, 33
    As, xs = random_band_columns(rng, nc, n, kl, ku)
    bm, rv = band_system(As, xs, nlevs=40, jtop=-3, nlevsno_off=5, kl=kl, ku=ku)
    jtop = np.full(nc, -3, dtype=np.int64)
    jbot = np.full(nc, n - 4, dtype=np.int64)
    want = band_solve(bm, rv, jtop, jbot, kl, ku, nlevsno_off=5)
    t = api.to_device(dict(bm=bm, rv=rv, jtop=jtop, jbot=jbot, tv=np.full((40, nc), np.nan)))
    api.batched_band_solve(t["tv"], t["bm"], t["rv"], t["jtop"], t["jbot"], None, kl, ku, 5)
    torch.cuda.synchronize()
    got = t["tv"].cpu().numpy()
    assert rel_err(got, want, floor=1e-12) <= 1e-12
    for col in range(0, nc, 97):
        ref = np.linalg.solve(As[col], As[col] @ xs[col])
        assert rel_err(got[2:2 + n, col], ref, floor=1.0) < 1e-8
