"""Hand-written adjoints vs the mathematics that fixes them.

* tridiagonal VJP: the reference's own adjoint known-answer case
  (test/test_enzyme_smoke.jl:187-246: n = 9, jtop = 3, all four bands) against
  a dense transposed solve (numpy) to 1e-10 and central differences to 1e-6;
  the entries the sweep never reads (a[1..jtop], c[n]) must come back exactly 0.
* band VJP: lambda = A^-T tbar, rbar = lambda, Abar = -lambda x^T on the band,
  against numpy dense solves, with and without row swaps, for kl != ku too.
* Richards step and multi-day calibration adjoints: central differences with
  relative step 1e-6, bar 1e-6 in the reference's norm-wise relative error
  max|adj - fd| / max|fd| per parameter (test_enzyme_smoke.jl:238), skipping any
  column whose accept/reject path changes under the perturbation. (Element-wise,
  a few entries whose gradient is ~1e-3 of the largest sit at ~1e-5: the FD
  noise of a loss summing O(100) kg/m2 water amounts whose per-step change is
  ~1e-6 of that.)
"""
import numpy as np
import pytest

from clmsoil import adjoint, api, calibration, inputs
from test_oracle_kat import band_system

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


# ---------------------------------------------------------------- tridiagonal
def _tri_dense(a, b, c, j0, n):
    A = np.zeros((n - j0 + 1, n - j0 + 1))
    for j in range(j0, n + 1):
        i = j - j0
        A[i, i] = b[j - 1]
        if j > j0:
            A[i, i - 1] = a[j - 1]
        if j < n:
            A[i, i + 1] = c[j - 1]
    return A


def test_tridiagonal_vjp_reference_case(gpu):
    import torch
    n, jtop = 9, 3
    a0 = np.concatenate([np.zeros(jtop), [0.31, -0.22, 0.44, -0.17, 0.28, -0.39]])
    b0 = np.array([3.4, 3.1, 3.9, 3.2, 3.7, 3.3, 3.5, 3.8, 3.6])
    c0 = np.concatenate([[-0.25, 0.33, -0.41, 0.19, -0.29, 0.37, -0.21, 0.24], [0.0]])
    r0 = np.array([0.7, -1.2, 0.4, 1.9, -0.6, 0.3, -1.1, 0.8, 0.2])
    w = np.array([0.9, -0.4, 1.3, 0.2, -1.7, 0.6, 0.5, -0.8, 1.1])
    # columns: the case, the case with jtop = 1, and a masked-out copy
    cols = 3
    A = lambda v: np.repeat(v[:, None], cols, axis=1)
    jt = np.array([jtop, 1, jtop], dtype=np.int64)
    mask = np.array([1, 1, 0], dtype=np.uint8)
    t = {k: _dev(v) for k, v in dict(a=A(a0), b=A(b0), c=A(c0), r=A(r0), jt=jt, mask=mask).items()}
    u = torch.zeros((n, cols), dtype=torch.float64, device="cuda")
    api.tridiagonal_multi(u, t["a"], t["b"], t["c"], t["r"], t["jt"], t["mask"], cols, n)
    ubar = _dev(A(w))
    bars = {k: torch.zeros((n, cols), dtype=torch.float64, device="cuda") for k in "abcr"}
    adjoint.tridiagonal_multi_vjp(ubar, u, t["a"], t["b"], t["c"], t["jt"], t["mask"], cols, n,
                                  bars["a"], bars["b"], bars["c"], bars["r"])
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    got = {k: v.cpu().numpy() for k, v in bars.items()}
    for col, j0 in ((0, jtop), (1, 1)):
        M = _tri_dense(a0, b0, c0, j0, n)
        x = np.linalg.solve(M, r0[j0 - 1:])
        lam = np.linalg.solve(M.T, w[j0 - 1:])
        ref = {"r": lam, "b": -lam * x,
               "a": np.concatenate([[0.0], -lam[1:] * x[:-1]]),
               "c": np.concatenate([-lam[:-1] * x[1:], [0.0]])}
        for k in "abcr":
            g = got[k][j0 - 1:, col]
            assert np.max(np.abs(g - ref[k])) <= 1e-10 * max(np.max(np.abs(ref[k])), 1e-12), k
        # central differences of w . u(a, b, c, r), step 1e-6 * max(|v|, 1)
        def obj(aa, bb, cc, rr):
            return float(w[j0 - 1:] @ np.linalg.solve(_tri_dense(aa, bb, cc, j0, n), rr[j0 - 1:]))
        base = dict(a=a0, b=b0, c=c0, r=r0)
        for k in "abcr":
            fd = np.zeros(n)
            for i in range(n):
                h = 1e-6 * max(abs(base[k][i]), 1.0)
                vp, vm = dict(base), dict(base)
                vp[k] = base[k].copy(); vp[k][i] += h
                vm[k] = base[k].copy(); vm[k][i] -= h
                fd[i] = (obj(vp["a"], vp["b"], vp["c"], vp["r"]) -
                         obj(vm["a"], vm["b"], vm["c"], vm["r"])) / (2 * h)
            assert np.max(np.abs(got[k][:, col] - fd)) <= 1e-6 * max(np.max(np.abs(fd)), 1e-12), k
        # entries the forward sweep never reads carry exactly zero
        assert np.all(got["a"][:j0, col] == 0.0)
        assert got["c"][n - 1, col] == 0.0
    # ubar consumed on the system rows; masked column untouched
    ub = ubar.cpu().numpy()
    assert np.all(ub[jtop - 1:, 0] == 0.0) and np.all(ub[:, 1] == 0.0)
    assert np.array_equal(ub[:, 2], w)
    for k in "abcr":
        assert np.all(got[k][:, 2] == 0.0)


# ---------------------------------------------------------------- band
@pytest.mark.parametrize("kl,ku", [(2, 2), (1, 2), (2, 1)])
def test_band_vjp_matches_dense(gpu, kl, ku):
    import torch
    rng = np.random.default_rng(7 + kl + 3 * ku)
    nc, n, nlev, jtop, off = 40, 17, 24, -2, 4
    As, xs = [], []
    for col in range(nc):
        A = np.zeros((n, n))
        for d in range(-ku, kl + 1):
            A += np.diag(rng.normal(size=n - abs(d)), d)
        if col % 2:
            A += np.diag(np.full(n, 8.0))
        As.append(A); xs.append(rng.normal(size=n))
    bm, rv = band_system(As, xs, nlevs=nlev, jtop=jtop, nlevsno_off=off, kl=kl, ku=ku)
    jt = jtop + off + 1
    tbar = np.zeros((nlev, nc))
    tbar[jt - 1:jt - 1 + n] = rng.normal(size=(n, nc))
    jtv = np.full(nc, jtop, dtype=np.int64)
    jbv = np.full(nc, jtop + n - 1, dtype=np.int64)
    t = api.to_device(dict(bm=bm, rv=rv, jt=jtv, jb=jbv, tv=np.zeros((nlev, nc)), tbar=tbar.copy(),
                           bmb=np.zeros_like(bm), rvb=np.zeros_like(rv)))
    api.batched_band_solve(t["tv"], t["bm"], t["rv"], t["jt"], t["jb"], None, kl, ku, off)
    adjoint.batched_band_solve_vjp(t["tbar"], t["tv"], t["bm"], t["jt"], t["jb"], None, kl, ku, off,
                                   t["bmb"], t["rvb"])
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    bmb, rvb = t["bmb"].cpu().numpy(), t["rvb"].cpu().numpy()
    tv = t["tv"].cpu().numpy()
    assert np.all(t["tbar"].cpu().numpy() == 0.0)
    for col in range(nc):
        A = As[col]
        x = tv[jt - 1:jt - 1 + n, col]
        lam = np.linalg.solve(A.T, tbar[jt - 1:jt - 1 + n, col])
        scale = max(np.max(np.abs(lam)), 1e-12)
        assert np.max(np.abs(rvb[jt - 1:jt - 1 + n, col] - lam)) <= 1e-10 * scale
        want = np.zeros((nlev, kl + ku + 1))
        for i in range(n):
            for bb in range(1, kl + ku + 2):
                j = i + kl + 1 - bb
                if 0 <= j < n:
                    want[jt - 1 + i, bb - 1] = -lam[i] * x[j]
        assert np.max(np.abs(bmb[:, :, col] - want)) <= 1e-10 * scale * max(np.max(np.abs(x)), 1.0)


# ---------------------------------------------------------------- Richards step
def _mf(t, liq=None, dtsub0=None):
    st = api.MfState(t["h2osoi_liq"] if liq is None else liq, t["smp_l"], t["hk_l"], t["qcharge"],
                     t["nsubsteps"], t["nrejected"])
    mi = api.MfIn(t["z"], t["zi"], t["dz"], t["qflx_rootsoi"], t["watsat"], t["hksat"], t["bsw"],
                  t["sucsat"], t["icefrac"], t["zwt"], t["qflx_infl"], None, dtsub0)
    return st, mi


@pytest.mark.parametrize("lower_bc", [api.BC_FLUX, api.BC_ZERO_FLUX])
def test_richards_step_adjoint_vs_central_differences(gpu, lower_bc):
    import torch
    nc = 64
    d = inputs.richards_config0(nc=nc, seed=31, lower_bc=lower_bc, nbedrock="random", fill=0.0)
    d["nrejected"] = np.zeros(nc)
    base = api.to_device(d)
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=lower_bc)
    rng = np.random.default_rng(2)
    wl = _dev(rng.normal(size=d["h2osoi_liq"].shape))
    wq = _dev(rng.normal(size=nc) * 1e4)

    def run(t):
        t = {k: (v.clone() if hasattr(v, "clone") else v) for k, v in t.items()}
        st, mi = _mf(t)
        api.soilwm_moisture_form_solve(st, mi, None, t["nbedrock"], 12, 13, 20, 1800.0, cfg, 25)
        Lc = (wl * t["h2osoi_liq"]).sum(0) + wq * t["qcharge"]
        return Lc, t["nsubsteps"].clone(), t["nrejected"].clone()

    L0, ns0, nr0 = run(base)
    st, mi = _mf(base)
    lbar = wl.clone()
    g = torch.zeros((3, nc), dtype=torch.float64, device="cuda")
    ws = adjoint.new_workspace(nc)
    adjoint.soilwm_moisture_form_adjoint(st, mi, None, base["nbedrock"], 12, 20, 25, 1800.0, cfg,
                                         lbar, wq, ws, grad_scale=g)
    torch.cuda.synchronize()
    assert api.device_errors() == 0
    h = 1e-6
    checks = [("bsw", 0), ("hksat", 1), ("watsat", 2)]
    worst, skipped = 0.0, 0
    for name, q in checks + [("liq", j) for j in (0, 3, 9)]:
        tp = dict(base); tm = dict(base)
        if name == "liq":
            row = 12 + q
            tp["h2osoi_liq"] = base["h2osoi_liq"].clone(); tp["h2osoi_liq"][row] *= 1 + h
            tm["h2osoi_liq"] = base["h2osoi_liq"].clone(); tm["h2osoi_liq"][row] *= 1 - h
            adj = lbar[row] * base["h2osoi_liq"][row]          # dL/dlog(liq)
        else:
            tp[name] = base[name] * (1 + h)
            tm[name] = base[name] * (1 - h)
            adj = g[q]
        Lp, nsp, nrp = run(tp)
        Lm, nsm, nrm = run(tm)
        fd = (Lp - Lm) / (2 * h)
        same_path = ((nsp == ns0) & (nsm == ns0) & (nrp == nr0) & (nrm == nr0)).cpu().numpy()
        skipped += int((~same_path).sum())
        ok = same_path.copy()
        if name == "liq":
            ok &= (q < base["nbedrock"]).cpu().numpy()
        fd, adj = fd.cpu().numpy(), adj.cpu().numpy()
        # norm-wise relative error, the reference's metric (test_enzyme_smoke.jl:238)
        err = np.max(np.abs(adj - fd)[ok]) / max(np.max(np.abs(fd[ok])), 1e-12)
        worst = max(worst, float(err))
    assert worst <= 1e-6, worst
    assert skipped <= 6 * 3


def test_calibration_gradient_vs_central_differences(gpu):
    r = calibration.fd_check(nc=64, days=2)
    assert r["max_rel_err"] is not None and r["max_rel_err"] <= 1e-6, r["max_rel_err"]
    assert r["excluded_columns"] <= 3
    assert np.all(np.isfinite(r["grad"]))


def test_calibration_gradient_30_days(gpu):
    """The calibration workload's full horizon: 30 days (1,440 steps) on 64
    columns, adjoint through one checkpoint per day against central differences."""
    r = calibration.fd_check(nc=64, days=30)
    assert r["max_rel_err"] is not None and r["max_rel_err"] <= 1e-6, r["max_rel_err"]
    assert r["excluded_columns"] <= 4
    assert np.all(np.isfinite(r["grad"]))


def test_calibration_gradient_with_freezing_columns(gpu):
    """Columns around the freezing point: layers freeze and thaw, water moves
    between h2osoi_liq and h2osoi_ice in the heat step, and the gradient runs
    through the heat-step adjoint as well (2 days, 96 steps). Layers that sit at
    their supercooled-water limit switch between frozen and unfrozen labels
    under a 1e-6 perturbation of the multipliers; such columns are excluded,
    and at least 16 columns with phase change must remain."""
    import torch
    cal = calibration.Calibration(64, 2, seed=5, freezing=True)
    ice0 = cal.base["h2osoi_ice"].clone()
    D, ck, _ = cal.forward(torch.ones((3, 64), dtype=torch.float64, device="cuda"),
                           keep_checkpoints=True)
    froze = (ck[1]["h2osoi_ice"] != ice0).any(dim=0).cpu().numpy()
    assert froze.mean() > 0.3
    r = calibration.fd_check(nc=64, days=2, freezing=True)
    assert r["max_rel_err"] is not None and r["max_rel_err"] <= 1e-6, r["max_rel_err"]
    assert (r["valid"].all(axis=0) & froze).sum() >= 16
    assert np.all(np.isfinite(r["grad"]))
