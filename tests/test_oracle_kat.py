"""The oracle against CLM.jl's own known-answer tests, exact solutions,
independent library solves and invariants (CPU only)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from _util import clone, rel_err
from clmsoil import bind, inputs

KAT = json.loads((Path(__file__).parent / "golden" / "reference_kats.json").read_text())


def _hk(hksat, bsw, s, imped):
    hk, d = C.c_double(), C.c_double()
    oracle.lib().orc_soil_hk(hksat, bsw, s, imped, C.byref(hk), C.byref(d))
    return hk.value, d.value


def _smp(bsw, sucsat, s):
    v, d = C.c_double(), C.c_double()
    oracle.lib().orc_soil_suction(bsw, sucsat, s, C.byref(v), C.byref(d))
    return v.value, d.value


def _ss(c, j):
    s = KAT["swrc_soilstate"]
    return s["hksat"][c - 1][j - 1], s["bsw"][c - 1][j - 1], s["sucsat"][c - 1][j - 1]


def test_swrc_values():
    v = KAT["swrc_values"]
    hks, b, suc = _ss(1, 1)
    hk, dhk = _hk(hks, b, 0.5, 1.0)
    assert hk == v["hk_c1j1_s05"] and dhk == v["dhkds_c1j1_s05"]
    smp, dsmp = _smp(b, suc, 0.5)
    assert smp == v["smp_c1j1_s05"] and dsmp == v["dsmpds_c1j1_s05"]
    hk, dhk = _hk(hks, b, 1.0, 1.0)
    assert hk == v["hk_c1j1_s1"] and dhk == pytest.approx(v["dhkds_c1j1_s1"], rel=1e-15)
    assert _smp(b, suc, 1.0)[0] == v["smp_c1j1_s1"]
    assert _hk(hks, b, 0.5, 0.1)[0] == pytest.approx(v["hk_c1j1_s05_imped01"], rel=1e-15)
    hks2, b2, suc2 = _ss(2, 1)
    assert _hk(hks2, b2, 0.5, 1.0)[0] == v["hk_c2j1_s05"]
    assert _smp(b2, suc2, 0.5)[0] == v["smp_c2j1_s05"]


def test_swrc_inverse_roundtrip_and_fd():
    v = KAT["swrc_values"]
    L = oracle.lib()
    for c, j in ((1, 1), (2, 1)):
        hks, b, suc = _ss(c, j)
        for s in v["roundtrip_s"]:
            smp = _smp(b, suc, s)[0]
            assert abs(L.orc_soil_suction_inverse(b, suc, smp) - s) <= v["roundtrip_atol"]
    eps = v["fd_eps"]
    for c, j, s in v["fd_cases"]:
        hks, b, suc = _ss(c, j)
        fd = (_hk(hks, b, s + eps, 1.0)[0] - _hk(hks, b, s - eps, 1.0)[0]) / (2 * eps)
        assert _hk(hks, b, s, 1.0)[1] == pytest.approx(fd, rel=v["fd_rtol"])
        fd = (_smp(b, suc, s + eps)[0] - _smp(b, suc, s - eps)[0]) / (2 * eps)
        assert _smp(b, suc, s)[1] == pytest.approx(fd, rel=v["fd_rtol"])


def test_ice_impedance():
    for f, e, want in KAT["ice_impedance"]["cases"]:
        assert oracle.lib().orc_ice_impedance(f, e) == pytest.approx(want, rel=1e-15)


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def test_tridiagonal_col():
    k = KAT["tridiagonal_col"]["case3"]
    a, b, c, r = map(_arr, (k["a"], k["b"], k["c"], k["r"]))
    u = np.zeros(3)
    oracle.lib().orc_tridiagonal_col(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                     r.ctypes.data, 1, 1, 3)
    np.testing.assert_allclose(u, k["u"], atol=k["atol"])
    k = KAT["tridiagonal_col"]["jtop2"]
    a, b, c, r = map(_arr, (k["a"], k["b"], k["c"], k["r"]))
    u = np.zeros(4)
    oracle.lib().orc_tridiagonal_col(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                     r.ctypes.data, 2, 1, 4)
    assert abs(2 * u[1] - u[2] - 1) < 1e-12
    assert abs(-u[1] + 2 * u[2] - u[3]) < 1e-12
    assert abs(-u[2] + 2 * u[3] - 1) < 1e-12


def test_rhs_lhs_moisture_form():
    k = KAT["rhs_moisture_form"]
    vts, qin, qout, dtdz = map(_arr, (k["vert_trans_sink"], k["qin"], k["qout"], k["dt_dz"]))
    rmx = np.zeros(3)
    oracle.lib().orc_compute_rhs_moisture_form(3, vts.ctypes.data, qin.ctypes.data, qout.ctypes.data,
                                               dtdz.ctypes.data, rmx.ctypes.data)
    np.testing.assert_allclose(rmx, -(qin - qout - vts) * dtdz, atol=1e-15)
    k = KAT["lhs_moisture_form"]
    dtdz, d0, d1, o1, o2 = map(_arr, (k["dt_dz"], k["dqidw0"], k["dqidw1"], k["dqodw1"], k["dqodw2"]))
    amx, bmx, cmx = np.zeros(4), np.zeros(4), np.zeros(4)
    oracle.lib().orc_compute_lhs_moisture_form(4, *(x.ctypes.data for x in (dtdz, d0, d1, o1, o2,
                                                                             amx, bmx, cmx)))
    assert amx[0] == 0.0 and cmx[3] == 0.0
    np.testing.assert_allclose(bmx, -1.0 - (-d1 + o1) * dtdz, rtol=1e-15)
    np.testing.assert_allclose(amx[1:], d0[1:] * dtdz[1:], rtol=1e-15)
    np.testing.assert_allclose(cmx[:3], -o2[:3] * dtdz[:3], rtol=1e-15)


def conservation_case():
    """The 12-column x 10-layer zero-flux case of the reference's conservation test."""
    k = KAT["moisture_form_conservation"]
    nc, nlev, nlevsno = k["nc"], k["nlev"], k["nlevsno"]
    dzv = k["dz"]
    nl = nlevsno + nlev
    z = np.zeros((nl, nc)); dz = np.zeros((nl, nc)); zi = np.zeros((nl + 1, nc))
    for j in range(1, nlev + 1):
        dz[nlevsno + j - 1] = dzv
        z[nlevsno + j - 1] = (j - 0.5) * dzv
        zi[nlevsno + j] = j * dzv
    liq = np.zeros((nl, nc))
    for j in range(1, nlev + 1):
        liq[nlevsno + j - 1] = (k["liq_frac_base"] + k["liq_frac_step"] * j) * k["watsat"] * dzv * 1000.0
    full = lambda v: np.full((nlev, nc), v)
    return dict(nc=nc, nlevsno=nlevsno, nlevsoi=nlev, nlevgrnd=nlev, dtime=k["dtime"],
                nbedrock=np.full(nc, nlev, dtype=np.int64), h2osoi_liq=liq,
                smp_l=full(0.0), hk_l=full(0.0), qcharge=np.zeros(nc), nsubsteps=np.zeros(nc),
                z=z, zi=zi, dz=dz, qflx_rootsoi=full(0.0), watsat=full(k["watsat"]),
                hksat=full(k["hksat"]), bsw=full(k["bsw"]), sucsat=full(k["sucsat"]),
                icefrac=full(0.0), zwt=np.full(nc, nlev * dzv), qflx_infl=np.full(nc, k["qflx_infl"]))


def test_moisture_form_conservation_kat():
    k = KAT["moisture_form_conservation"]
    d = conservation_case()
    liq0 = d["h2osoi_liq"].copy()
    nlev, ns = k["nlev"], k["nlevsno"]
    dbg = dict(qin=np.zeros((nlev, k["nc"])), qout=np.zeros((nlev, k["nc"])))
    oracle.soilwm_moisture_form_solve(d, bind.swm_config(lower_bc=k["lower_boundary_condition"]), dbg)
    dliq = d["h2osoi_liq"][ns:ns + nlev].sum(0) - liq0[ns:ns + nlev].sum(0)
    want = d["qflx_infl"] * k["dtime"]
    np.testing.assert_allclose(dliq, want, rtol=k["rtol"])
    assert np.abs(dliq - want).max() < k["max_abs_mass_err"]
    assert np.array_equal(dbg["qin"][1:], dbg["qout"][:-1])     # exact telescoping
    assert np.array_equal(dbg["qin"][0], d["qflx_infl"])


def test_tridiagonal_solve_kats():
    L = oracle.lib()
    k = KAT["tridiagonal_solve"]["simple3"]
    a, b, c, r = map(_arr, (k["a"], k["b"], k["c"], k["r"]))
    u = np.zeros(3)
    L.orc_tridiagonal_solve(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data, r.ctypes.data, 1, 3)
    np.testing.assert_allclose(u, k["u"], atol=k["atol"])
    r = _arr(KAT["tridiagonal_solve"]["identity5"]["r"])
    u = np.zeros(5); z5 = np.zeros(5); o5 = np.ones(5)
    L.orc_tridiagonal_solve(u.ctypes.data, z5.ctypes.data, o5.ctypes.data, z5.ctypes.data, r.ctypes.data, 1, 5)
    np.testing.assert_allclose(u, r, atol=1e-14)
    a = np.zeros(5); b = np.ones(5); c = np.zeros(5); r = np.zeros(5)
    b[2:] = 2.0; a[3:] = -1.0; c[2:4] = -1.0; r[2] = 1.0; r[4] = 1.0
    u = np.zeros(5)
    L.orc_tridiagonal_solve(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data, r.ctypes.data, 3, 5)
    np.testing.assert_allclose(u[2:], [1.0, 1.0, 1.0], atol=1e-12)
    k = KAT["tridiagonal_solve"]["residual"]
    n, co = k["nlevs"], k["coeff"]
    a = np.full(n, -co); b = np.full(n, 1 + 2 * co); c = np.full(n, -co); a[0] = 0; c[-1] = 0
    r = np.array([k["r0"] + k["rstep"] * j for j in range(1, n + 1)])
    u = np.zeros(n)
    L.orc_tridiagonal_solve(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data, r.ctypes.data, 1, n)
    res = b * u - r
    res[1:] += a[1:] * u[:-1]
    res[:-1] += c[:-1] * u[1:]
    assert np.abs(res).max() < k["max_residual"]


def _tri_case(rng, ncols, nlevs):
    a = -0.5 * np.ones((nlevs, ncols)); c = -0.5 * np.ones((nlevs, ncols))
    b = 2 + 0.5 * rng.random((nlevs, ncols)); r = rng.random((nlevs, ncols))
    a[0] = 0.0; c[-1] = 0.0
    return a, b, c, r


def test_tridiagonal_multi_vs_dense():
    rng = np.random.default_rng(2)
    for ncols, nlevs, jt in KAT["tridiagonal_multi_deep"]["cases"]:
        a, b, c, r = _tri_case(rng, ncols, nlevs)
        u = np.zeros((nlevs, ncols))
        jtop = np.full(ncols, jt, dtype=np.int64)
        st = oracle.lib().orc_tridiagonal_multi(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                                r.ctypes.data, jtop.ctypes.data, None, ncols, nlevs, ncols)
        assert st == 0
        for col in range(ncols):
            M = (np.diag(b[jt - 1:, col]) + np.diag(a[jt:, col], -1) + np.diag(c[jt - 1:-1, col], 1))
            np.testing.assert_allclose(u[jt - 1:, col], np.linalg.solve(M, r[jt - 1:, col]),
                                       atol=KAT["tridiagonal_multi_deep"]["atol"])


def test_tridiagonal_multi_mask():
    a = np.zeros((3, 3)); b = 2 * np.ones((3, 3)); c = np.zeros((3, 3)); r = np.ones((3, 3))
    u = np.zeros((3, 3)); jtop = np.ones(3, dtype=np.int64)
    mask = np.array([1, 0, 1], dtype=np.uint8)
    oracle.lib().orc_tridiagonal_multi(u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                       r.ctypes.data, jtop.ctypes.data, mask.ctypes.data, 3, 3, 3)
    assert np.all(u[:, 1] == 0.0) and np.all(u[:, 0] > 0) and np.all(u[:, 2] > 0)


def band_system(A_cols, x_cols, nlevs=None, jtop=0, nlevsno_off=0, kl=2, ku=2):
    """Pack dense per-column matrices into the reference's row-oriented band
    storage bmatrix[c, b, row] = A[row, row+kl+1-b] (CLM layout (nlev, nband, nc))."""
    n = A_cols[0].shape[0]
    nc = len(A_cols)
    nlevs = nlevs or n
    nband = kl + ku + 1
    bm = np.zeros((nlevs, nband, nc)); rv = np.zeros((nlevs, nc))
    jt = jtop + nlevsno_off + 1
    for col, (A, x) in enumerate(zip(A_cols, x_cols)):
        rhs = A @ x
        for i in range(n):          # 0-based i, 1-based row
            row = jt + i
            for bb in range(1, nband + 1):
                j = i + kl + 1 - bb
                if 0 <= j < n:
                    bm[row - 1, bb - 1, col] = A[i, j]
            rv[row - 1, col] = rhs[i]
    return bm, rv


def band_solve(bm, rv, jtop, jbot, kl=2, ku=2, nlevsno_off=0):
    nlev, nband, nc = bm.shape
    tv = np.full((nlev, nc), np.nan)
    st = oracle.lib().orc_batched_band_solve(tv.ctypes.data, bm.ctypes.data, rv.ctypes.data,
                                             jtop.ctypes.data, jbot.ctypes.data, None, kl, ku,
                                             nlevsno_off, nc, nlev, nc)
    assert st == 0
    return tv


def _banded(n, diag, off1, off2=0.0):
    A = np.diag(np.full(n, diag)) + np.diag(np.full(n - 1, off1), 1) + np.diag(np.full(n - 1, off1), -1)
    if off2:
        A += np.diag(np.full(n - 2, off2), 2) + np.diag(np.full(n - 2, off2), -2)
    return A


@pytest.mark.parametrize("case", ["tri4", "penta5", "penta6"])
def test_band_kats(case):
    k = KAT["band_diagonal"][case]
    x = np.array(k["x"]); n = len(x)
    A = _banded(n, k["diag"], k["off1"], k.get("off2", 0.0))
    kl = 1 if case == "tri4" else 2
    bm, rv = band_system([A, A], [x, x], kl=kl, ku=kl)
    # batched convention: rows jtop+nlevsno_off+1 .. jbot+nlevsno_off+1
    tv = band_solve(bm, rv, np.zeros(2, dtype=np.int64), np.full(2, n - 1, dtype=np.int64), kl, kl)
    np.testing.assert_allclose(tv[:, 0], x, atol=k["atol"])
    np.testing.assert_allclose(tv[:, 1], x, atol=k["atol"])


def test_band_partial_range():
    k = KAT["band_diagonal"]["partial"]
    x = np.array(k["x"]); n = len(x)
    A = _banded(n, k["diag"], k["off1"])
    bm, rv = band_system([A], [x], nlevs=k["nlevs"], jtop=k["jtop"] - 1, kl=1, ku=1)
    jtop = np.array([k["jtop"] - 1], dtype=np.int64)   # nlevsno_off = 0: jt = jtop + 1
    tv = band_solve(bm, rv, jtop, np.array([k["jbot"] - 1], dtype=np.int64), 1, 1)
    np.testing.assert_allclose(tv[k["jtop"] - 1:k["jbot"], 0], x, atol=k["atol"])
    assert np.all(np.isnan(tv[:k["jtop"] - 1, 0])) and np.all(np.isnan(tv[k["jbot"]:, 0]))


def test_band_pivoting_vs_lapack():
    """Random non-dominant pentadiagonal systems (rows must interchange): the
    reference's partial-pivoting elimination agrees with LAPACK gesv."""
    rng = np.random.default_rng(7)
    As, xs = [], []
    for _ in range(64):
        n = 30
        A = np.zeros((n, n))
        for d in range(-2, 3):
            A += np.diag(rng.normal(size=n - abs(d)), d)
        As.append(A); xs.append(rng.normal(size=n))
    bm, rv = band_system(As, xs)
    tv = band_solve(bm, rv, np.zeros(64, dtype=np.int64), np.full(64, 29, dtype=np.int64))
    for col in range(64):
        want = np.linalg.solve(As[col], As[col] @ xs[col])
        assert rel_err(tv[:, col], want, floor=1.0) < 1e-9


def test_heat_step_mass_identity_and_finite():
    """Phase change only moves water between phases: per active level,
    ice + liq (+ excess ice) is unchanged by the heat step without h2osfc."""
    h = inputs.heat_config1(nc=2048, seed=11)
    h.update(inputs.heat_outputs(h))
    m0 = h["h2osoi_ice"] + h["h2osoi_liq"]
    oracle.soil_temperature(h, bind.heat_config())
    m1 = h["h2osoi_ice"] + h["h2osoi_liq"]
    act = ~np.isnan(m0)
    k = KAT["phase_change_mass"]
    np.testing.assert_allclose(m1[act], m0[act], atol=k["atol"], rtol=k["rtol"])
    assert np.all(np.isfinite(h["t_soisno"][act]))
    assert np.all(h["h2osoi_liq"][act] >= 0) and np.all(h["h2osoi_ice"][act] >= 0)
    ns = 12
    fact = h["fact"]
    assert np.all(fact[ns:] > 0)                              # fact > 0 for soil layers
    assert np.all(h["thk"][ns:] > 0)


def test_heat_step_no_phase_change_above_freezing():
    """No ice, every layer above freezing: no phase change (test_soil_temperature.jl:534-564)."""
    h = inputs.heat_config1(nc=256, seed=3)
    ns = 12
    h["h2osoi_liq"][ns:] += h["h2osoi_ice"][ns:] * 1000.0 / 917.0
    h["h2osoi_ice"][ns:] = 0.0
    h["t_soisno"][ns:] = 285.0
    h["snl"][:] = 0
    h["frac_sno_eff"][:] = 0.0
    h.update(inputs.heat_outputs(h))
    oracle.soil_temperature(h, bind.heat_config())
    assert np.all(h["xmf"] == 0.0) and np.all(h["qflx_snomelt"] == 0.0)
    assert np.all(h["qflx_snofrz"] == 0.0)
    assert np.all(h["imelt"][ns:] == 0)


def test_richards_config0_balance():
    """Water balance per column: sum(dliq) = (infl - root - qcharge) * dtime for
    the zero-flux and water-table lower boundaries (the linearized bottom flux the
    solve uses is the one qcharge accumulates; under BC_FLUX the reference
    accumulates hk + dhkdw*dwat without the 1/watsat of the solve's derivative)."""
    for bc in (2, 3):
        d = inputs.richards_config0(nc=512, seed=5)
        liq0 = d["h2osoi_liq"].copy()
        oracle.soilwm_moisture_form_solve(d, bind.swm_config(lower_bc=bc))
        ns = 12
        dl = (d["h2osoi_liq"][ns:ns + 20] - liq0[ns:ns + 20]).sum(0)
        bal = (d["qflx_infl"] - d["qflx_rootsoi"].sum(0) - d["qcharge"]) * d["dtime"]
        assert np.abs(dl - bal).max() < 1e-9


@pytest.mark.parametrize("kl,ku", [(1, 2), (2, 1)])
def test_band_solve_unequal_bandwidths(kl, ku):
    """band_diagonal.jl:96-100 reads A[i, j] for j - kl <= i <= j + ku: ku
    sub-diagonals and kl super-diagonals. The oracle must solve such systems."""
    rng = np.random.default_rng(40 + kl)
    n = 11
    A = np.zeros((n, n))
    for d in range(-ku, kl + 1):
        A += np.diag(rng.normal(size=n - abs(d)), d)
    x = rng.normal(size=n)
    bm, rv = band_system([A], [x], kl=kl, ku=ku)
    tv = band_solve(bm, rv, np.zeros(1, dtype=np.int64), np.full(1, n - 1, dtype=np.int64), kl, ku)
    np.testing.assert_allclose(tv[:, 0], x, rtol=1e-9, atol=1e-9)
