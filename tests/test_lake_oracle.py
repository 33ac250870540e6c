"""Pins of the oracle's lake temperature solve (orc_lake_temperature_solve,
following lake_temperature.jl:632-785, 1205-1225 and tridiagonal.jl:86-142).

The reference's own lake tests check only that the full lake_temperature! runs
and conserves energy (test_lake_temperature.jl:285-336). What the mathematics
fixes, so that a plausible mistake fails one of them:
* energy: the Crank-Nicolson rows telescope, so the heat stored over the step,
  sum(cv * dT), equals dt * (fin + the solar sources of every level but the
  bottom) to rounding (a dropped term, a wrong sign or factor fails it);
* the solve: the new temperatures satisfy the assembled tridiagonal system
  (residual vs a dense numpy solve of the same rows);
* the interface conductivities: with one conductivity everywhere, every
  harmonic-mean interface conductivity equals it;
* the extended column's depths: snow nodes above 0, lake nodes at the lake
  grid, and the first soil node at the lake depth (sum of dz_lake) plus its own
  depth below the lake bottom;
* a column at one temperature with no heat input stays at it (to rounding).
"""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import inputs

DT = 1800.0


def ext_temps(d, c, jtop):
    """The extended column jtop..nlevlak+nlevgrnd of column c: (ext index, T)."""
    ns, nl, ng = d["nlevsno"], d["nlevlak"], d["nlevgrnd"]
    out = []
    for je in range(jtop, nl + ng + 1):
        if je < 1:
            out.append(d["t_soisno"][je + ns - 1, c])
        elif je <= nl:
            out.append(d["t_lake"][je - 1, c])
        else:
            out.append(d["t_soisno"][je - nl + ns - 1, c])
    return np.array(out)


def test_energy_conserved_and_rows_solved():
    d = inputs.lake_columns(nc=600, seed=92)
    out = clone(d)
    g = oracle.lake_temperature_solve(out, DT, dbg=True)
    ns = d["nlevsno"]
    for c in np.flatnonzero(d["mask_lakec"]):
        jtop = int(d["snl"][c]) + 1
        lo = jtop + ns - 1                               # 0-based extended row of the top
        t_old = g["tx"][lo:, c]
        t_new = ext_temps(out, c, jtop)
        cvx, phix = g["cvx"][lo:, c], g["phix"][lo:, c]
        stored = np.sum(cvx * (t_new - t_old))
        source = DT * (d["fin"][c] + phix[:-1].sum())
        scale = DT * (abs(d["fin"][c]) + np.abs(phix).sum()) + np.sum(cvx * np.abs(t_new)) * 1e-4
        assert abs(stored - source) <= 1e-11 * scale, (c, stored, source)
        n = len(t_new)
        A = np.diag(g["b"][lo:, c]) + np.diag(g["a"][lo + 1:, c], -1) + np.diag(g["c1"][lo:-1, c], 1)
        assert g["a"][lo, c] == 0.0 and g["c1"][-1, c] == 0.0
        x = np.linalg.solve(A, g["r"][lo:, c])
        np.testing.assert_allclose(t_new, x, rtol=1e-12, atol=0)
        assert n == d["nlevlak"] + d["nlevgrnd"] - jtop + 1
    off = d["mask_lakec"] == 0
    assert np.array_equal(out["t_lake"][:, off], d["t_lake"][:, off])


def test_uniform_conductivity_interfaces():
    d = inputs.lake_columns(nc=300, seed=93, uniform_tk=1.7)
    out = clone(d)
    g = oracle.lake_temperature_solve(out, DT, dbg=True)
    ns = d["nlevsno"]
    for c in np.flatnonzero(d["mask_lakec"]):
        lo = int(d["snl"][c]) + ns
        np.testing.assert_allclose(g["tkix"][lo:, c], 1.7, rtol=1e-14)


def test_extended_column_depths():
    d = inputs.lake_columns(nc=300, seed=94)
    g = oracle.lake_temperature_solve(clone(d), DT, dbg=True)
    ns, nl = d["nlevsno"], d["nlevlak"]
    for c in np.flatnonzero(d["mask_lakec"]):
        jtop = int(d["snl"][c]) + 1
        zx = g["zx"][jtop + ns - 1:, c]
        assert np.all(np.diff(zx) > 0)
        assert np.all(zx[:1 - jtop] < 0)                               # snow above the surface
        np.testing.assert_array_equal(zx[1 - jtop:1 - jtop + nl], d["z_lake"][:, c])
        first_soil = zx[1 - jtop + nl]
        assert abs(first_soil - (d["dz_lake"][:, c].sum() + d["z"][ns, c])) < 1e-12 * first_soil


def test_isothermal_column_without_input_stays():
    d = inputs.lake_columns(nc=200, seed=95)
    d["t_soisno"][:] = 280.0
    d["t_lake"][:] = 280.0
    d["fin"][:] = 0.0
    d["phi"][:] = 0.0
    d["phi_soil"][:] = 0.0
    d["sabg_lyr_col"][:] = 0.0
    out = clone(d)
    oracle.lake_temperature_solve(out, DT)
    np.testing.assert_allclose(out["t_lake"], 280.0, rtol=1e-13)
    np.testing.assert_allclose(out["t_soisno"], 280.0, rtol=1e-13)


def test_bad_snl_rejected():
    d = inputs.lake_columns(nc=8, seed=96)
    d["mask_lakec"][:] = 1
    d["snl"][2] = -d["nlevsno"] - 1
    with pytest.raises(RuntimeError):
        oracle.lake_temperature_solve(d, DT)
