"""Pins of the oracle's vertical transport of the soil organic matter pools
(orc_litter_vert_transp, following litter_vert_transp.jl:93-639).

* the reference's known answers (test_litter_vert_transp.jl): patankar_A
  (:133-151), the cryoturbation diffusivity above the active layer and no
  advection (:467-503), frozen columns without mixing (:239-270), coarse woody
  debris taking only its source (:339-383), masked columns untouched
  (:385-425), the transport tendency (:427-465) and the bedrock correction;
* what the mathematics fixes: with zero-gradient ghost rows and no advection
  the transport conserves each pool's mass sum(conc * dz) plus its source to
  rounding (on columns with bedrock at the bottom level); a uniform profile without source stays uniform; Julia's x^5 is
  within one ulp of the exact fifth power (checked against rational arithmetic).
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import inputs

DT = 1800.0


def run(d, dtime=DT, **kw):
    p = dict(nlevdecomp=d["nlevdecomp"], nlevdecomp_full=d["nlevdecomp_full"],
             npools=d["npools"], ntypes=d["ntypes"], is_cwd=d["is_cwd"])
    p.update(kw)
    oracle.litter_vert_transp(d, dtime, **p)
    return d


def test_patankar_and_pow5():
    assert oracle.jl_pow5(1.0) == 1.0
    assert oracle.jl_pow5(0.5) == 0.5 ** 5                  # patankar_A(5) (exact)
    assert oracle.jl_pow5(0.0) == 0.0 and oracle.jl_pow5(-1.0) == -1.0
    rng = np.random.default_rng(3)
    for x in rng.uniform(-1.5, 1.5, 2000):
        exact = Fraction(float(x)) ** 5
        got = oracle.jl_pow5(float(x))
        ulp = np.spacing(abs(float(exact)))
        assert abs(Fraction(got) - exact) <= Fraction(ulp), x
    # patankar_A(pe) = max(0, jl_pow5(1 - 0.1|pe|)): 0 for |pe| >= 10
    assert oracle.jl_pow5(1.0 - 0.1 * 10.0) == 0.0


def test_uniform_profile_without_source_stays():
    d = inputs.lvt_columns(nc=500, seed=111)
    for t in range(d["ntypes"]):
        d["conc"][t][:] = 250.0
        d["source"][t][:] = 0.0
    d["nbedrock"][:] = d["nlevdecomp"]
    out = run(clone(d), dtime=3.0e7, som_diffus=1e-8, cryoturb_diffusion_k=5e-8)
    m = d["mask_bgc_soilc"] == 1
    for t in range(d["ntypes"]):
        np.testing.assert_allclose(out["conc"][t][:, :d["nlevdecomp"], m], 250.0, rtol=1e-12)


def clone_lvt(d):
    out = clone(d)
    for k in ("conc", "source", "tendency"):
        out[k] = [x.copy() for x in d[k]]
    return out


def test_mass_conserved_and_tendency():
    """Mass: on columns whose bedrock is the bottom decomposition level (the
    reference's own conservation test, test_litter_vert_transp.jl:272-337; below
    a shallower bedrock the rows couple to the level beneath with the
    conductance of the level above, and the bedrock move then collects what
    reached there)."""
    d = inputs.lvt_columns(nc=800, seed=112)
    nd, dz = d["nlevdecomp"], d["dzsoi_decomp"]
    d["nbedrock"][::2] = nd
    out = run(clone_lvt(d), dtime=3.0e7, som_diffus=1e-8, cryoturb_diffusion_k=5e-8)
    m = d["mask_bgc_soilc"] == 1
    full = m & (d["nbedrock"] == nd)
    moved = 0
    for t in range(d["ntypes"]):
        c0, c1, src = d["conc"][t][:, :nd], out["conc"][t][:, :nd], d["source"][t][:, :nd]
        before = np.einsum("sjc,j->sc", c0 + src, dz)
        after = np.einsum("sjc,j->sc", c1, dz)
        np.testing.assert_allclose(after[:, full], before[:, full], rtol=1e-12)
        for s in range(d["npools"]):
            if d["is_cwd"][s]:
                np.testing.assert_array_equal(c1[s][:, m & (d["nbedrock"] == nd)],
                                              (c0[s] + src[s])[:, m & (d["nbedrock"] == nd)])
                assert np.all(out["tendency"][t][s] == 7.0)            # untouched
                continue
            moved += np.sum(np.abs(c1[s][:, m] - (c0[s] + src[s])[:, m]) > 1e-6 * np.abs(c0[s][:, m]))
        # below bedrock everything has moved into the bedrock layer
        j = np.arange(d["nlevdecomp_full"])[:, None]
        below = (j[:nd] >= d["nbedrock"][None, :]) & m[None, :]
        assert np.all(c1[:, below] == 0.0)
    assert moved > 1000
    off = ~m
    assert np.array_equal(out["conc"][0][:, :, off], d["conc"][0][:, :, off])


def test_tendency_is_the_concentration_change():
    """test_litter_vert_transp.jl:427-465: the tendency equals (new - old - source)
    / dtime, where new is the solved concentration before the bedrock move."""
    d = inputs.lvt_columns(nc=300, seed=113)
    d["nbedrock"][:] = d["nlevdecomp"]
    out = run(clone_lvt(d), dtime=3.0e7, som_diffus=1e-8)
    nd = d["nlevdecomp"]
    m = d["mask_bgc_soilc"] == 1
    for s in range(d["npools"]):
        if d["is_cwd"][s]:
            continue
        exp = (out["conc"][0][s, :nd] - d["conc"][0][s, :nd] - d["source"][0][s, :nd]) / 3.0e7
        np.testing.assert_allclose(out["tendency"][0][s, :nd][:, m], exp[:, m], rtol=1e-9,
                                   atol=1e-15 * np.abs(d["conc"][0][s, :nd][:, m]).max() / 3.0e7)


def test_coefficients_cryoturbation_and_frozen():
    d = inputs.lvt_columns(nc=400, seed=114)
    d["mask_bgc_soilc"][:] = 1
    k = 5.0e-4
    out = run(clone_lvt(d), cryoturb_diffusion_k=k, som_diffus=1.0e-4, max_altdepth_cryoturbation=2.0,
              som_adv_flux=0.0)
    nd = d["nlevdecomp"]
    zi = d["zisoi"]
    alt = np.maximum(d["altmax"], d["altmax_lastyear"])
    for c in range(d["nc"]):
        nb = d["nbedrock"][c]
        dif, adv = out["som_diffus_coef"][:, c], out["som_adv_coef"][:, c]
        if alt[c] == 0.0:                                  # frozen: no mixing at all
            assert np.all(dif == 0.0) and np.all(adv == 0.0)
        elif alt[c] <= 2.0:                                # cryoturbation, no advection
            assert np.all(adv == 0.0)
            for j in range(1, nd + 2):
                if j <= nb + 1 and zi[j] < alt[c]:
                    assert dif[j - 1] == k
                if j > nb + 1:
                    assert dif[j - 1] == 0.0
            assert np.all(np.diff(dif[:nb + 1]) <= 0.0)    # decreasing with depth
        else:                                              # bioturbation
            assert np.all(dif[:nb + 1] == 1.0e-4) and np.all(dif[nb + 1:] == 0.0)


def test_bedrock_correction_and_bad_input():
    d = inputs.lvt_columns(nc=50, seed=115, npools=2, cwd_pools=())
    d["mask_bgc_soilc"][:] = 1
    d["nbedrock"][:] = 3
    d["altmax"][:] = 0.0                                   # no mixing: the leak alone
    d["altmax_lastyear"][:] = 0.0
    out = run(clone_lvt(d))
    c0 = d["conc"][0] + d["source"][0]
    dz = d["dzsoi_decomp"]
    nd = d["nlevdecomp"]
    exp3 = c0[:, 2] + sum(c0[:, j] * (dz[j] / dz[2]) for j in range(3, nd))
    np.testing.assert_allclose(out["conc"][0][:, 2], exp3, rtol=1e-12)
    assert np.all(out["conc"][0][:, 3:nd] == 0.0)
    d["nbedrock"][4] = 0
    with pytest.raises(RuntimeError):
        run(clone_lvt(d))
    with pytest.raises(RuntimeError):
        run(clone_lvt(d), use_soil_matrixcn=True)
