"""HIP first-substep predictor vs the float64 oracle (oracle/mlp_oracle.py).

Bars (FP32 network, FP32 accumulation): forward max|dy| <= 1e-4 * max|y|;
gradient max|dg| <= 2e-4 * max|g| and loss within 1e-4 relative; Adam within
1e-5 relative of the oracle step; features within FP32 rounding (1e-6 relative).
The learned controller only changes the first substep, so the water balance
of every column must close as tightly as with the plain controller."""
import numpy as np
import pytest

import mlp_oracle as M
from clmsoil import api, inputs, learned

pytestmark = pytest.mark.gpu


def _rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = learned.init_params(seed)
    p[learned.OB1:learned.OW2] = rng.normal(0, 0.1, 64)
    p[learned.OB2:learned.OW3] = rng.normal(0, 0.1, 64)
    p[learned.OB3] = 0.3
    X = rng.normal(0, 1, (24, n)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    return p, X, t


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("n", [1, 63, 200, 4096])
def test_forward_matches_oracle(gpu, n):
    p, X, t = _rand_problem(n, 5 + n)
    y = learned.mlp_forward(_dev(p), _dev(X)).cpu().numpy()
    yr, _ = M.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(y - yr)) <= 1e-4 * max(1.0, np.max(np.abs(yr)))


@pytest.mark.parametrize("n", [1, 64, 200, 4096])
def test_gradient_matches_oracle(gpu, n):
    p, X, t = _rand_problem(n, 11 + n)
    tr = learned.AdamTrainer(_dev(p), max_batch=4096)
    loss, g = tr.gradient(_dev(X), _dev(t))
    Lr, gr = M.loss_and_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    g = g.cpu().numpy()
    assert abs(float(loss) - Lr) <= 1e-4 * Lr
    assert np.max(np.abs(g - gr)) <= 2e-4 * np.max(np.abs(gr))
    # parameters are unchanged by the gradient query
    assert np.array_equal(tr.params.cpu().numpy(), p)


def test_adam_steps_match_oracle(gpu):
    p, X, t = _rand_problem(512, 21)
    tr = learned.AdamTrainer(_dev(p), lr=1e-2, max_batch=512)
    pr, m, v = p.astype(np.float64), np.zeros(M.NPARAM), np.zeros(M.NPARAM)
    for step in range(1, 4):
        tr.step(_dev(X), _dev(t))
        _, gr = M.loss_and_grad(pr, X.astype(np.float64), t.astype(np.float64))
        pr, m, v = M.adam(pr, m, v, gr, step, lr=1e-2)
    got = tr.params.cpu().numpy()
    assert np.max(np.abs(got - pr)) <= 1e-4 * np.max(np.abs(pr))


def test_training_reduces_loss(gpu):
    import torch
    rng = np.random.default_rng(8)
    X = rng.normal(0, 1, (24, 16384)).astype(np.float32)
    y = (np.tanh(X[0] - 0.5 * X[3]) + 0.3 * X[22]).astype(np.float32)
    tr = learned.AdamTrainer(_dev(learned.init_params(1)), lr=3e-3, max_batch=1024)
    Xd, yd = _dev(X), _dev(y)
    first = float(tr.gradient(Xd[:, :1024].contiguous(), yd[:1024].contiguous())[0])
    for ep in range(5):
        for i in range(0, 16384, 1024):
            tr.step(Xd[:, i:i + 1024].contiguous(), yd[i:i + 1024].contiguous())
    torch.cuda.synchronize()
    last = float(tr.gradient(Xd[:, :1024].contiguous(), yd[:1024].contiguous())[0])
    assert last < 0.3 * first


def _state(nc, seed):
    d = inputs.richards_config0(nc=nc, seed=seed, fill=0.0)
    d["nbedrock"] = np.random.default_rng(seed).integers(5, 21, nc).astype(np.int64)
    return d


def test_features_match_oracle(gpu):
    d = _state(777, 3)
    t = api.to_device(d)
    st, mi = learned.mf_views(t)
    x = learned.richards_features(st, mi, None, t["nbedrock"], 12, 20, 25, 1800.0).cpu().numpy()
    ref = M.features(d, 1800.0)
    np.testing.assert_allclose(x, ref, rtol=1e-6, atol=1e-7)
    norm = np.concatenate([ref.mean(1), ref.std(1) + 1.0, [-1.0, 0.5]]).astype(np.float32)
    xs = learned.richards_features(st, mi, None, t["nbedrock"], 12, 20, 25, 1800.0,
                                   norm=_dev(norm)).cpu().numpy()
    np.testing.assert_allclose(xs, M.features(d, 1800.0, norm), rtol=1e-5, atol=1e-5)


def test_predict_dtsub0_matches_oracle(gpu):
    d = _state(1000, 4)
    t = api.to_device(d)
    st, mi = learned.mf_views(t)
    p = learned.init_params(2)
    ref = M.features(d, 1800.0)
    norm = np.concatenate([ref.mean(1), ref.std(1) + 1.0, [-1.2, 0.7]]).astype(np.float32)
    pred = learned.DtPredictor(_dev(p), _dev(norm))
    dt0 = pred.predict_dtsub0(st, mi, None, t["nbedrock"], 12, 20, 25, 1800.0).cpu().numpy()
    y, _ = M.forward(p.astype(np.float64), M.features(d, 1800.0, norm))
    want = 1800.0 * np.exp(-1.2 + 0.7 * y)
    np.testing.assert_allclose(dt0, want, rtol=1e-4)


def test_trained_model_cuts_rejected_substeps(gpu):
    """The committed network (trained by `python -m clmsoil.learned` on the GPU)
    against the plain controller on held-out config-0 columns: fewer rejected
    substeps and fewer iterations, same water balance."""
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    pred = learned.DtPredictor.load()
    r = learned.compare_controllers(pred, 32768, 6, 99, cfg, two_launch=True)
    p, l, l2 = r["plain"], r["learned"], r["learned_two_launch"]
    assert l["rejected_rate"] < 0.5 * p["rejected_rate"]
    assert l["mean_iterations"] < p["mean_iterations"]
    assert p["max_balance_residual"] < 1e-9 and l["max_balance_residual"] < 1e-9
    # the two forms take the same first substeps, so the same substep paths
    assert l2["mean_iterations"] == l["mean_iterations"]
    assert l2["rejected_rate"] == l["rejected_rate"]


@pytest.mark.parametrize("persistent", [True, False])
def test_fused_inference_equals_predict_then_solve(gpu, persistent):
    """The network fused into the Richards kernel's column load gives the same
    bits as the prediction launch followed by the solve from its first
    substeps, on the resident-grid and the one-column-per-group launches, with
    masked columns and a partial last wavefront (1000 columns)."""
    import torch
    lib = clmsoil_lib()
    cfg = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                      lower_boundary_condition=api.BC_ZERO_FLUX)
    pred = learned.DtPredictor.load()
    nc = 1000
    d = _state(nc, 17)
    d["nrejected"] = np.zeros(nc)
    mask = np.random.default_rng(5).random(nc) > 0.1
    keys = ("h2osoi_liq", "smp_l", "hk_l", "qcharge", "nsubsteps", "nrejected")
    assert lib.clmsoil_set_richards_schedule(1 if persistent else 0) == 0
    try:
        ta, tb, tc = api.to_device(d), api.to_device(d), api.to_device(d)
        m = _dev(mask)
        st, mi = learned.mf_views(ta)
        dt0 = pred.predict_dtsub0(st, mi, m, ta["nbedrock"], 12, 20, 25, 1800.0)
        st, mi = learned.mf_views(ta, dt0)
        api.soilwm_moisture_form_solve(st, mi, m, ta["nbedrock"], 12, 13, 20, 1800.0, cfg, 25)
        st, mi = learned.mf_views(tb)
        pred.solve(st, mi, m, tb["nbedrock"], 12, 13, 20, 1800.0, cfg, 25)
        st, mi = learned.mf_views(tc)
        api.soilwm_moisture_form_solve(st, mi, m, tc["nbedrock"], 12, 13, 20, 1800.0, cfg, 25)
        torch.cuda.synchronize()
    finally:
        lib.clmsoil_set_richards_schedule(1)
    assert api.device_errors() == 0
    for k in keys:
        assert np.array_equal(ta[k].cpu().numpy(), tb[k].cpu().numpy()), k
    # masked columns are untouched, and the network changed the substep paths
    assert np.array_equal(tb["h2osoi_liq"].cpu().numpy()[:, ~mask], d["h2osoi_liq"][:, ~mask])
    nb, nc_plain = tb["nsubsteps"].cpu().numpy(), tc["nsubsteps"].cpu().numpy()
    assert np.sum(nb[mask] != nc_plain[mask]) > nc // 10
    assert nb[mask].sum() < nc_plain[mask].sum()


def test_fused_inference_host_validation(gpu):
    import ctypes as C
    from clmsoil import bind
    lib = clmsoil_lib()
    d = _state(8, 1)
    t = api.to_device(d)
    t.update(nc=8, nlevsno=12, nlevsoi=20, nlevgrnd=25, mask=None, smpmin=None,
             nrejected=None, dtsub0=None, workspace=None)
    a, lev = bind.mf_args(t), bind.levels(t)
    c = api.init_soilwater_movement(soilwater_movement_method=api.MOISTURE_FORM,
                                    lower_boundary_condition=api.BC_ZERO_FLUX).to_abi()
    pred = learned.DtPredictor.load()
    pp, pn = C.c_void_p(pred.params.data_ptr()), C.c_void_p(pred.norm.data_ptr())
    f = lib.clmsoil_soilwm_moisture_form_solve_learned
    assert f(C.byref(a), C.byref(lev), 1800.0, C.byref(c), None, pn, None) == 8   # not trained
    t["dtsub0"] = t["qcharge"]
    a2 = bind.mf_args(t)
    assert f(C.byref(a2), C.byref(lev), 1800.0, C.byref(c), pp, pn, None) == 7   # dtsub0 given


def clmsoil_lib():
    import clmsoil
    return clmsoil.load_library()
