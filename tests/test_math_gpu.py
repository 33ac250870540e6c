"""The device build of the kernels' FP64 pow/log/log10/exp (clm_math.h) against
the host build the oracle links: bit for bit over random arguments in the
ranges the kernels use, wide ranges and the C99 special values. With this,
every transcendental the Richards and heat kernels evaluate rounds as the
oracle's does (the functions themselves are pinned in test_clm_math.py)."""
import numpy as np
import pytest

import oracle
from _util import bitwise_mismatch
from clmsoil import api


def _args(rng, n):
    sv = [0.0, -0.0, 1.0, -1.0, 0.5, -2.0, 3.0, 1e-310, np.inf, -np.inf, np.nan, 1e308,
          709.0, -745.5, 2.5]
    sx = np.array([a for a in sv for _ in sv])
    sy = np.array([b for _ in sv for b in sv])
    x = np.concatenate([rng.uniform(0.01, 1.0, n), np.exp(rng.uniform(-700, 700, n)),
                        np.full(n, 10.0), np.full(n, 0.57), rng.uniform(1e-7, 1.0, n), sx])
    y = np.concatenate([-rng.uniform(2.5, 15, n), rng.uniform(-2, 2, n), -6 * rng.uniform(0, 1, n),
                        rng.uniform(0, 0.95, n), rng.uniform(8, 33, n), sy])
    return x, y


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["pow", "log", "log10", "exp"])
def test_device_equals_host(gpu, op):
    import torch
    rng = np.random.default_rng(7)
    x, y = _args(rng, 200_000)
    if op == "exp":
        x = np.concatenate([x, rng.uniform(-746, 710, 200_000)])
        y = np.concatenate([y, np.zeros(200_000)])
    host = oracle.clm_math(op, x, y)
    dev = api.math_eval(op, torch.from_numpy(x).cuda(),
                        torch.from_numpy(y).cuda() if op == "pow" else None).cpu().numpy()
    assert bitwise_mismatch(dev, host) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("op,general", [("pow_w", "pow"), ("pow_base_w", "pow"),
                                        ("log10_w", "log10")])
def test_wavefront_fast_paths_equal_general(gpu, op, general):
    """The kernels' fast-path forms against the host's general functions: whole
    wavefronts in the fast path's domain, wavefronts with one lane outside it
    (the general function then runs for the wavefront) and special values."""
    import torch
    rng = np.random.default_rng(8)
    x, y = _args(rng, 64 * 3000)
    if op == "pow_base_w":   # fixed bases, one per wavefront of 64 lanes
        x = np.repeat(rng.choice([0.57, 2.29, 10.0, 0.3, 1.0, -2.0, np.nan], x.size // 64 + 1),
                      64)[:x.size]
    host = oracle.clm_math(general, x, y)
    dev = api.math_eval(op, torch.from_numpy(x).cuda(),
                        None if op == "log10_w" else torch.from_numpy(y).cuda()).cpu().numpy()
    assert bitwise_mismatch(dev, host) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["jl_max", "jl_min"])
def test_julia_max_min_device(gpu, op):
    """Julia's Float64 max/min as the kernels evaluate them: NaN from either
    side, signed zeros ordered, bit for bit as the host build."""
    import torch
    sv = np.array([np.nan, -np.nan, 0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 5e-324, 2.5])
    x = np.repeat(sv, sv.size)
    y = np.tile(sv, sv.size)
    host = oracle.clm_math(op, x, y)
    dev = api.math_eval(op, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()).cpu().numpy()
    same = (np.isnan(dev) & np.isnan(host)) | ((dev == host) & (np.signbit(dev) == np.signbit(host)))
    assert same.all()
