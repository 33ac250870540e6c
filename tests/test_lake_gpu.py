"""HIP lake temperature solve vs the oracle (same seeded inputs). Bar:
bit-exact. The fused kernel forms the same rows with the same operations
(sums, products, quotients, the 1e-30 pivot guard) as the reference's five
assembly kernels and tridiagonal_multi!, with no transcendental function."""
import numpy as np
import pytest

import oracle
from _util import clone
from clmsoil import api, inputs
from clmsoil._abi import LAKE_FIELDS


def run_gpu(d, dtime):
    import torch
    t = api.to_device(d)
    api.lake_temperature_solve({k: t[k] for k in LAKE_FIELDS} | {"nc": d["nc"],
                                                               "nlevlak": d["nlevlak"]},
                               dtime, d["nlevsno"], d["nlevsoi"], d["nlevgrnd"])
    torch.cuda.synchronize()
    return api.to_host(t)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,max_snow,uniform", [(97, 6, None), (98, 12, None), (99, 0, 1.3)])
def test_lake_solve_vs_oracle(gpu, seed, max_snow, uniform):
    d = inputs.lake_columns(nc=3001, seed=seed, max_snow=max_snow, uniform_tk=uniform)
    api.device_errors()
    got = run_gpu(d, 1800.0)
    assert api.device_errors() == 0
    ref = clone(d)
    oracle.lake_temperature_solve(ref, 1800.0)
    assert np.array_equal(got["t_lake"], ref["t_lake"])
    assert np.array_equal(got["t_soisno"], ref["t_soisno"])


@pytest.mark.gpu
def test_lake_solve_bad_snl_flags(gpu):
    d = inputs.lake_columns(nc=300, seed=100)
    d["mask_lakec"][:] = 1
    d["snl"][7] = 1
    api.device_errors()
    got = run_gpu(d, 1800.0)
    assert api.device_errors() & 2                        # CLMSOIL_DEVERR_SNL
    assert np.array_equal(got["t_lake"][:, 7], d["t_lake"][:, 7])
    ref = clone(d)
    ref["mask_lakec"][7] = 0
    oracle.lake_temperature_solve(ref, 1800.0)
    assert np.array_equal(got["t_lake"], ref["t_lake"])
