"""Pins of the FP64 pow/log/log10/exp shared by the HIP kernels and the oracle
(clm.jl_amd/csrc/clm_math.h: Tang's table-driven log and exp, double-double
where pow needs it).

CLM.jl calls Julia's `^`, `log10`, `log` and `exp` (Base.Math, Julia >= 1.10,
not vendored in the reference tree); those are documented as faithfully
rounded and are correctly rounded in nearly all cases. The shared functions
are pinned against what the mathematics fixes:
  * values computed at 60 significant digits (Python's decimal module) and
    rounded once: every result within one ulp, >= 99.9 % exactly the correctly
    rounded double, over the argument ranges the soil kernels use and wide ones;
  * exactly representable results (x^1, 2^k, 4^0.5, 10^k, log10(10^k),
    exp(0), log(1)) returned exactly;
  * C99 Annex F special values, against the host C library's pow;
  * monotonicity in x, and agreement with the host C library (glibc) on
    >= 99.8 % of inputs.
A negative control shows the bar has teeth: pow computed as exp(y*log(x)) in
plain double fails it.
"""
from __future__ import annotations

import ctypes
import math
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle

getcontext().prec = 60

_libm = ctypes.CDLL("libm.so.6")
for _f, _n in (("pow", 2), ("log", 1), ("log10", 1), ("exp", 1)):
    getattr(_libm, _f).restype = ctypes.c_double
    getattr(_libm, _f).argtypes = [ctypes.c_double] * _n


def _ord(a):
    """Integers ordered like the doubles (for ulp distances)."""
    a = np.asarray(a, dtype=np.float64)
    i = a.view(np.int64)
    return np.where(i < 0, np.int64(-0x8000000000000000) - i, i)


def ulp_dist(a, b):
    return np.abs(_ord(a) - _ord(b))


def exact(op, x, y=None):
    out = np.empty(len(x))
    for k in range(len(x)):
        dx = Decimal(float(x[k]))
        if op == "pow":
            v = (Decimal(float(y[k])) * dx.ln()).exp()
        elif op == "log":
            v = dx.ln()
        elif op == "log10":
            v = dx.log10()
        else:
            v = dx.exp()
        out[k] = float(v)
    return out


def glibc(op, x, y=None):
    f = getattr(_libm, op)
    if op == "pow":
        return np.array([f(float(a), float(b)) for a, b in zip(x, y)])
    return np.array([f(float(a)) for a in x])


RNG = np.random.default_rng(20261204)
N = 6000
# (name, op, x, y): argument ranges of the kernels' calls and wide ones
CASES = [
    ("suction s^-b", "pow", RNG.uniform(0.01, 1.0, N), -RNG.uniform(2.5, 15.0, N)),
    ("conductivity s^(2b+3)", "pow", RNG.uniform(0.01, 1.0, N), RNG.uniform(8.0, 33.0, N)),
    ("ice impedance 10^(-6f)", "pow", np.full(N, 10.0), -6.0 * RNG.uniform(0.0, 1.0, N)),
    ("Johansen 0.57^(fl*ws)", "pow", np.full(N, 0.57), RNG.uniform(0.0, 0.95, N)),
    ("Johansen 2.29^((1-fl)*ws)", "pow", np.full(N, 2.29), RNG.uniform(0.0, 0.95, N)),
    ("supercooled (smp/psi)^(-1/b)", "pow", np.exp(RNG.uniform(-3, 12, N)), -1.0 / RNG.uniform(2.5, 15, N)),
    ("wide pow", "pow", np.exp(RNG.uniform(-300, 300, N)), RNG.uniform(-2.0, 2.0, N)),
    ("Kersten log10(satw)", "log10", RNG.uniform(1e-7, 1.0, N), None),
    ("wide log10", "log10", np.exp(RNG.uniform(-700, 700, N)), None),
    ("wide log", "log", np.exp(RNG.uniform(-700, 700, N)), None),
    ("log near 1", "log", 1.0 + RNG.uniform(-0.02, 0.02, N), None),
    ("wide exp", "exp", RNG.uniform(-744.0, 709.0, N), None),
    ("exp near 0", "exp", RNG.uniform(-0.01, 0.01, N), None),
]


@pytest.mark.parametrize("name,op,x,y", CASES, ids=[c[0] for c in CASES])
def test_against_60_digit_values(name, op, x, y):
    got = oracle.clm_math(op, x, y)
    ref = exact(op, x, y)
    d = ulp_dist(got, ref)
    assert d.max() <= 1, f"{name}: max {d.max()} ulp"
    assert np.mean(d == 0) >= 0.999, f"{name}: only {np.mean(d == 0):.5f} correctly rounded"
    # the host C library, for scale: at least as close on these samples
    g = ulp_dist(glibc(op, x, y), ref)
    assert np.count_nonzero(d) <= max(np.count_nonzero(g), 6)


def test_negative_control_plain_double_pow_fails():
    """pow as exp(y*log(x)) in plain double (no extra precision) misses the bar:
    the test above would catch a lost low part in log or in y*log(x)."""
    _, _, x, y = CASES[1]
    naive = np.array([math.exp(b * math.log(a)) for a, b in zip(x, y)])
    ref = exact("pow", x, y)
    d = ulp_dist(naive, ref)
    assert d.max() > 1 and np.mean(d == 0) < 0.999


def test_exact_results():
    xs = np.array([0.3, 1.7, 1e-300, 123.456, 0.01])
    assert np.array_equal(oracle.clm_math("pow", xs, np.ones_like(xs)), xs)
    k = np.arange(-1074, 1024, dtype=np.float64)
    assert np.array_equal(oracle.clm_math("pow", np.full(k.size, 2.0), k), np.ldexp(1.0, k.astype(int)))
    assert oracle.clm_math("pow", np.array([4.0]), np.array([0.5]))[0] == 2.0
    kk = np.arange(0, 23, dtype=np.float64)
    assert np.array_equal(oracle.clm_math("pow", np.full(kk.size, 10.0), kk), 10.0 ** kk)
    assert np.array_equal(oracle.clm_math("log10", 10.0 ** kk), kk)
    assert oracle.clm_math("exp", np.array([0.0]))[0] == 1.0
    assert oracle.clm_math("log", np.array([1.0]))[0] == 0.0
    assert oracle.clm_math("log", np.array([2.0]))[0] == math.log(2.0)


def _same(a, b):
    return (np.isnan(a) & np.isnan(b)) | ((a == b) & (np.signbit(a) == np.signbit(b)))


def test_special_values_c99():
    sv = [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, -2.0, 3.0, -3.0, 0.25, 1e-310, -1e-310,
          math.inf, -math.inf, math.nan, 1e308, 1e-308, 709.0, -745.5, 2.5, -2.5]
    x = np.array([a for a in sv for _ in sv])
    y = np.array([b for _ in sv for b in sv])
    got = oracle.clm_math("pow", x, y)
    ref = glibc("pow", x, y)
    bad = ~_same(got, ref)
    assert not bad.any(), list(zip(x[bad], y[bad], got[bad], ref[bad]))
    for op in ("log", "log10", "exp"):
        xs = np.array(sv)
        g, r = oracle.clm_math(op, xs), glibc(op, xs)
        bad = ~_same(g, r)
        assert not bad.any(), (op, xs[bad], g[bad], r[bad])


def test_subnormal_and_overflow_edges():
    x = np.array([5e-324, 1e-320, 2.2250738585072014e-308 / 3])
    ref = exact("log", x)
    assert ulp_dist(oracle.clm_math("log", x), ref).max() <= 1
    e = np.array([709.782712893384, 709.7827128933841, -745.1332191019411, -745.1332191019412,
                  -708.4, -720.0, -740.0])
    got = oracle.clm_math("exp", e)
    ref = exact("exp", e)
    # results in the subnormal range are rounded twice: one ulp of the subnormal grid
    assert ulp_dist(got, ref).max() <= 1, (got, ref)


def test_monotone_in_x():
    x = np.sort(RNG.uniform(0.01, 1.0, 200000))
    for b in (3.7, -5.2, 27.0):
        p = oracle.clm_math("pow", x, np.full(x.size, b))
        dp = np.diff(p)
        assert (dp >= 0).all() if b > 0 else (dp <= 0).all()


def test_bits_helper_roundtrip():
    # the ulp distance helper itself: adjacent doubles are one apart
    a = np.array([1.0, -1.0, 0.0, 5e-324])
    b = np.nextafter(a, np.inf)
    assert (ulp_dist(a, b) == 1).all()
    assert ulp_dist(np.array([0.0]), np.array([-0.0]))[0] == 0
    assert struct.pack("<d", 1.0) == np.float64(1.0).tobytes()


def test_fast_paths_equal_general_functions():
    """clm_pow_pos, clm_pow_base and clm_log10_pos return the bits of clm_pow /
    clm_log10 on their domains (positive normal x, finite y): the kernels take
    them for a batch whose arguments all qualify, the general functions else."""
    rng = np.random.default_rng(11)
    n = 400_000
    x = np.concatenate([rng.uniform(0.01, 1.0, n),                     # saturations
                        np.exp(rng.uniform(-700.0, 700.0, n)),         # wide range
                        [2.0 ** -1022, 1.0, 1.0 + 2.0 ** -52, np.finfo(float).max]])
    y = np.concatenate([rng.uniform(-20.0, 30.0, n), rng.uniform(-2.0, 2.0, n),
                        [0.5, 7.0, 0.0, -0.0]])
    y[::97] = 0.0
    y[1::101] = -0.0
    a = oracle.clm_math("pow", x, y)
    b = oracle.clm_math("pow_pos", x, y)
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
    for base in (0.57, 2.29, 10.0, 0.01, 1.5):
        yb = np.concatenate([rng.uniform(-300.0, 300.0, n // 4), [0.0, -0.0, 1.0, -1.0]])
        g = oracle.clm_math("pow", np.full_like(yb, base), yb)
        f = oracle.clm_math("pow_base", np.full_like(yb, base), yb)
        assert np.array_equal(g.view(np.int64), f.view(np.int64)), base
    g = oracle.clm_math("log10", x)
    f = oracle.clm_math("log10_pos", x)
    assert np.array_equal(g.view(np.int64), f.view(np.int64))
    # outside the domain they differ: the domain test is needed
    assert oracle.clm_math("pow_pos", np.array([-2.0]), np.array([2.0]))[0] != 4.0
