"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np


def rel_err(a, b, floor=0.0):
    """max |a-b| / max(|a|, |b|, floor) over all elements (0 where both are 0,
    NaN positions must coincide). `floor` is a scalar or an array of a's shape."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return np.inf
    m = ~na
    a, b = a[m], b[m]
    if a.size == 0:
        return 0.0
    fl = np.broadcast_to(np.asarray(floor, dtype=np.float64), m.shape)[m] if np.ndim(floor) else floor
    den = np.maximum(np.maximum(np.abs(a), np.abs(b)), fl)
    d = np.abs(a - b)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(d == 0, 0.0, d / den)
    return float(r.max())


def bitwise_mismatch(a, b):
    """Number of elements whose float64 bits differ (NaN payloads aside: any
    NaN equals any NaN) between two arrays of the same shape."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return -1
    same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
    return int(np.count_nonzero(~same))


def assert_bitwise(got, ref, keys):
    """Every listed output equal bit for bit; the message names the fields that
    differ with their count of differing elements and largest relative error."""
    bad = {}
    for k in keys:
        if k not in ref:
            continue
        n = bitwise_mismatch(got[k], ref[k])
        if n:
            bad[k] = (n, rel_err(got[k], ref[k]))
    assert not bad, bad


def clone(d):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in d.items()}
