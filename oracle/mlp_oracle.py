"""float64 numpy restatement of the first-substep predictor (features, network,
loss gradient, Adam) defined in include/clmsoil_learned.h.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the clmsoil package. The
network has no counterpart in CLM.jl; the features restate the reference's node
saturation (src/biogeophys/soil_water_movement.jl:1273-1286).
"""
from __future__ import annotations

import numpy as np

NIN, NH, NPARAM = 24, 64, 5825
K = np.sqrt(2.0 / np.pi)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[0:1536].reshape(NH, NIN), p[1536:1600], p[1600:5696].reshape(NH, NH),
            p[5696:5760], p[5760:5824], p[5824])


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K * (x + 0.044715 * x ** 3)))


def gelu_d(x):
    th = np.tanh(K * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * K * (1.0 + 3.0 * 0.044715 * x * x)


def forward(p, X):
    """X [24, n] -> y [n] and the intermediates."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Z1 = W1 @ X + b1[:, None]
    H1 = gelu(Z1)
    Z2 = W2 @ H1 + b2[:, None]
    H2 = gelu(Z2)
    y = w3 @ H2 + b3
    return y, (Z1, H1, Z2, H2)


def loss_and_grad(p, X, t, n_total=None):
    """L = sum((y - t)^2) / n_total and dL/dp packed like p."""
    X = np.asarray(X, np.float64)
    t = np.asarray(t, np.float64)
    n_total = X.shape[1] if n_total is None else n_total
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (Z1, H1, Z2, H2) = forward(p, X)
    e = y - t
    L = float(np.sum(e * e) / n_total)
    dy = 2.0 * e / n_total
    D2 = (w3[:, None] * dy[None, :]) * gelu_d(Z2)
    D1 = (W2.T @ D2) * gelu_d(Z1)
    g = np.concatenate([(D1 @ X.T).ravel(), D1.sum(1), (D2 @ H1.T).ravel(), D2.sum(1),
                        H2 @ dy, [dy.sum()]])
    return L, g


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(d, dtime, norm=None):
    """Raw features [24, nc] of a moisture-form state dict (host arrays in the
    (nlev, nc) layout), standardised when norm[50] is given."""
    nc = int(d["nc"])
    nlevsno, nlevsoi = int(d["nlevsno"]), int(d["nlevsoi"])
    nb = np.clip(np.asarray(d["nbedrock"]), 0, nlevsoi)
    f = np.zeros((NIN, nc))
    for j in range(1, 21):
        if j > nlevsoi:
            break
        liq = d["h2osoi_liq"][nlevsno + j - 1]
        dz = d["dz"][nlevsno + j - 1]
        vwc = np.maximum(liq, 1e-6) / (dz * 1000.0)
        s = np.clip(vwc / d["watsat"][j - 1], 0.01, 1.0)
        f[j - 1] = np.where(j <= nb, s, 0.0)
    f[20] = np.log(d["hksat"][0])
    f[21] = d["bsw"][0]
    f[22] = d["qflx_infl"]
    f[23] = dtime
    if norm is not None:
        norm = np.asarray(norm, np.float64)
        f = (f - norm[:NIN, None]) / norm[NIN:2 * NIN, None]
    return f
