"""Trace-domain pieces of the path, restated in NumPy.  TEST INFRASTRUCTURE ONLY.

* ricker            fwi/forward.py:25-28
* cosine_taper_end  fwi/filters.py:26-32
* biquad / sosfilt  fwi/filters.py:14-23, 35-42 (direct form I, zero state,
                    coefficients normalised by a0)
* *_T               the exact transposes of the above, written as the
                    reverse-time recursion that reverse-mode AD of the DF-I
                    scan evaluates (fwi/gradient.py:208-210 takes jax.grad
                    through these)
* residual_prep     fwi/gradient.py:139-156: per-shot loss and receiver
                    cotangent
"""
from __future__ import annotations

import numpy as np


def ricker(t, fc, t0, scaling_factor=1.0):
    arg = (np.pi * fc * (np.asarray(t) - t0)) ** 2
    return scaling_factor * (1.0 - 2.0 * arg) * np.exp(-arg)


def taper_window(nt, n_taper):
    n = min(int(n_taper), int(nt))
    if n <= 0:
        return None
    return (np.cos(np.arange(1, n + 1) / n * np.pi) + 1) / 2


def cosine_taper_end(x, n_taper):
    x = np.array(x, copy=True)
    w = taper_window(x.shape[0], n_taper)
    if w is None:
        return x
    x[-w.size:] *= w.reshape((-1,) + (1,) * (x.ndim - 1))
    return x


# the taper is diagonal, so it is its own transpose
cosine_taper_end_T = cosine_taper_end


def _norm(s):
    s = np.asarray(s, np.float64)
    return s[0] / s[3], s[1] / s[3], s[2] / s[3], s[4] / s[3], s[5] / s[3]


def biquad(x, b0, b1, b2, a1, a2):
    x = np.asarray(x)
    y = np.zeros_like(x)
    x1 = np.zeros_like(x[0]); x2 = np.zeros_like(x[0]); y1 = np.zeros_like(x[0]); y2 = np.zeros_like(x[0])
    for i in range(x.shape[0]):
        xi = x[i]
        yi = b0 * xi + b1 * x1 + b2 * x2 - a1 * y1 - a2 * y2
        y[i] = yi
        x2, x1, y2, y1 = x1, xi, y1, yi
    return y


def biquad_T(g, b0, b1, b2, a1, a2):
    """Transpose of biquad(): carry cotangents (cx1, cx2, cy1, cy2) of the
    carry (x[i-1], x[i-2], y[i-1], y[i-2]) run backwards in time."""
    g = np.asarray(g)
    out = np.zeros_like(g)
    z = np.zeros_like(g[0])
    cx1, cx2, cy1, cy2 = z, z, z, z
    for i in range(g.shape[0] - 1, -1, -1):
        gy = g[i] + cy1
        out[i] = cx1 + b0 * gy
        cx1, cx2, cy1, cy2 = cx2 + b1 * gy, b2 * gy, cy2 - a1 * gy, -a2 * gy
    return out


def sosfilt(x, sos):
    for s in np.asarray(sos):
        x = biquad(x, *_norm(s))
    return x


def sosfilt_T(g, sos):
    for s in np.asarray(sos)[::-1]:
        g = biquad_T(g, *_norm(s))
    return g


def preprocess(x, sos=None, n_taper=0):
    """taper then filter, the order of fwi/gradient.py:147-152."""
    if n_taper > 0:
        x = cosine_taper_end(x, n_taper)
    if sos is not None:
        x = sosfilt(x, sos)
    return x


def preprocess_T(g, sos=None, n_taper=0):
    if sos is not None:
        g = sosfilt_T(g, sos)
    if n_taper > 0:
        g = cosine_taper_end_T(g, n_taper)
    return g


def residual_prep(pred, obs, sos=None, n_taper=0, scale=1.0, n_shots=1):
    """Per-shot misfit (scale/n_shots) * mean((F T pred - F T obs)^2) and its
    gradient with respect to pred (fwi/gradient.py:139-156,208-211)."""
    est = preprocess(pred, sos, n_taper)
    obs_p = preprocess(obs, sos, n_taper)
    r = est - obs_p
    w = scale / n_shots
    loss = w * np.mean(r ** 2)
    cot = preprocess_T(2.0 * r * (w / r.size), sos, n_taper)
    return float(loss), cot
