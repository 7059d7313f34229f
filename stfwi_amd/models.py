"""Seeded synthetic earth models for the benchmark and the tests.

* layered_model: random layered earth with 12-20 interfaces (sinusoidal
  dips of amplitude <= 15 cells) and 3 faults (offset <= 10 cells);
  vp 1600-4700 m/s increasing with depth with +-10% noise per layer,
  vs = vp / U(1.6, 2.0) per layer, rho = 310 vp^0.25.
* smooth_initial: the drivers' initial model, a Gaussian smoothing of
  slowness for velocities and of rho itself (Main_strain_fwi.py:70-73).
* anomaly_model: constant background plus a Gaussian Vs anomaly.
* load_marmousi, marmousi_acquisition, stable_dt: the reference drivers'
  set-up of the Marmousi run (model files, shot and receiver lines, time
  step), so that a user of those drivers gets the same survey.

The reference ships only Marmousi; the generators are this project's own.
"""
from __future__ import annotations

import numpy as np


def layered_model(nz=152, nx=500, seed=0, n_faults=3, vp_range=(1600.0, 4700.0), vp_clip=None):
    rng = np.random.default_rng(seed)
    n_if = int(rng.integers(12, 21))
    depth = np.sort(rng.uniform(4, nz - 4, n_if))
    xs = np.arange(nx)
    tops = np.empty((n_if, nx))
    for i in range(n_if):
        amp = rng.uniform(0, 15)
        wl = rng.uniform(0.5, 3.0) * nx
        ph = rng.uniform(0, 2 * np.pi)
        tilt = rng.uniform(-5, 5)
        tops[i] = depth[i] + amp * np.sin(2 * np.pi * xs / wl + ph) + tilt * (xs / nx - 0.5)
    for _ in range(n_faults):
        xf = int(rng.integers(nx // 10, nx - nx // 10))
        off = rng.uniform(-10, 10)
        tops[:, xf:] += off
    tops = np.sort(tops, axis=0)
    nl = n_if + 1
    base = np.linspace(vp_range[0], vp_range[1], nl)
    vp_l = base * (1 + rng.uniform(-0.1, 0.1, nl))
    ratio = rng.uniform(1.6, 2.0, nl)
    zz = np.arange(nz)[:, None].astype(np.float64)
    layer = (zz >= tops[:, None, :]).sum(axis=0)          # (nz, nx) layer index
    vp = vp_l[layer]
    vs = vp / ratio[layer]
    if vp_clip is not None:
        vp = np.minimum(vp, vp_clip)
        vs = np.minimum(vs, vp / 1.6)
    rho = 310.0 * vp ** 0.25
    return vs, vp, rho


def smooth_initial(vs, vp, rho, sigma=10.0):
    from scipy.ndimage import gaussian_filter
    return (1.0 / gaussian_filter(1.0 / vs, sigma), 1.0 / gaussian_filter(1.0 / vp, sigma),
            gaussian_filter(rho, sigma))


def anomaly_model(nz=64, nx=128, seed=0, vp0=3000.0, vs0=1730.0, rho0=2300.0, amp=0.10, sigma=6.0):
    rng = np.random.default_rng(seed)
    cz, cx = nz // 2 + int(rng.integers(-5, 6)), nx // 2 + int(rng.integers(-5, 6))
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    g = np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * sigma ** 2))
    vs = vs0 * (1 + amp * g)
    return vs, np.full((nz, nx), vp0), np.full((nz, nx), rho0)


def ricker(nt, dt, fc, t0=None):
    """fwi/forward.py:25-28 with the callers' t = arange(nt) dt, t0 = 1.5/fc."""
    t = np.arange(nt) * dt
    t0 = 1.5 / fc if t0 is None else t0
    arg = (np.pi * fc * (t - t0)) ** 2
    return (1.0 - 2.0 * arg) * np.exp(-arg)


MARMOUSI_FILES = ("vs_true", "vp_true", "rho_true", "vs_init", "vp_init", "rho_init")


def load_marmousi(model_dir):
    """The Marmousi arrays of the reference's marmousi_models/ directory
    (Main_strain_fwi.py:64-67): a dict of the .npy files present among
    vs/vp/rho _true/_init, loaded without pickles.  The three true models
    are required and must share one 2-D shape."""
    import os
    out = {}
    for name in MARMOUSI_FILES:
        path = os.path.join(model_dir, name + ".npy")
        if os.path.exists(path):
            out[name] = np.load(path, allow_pickle=False)
    missing = [n for n in MARMOUSI_FILES[:3] if n not in out]
    if missing:
        raise FileNotFoundError(f"{model_dir}: missing {', '.join(n + '.npy' for n in missing)}")
    shapes = {a.shape for a in out.values()}
    if len(shapes) != 1 or len(next(iter(shapes))) != 2:
        raise ValueError(f"{model_dir}: the model arrays must share one 2-D shape, got {sorted(shapes)}")
    return out


def marmousi_acquisition(nx_dom, n_shots, src_step=15, src_first=2, rec_step=2):
    """Surface shot and receiver lines of the reference drivers
    (Main_strain_fwi.py:78-87, compute_first_gradient.py:86-93): sources
    every src_step cells from column src_first, n_shots of them picked at
    evenly spaced indices (linspace, truncated); receivers every rec_step
    cells from column 0; all at depth 0.  Returns (src_x, src_z, rec_x,
    rec_z) with int64 sources and int32 receivers."""
    full = np.arange(0, nx_dom - src_first, src_step) + src_first
    idx = np.linspace(0, len(full) - 1, n_shots).astype(int)
    src_x = full[idx]
    src_z = np.zeros(len(full), dtype=int)[idx]
    rec_x = np.arange(0, nx_dom, rec_step, dtype=np.int32)
    return src_x, src_z, rec_x, np.zeros(len(rec_x), dtype=np.int32)


def stable_dt(vp_max, dx, dz, fd_order=8, courant=0.9):
    """Time step of the reference drivers (Main_strain_fwi.py:89-98):
    courant / (S vp_max sqrt(1/dx^2 + 1/dz^2)), S the sum of the magnitudes
    of the staggered-grid coefficients (fd8: 1225/1024 + 245/3072 + 49/5120 +
    5/7168; fd4: 9/8 + 1/24; otherwise 1).  vp_max keeps its type, as in the
    drivers: the maximum of a float32 model makes S vp_max a float32 product,
    and the time step then matches theirs to the last bit.  The number of
    steps for a record of tmax seconds is int(tmax / dt)."""
    if fd_order == 8:
        S = 1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168
    elif fd_order >= 4:
        S = 9 / 8 + 1 / 24
    else:
        S = 1.0
    return courant / (S * vp_max * np.sqrt(1 / dx ** 2 + 1 / dz ** 2))
