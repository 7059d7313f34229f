"""Seeded synthetic surveys: the workloads the tests and the benchmark run.

Each function returns a dict with the model(s), the acquisition (sources,
per-shot receivers, record lengths, source types) and the time axis.  The
reference drives its surveys from scripts (Main_strain_fwi.py:60-118,
compute_gradient.py); these are the same kind of set-up on the seeded
layered and anomaly models of ``stfwi_amd.models``.

* ``anomaly_survey``: 64x128 grid, 20 m cells, fd4, 20-cell C-PML on every
  side, one moment source at depth 3 cells, x = 64, Ricker 8 Hz, 32
  receivers at depth 2, nt = 1000, dt = 2 ms; observed data from the
  Gaussian Vs anomaly model, the homogeneous background as initial model.
* ``layered_survey``: 152x500, 20 m, fd8, free surface, 20-cell C-PML, 25
  moment sources at depth 2 spaced 20 cells, 500 surface receivers, Ricker
  8 Hz, nt = 3000, dt = 1.5 ms; initial model the true one smoothed with
  sigma = 10 cells.
* ``batched_survey``: 64 shots on the 152x500 layered model with record
  lengths from {1500, 2000, 3000, 4000} steps, receiver spreads of 120-500
  traces centred on the source, moment and force sources 32/32 and finite
  difference orders 4 and 8 32/32, all drawn with one seed.
* ``large_survey``: 608x2000, 5 m cells, fd8, free surface, nt = 8000,
  dt = 0.5 ms, Ricker 15 Hz, 8 moment sources, 2000 receivers.
"""
from __future__ import annotations

import numpy as np

from .models import anomaly_model, layered_model, ricker, smooth_initial


def anomaly_survey(seed=0):
    nz, nx, nt, dt, fc = 64, 128, 1000, 2e-3, 8.0
    vs_t, vp_t, rho_t = anomaly_model(nz, nx, seed=seed)
    vs0, vp0, rho0 = (np.full((nz, nx), v) for v in (1730.0, 3000.0, 2300.0))
    rec_x = np.round(np.linspace(2, nx - 3, 32)).astype(np.int32)
    return dict(nz=nz, nx=nx, dx=20.0, dz=20.0, pad=20, free_surface=False, fd_order=4, nt=nt, dt=dt, fc=fc,
                true=(vs_t, vp_t, rho_t), init=(vs0, vp0, rho0), src_x=np.array([64]), src_z=np.array([3]),
                src_type=["moment"], rec_x=rec_x, rec_z=np.full(32, 2, np.int32), wavelet=ricker(nt, dt, fc))


def layered_survey(seed=0, n_shots=25):
    nz, nx, nt, dt, fc = 152, 500, 3000, 1.5e-3, 8.0
    true = layered_model(nz, nx, seed=seed)
    return dict(nz=nz, nx=nx, dx=20.0, dz=20.0, pad=20, free_surface=True, fd_order=8, nt=nt, dt=dt, fc=fc,
                true=true, init=smooth_initial(*true, 10.0), src_x=(10 + 20 * np.arange(n_shots)) % nx,
                src_z=np.full(n_shots, 2), src_type=["moment"] * n_shots, rec_x=np.arange(nx, dtype=np.int32),
                rec_z=np.zeros(nx, np.int32), wavelet=ricker(nt, dt, fc))


def batched_survey(seed=1, n_shots=64):
    """Uneven shots: per shot nt, receiver spread, source type and fd order."""
    nz, nx, dt, fc = 152, 500, 1.5e-3, 8.0
    rng = np.random.default_rng(seed)
    true = layered_model(nz, nx, seed=seed)
    nts = rng.choice([1500, 2000, 3000, 4000], n_shots)
    nrec = rng.integers(120, 501, n_shots)
    src_type = np.array(["moment", "force"] * (n_shots // 2))[rng.permutation(n_shots)]
    fd = np.array([4, 8] * (n_shots // 2))[rng.permutation(n_shots)]
    src_x = np.round(np.linspace(4, nx - 5, n_shots)).astype(np.int32)
    rec_x, rec_z = [], []
    for k in range(n_shots):
        start = int(np.clip(src_x[k] - nrec[k] // 2, 0, nx - nrec[k]))
        rec_x.append(np.arange(start, start + nrec[k], dtype=np.int32))
        rec_z.append(np.zeros(nrec[k], np.int32))
    return dict(nz=nz, nx=nx, dx=20.0, dz=20.0, pad=20, free_surface=True, nt=nts.astype(int), dt=dt, fc=fc,
                fd_order=fd, true=true, init=smooth_initial(*true, 10.0), src_x=src_x,
                src_z=np.full(n_shots, 2, np.int32), src_type=list(src_type), rec_x=rec_x, rec_z=rec_z,
                wavelet=ricker(int(nts.max()), dt, fc))


def large_survey(seed=0, n_shots=8):
    """vp is clipped at 4700 * 1.1 m/s so the fd8 Courant number at dx = 5 m,
    dt = 0.5 ms stays below 0.94."""
    nz, nx, nt, dt, fc = 608, 2000, 8000, 5e-4, 15.0
    true = layered_model(nz, nx, seed=seed, vp_clip=4700.0 * 1.1)
    return dict(nz=nz, nx=nx, dx=5.0, dz=5.0, pad=20, free_surface=True, fd_order=8, nt=nt, dt=dt, fc=fc,
                true=true, init=smooth_initial(*true, 40.0),
                src_x=np.round(np.linspace(100, nx - 101, n_shots)).astype(np.int32),
                src_z=np.full(n_shots, 2, np.int32), src_type=["moment"] * n_shots,
                rec_x=np.arange(nx, dtype=np.int32), rec_z=np.zeros(nx, np.int32), wavelet=ricker(nt, dt, fc))


def algorithmic_bytes(nz, nx, pad, free_surface, nt, nrec, s=4, k=1):
    """Minimal single-pass bytes of one shot-gradient (every array touched
    once per step): forward s (13 N + 10 N_pml + 3 N_p + 2 nrec) and adjoint
    s (13 N + 10 N_pml + 3 N_p + 2 k N_p + 2 nrec) per step, N padded cells,
    N_p physical cells, N_pml cells of the C-PML strips."""
    nz_t = nz + (pad if free_surface else 2 * pad)
    nx_t = nx + 2 * pad
    N, Np = nz_t * nx_t, nz * nx
    Npml = nz_t * 2 * pad + (pad if free_surface else 2 * pad) * nx_t
    fwd = s * (13 * N + 10 * Npml + 3 * Np + 2 * nrec)
    adj = s * (13 * N + 10 * Npml + 3 * Np + 2 * k * Np + 2 * nrec)
    return nt * (fwd + adj)
