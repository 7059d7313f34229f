"""Drop-in for fwi/forward.py: Ricker wavelet and the batched HIP forward.

Public API (same names, argument meaning and defaults as the reference):
    ricker_jax(t, fc, t0, scaling_factor=1)              fwi/forward.py:25
    forward_jax(Vs, Vp, rho, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx,
                dz, dt, nt, fc, pad, src_wavelet, ...)      fwi/forward.py:119
    build_forward_fn(nz, nx, dx, dz, dt, nt, freq, pad, block_size, rec_x,
                     rec_z, fd_order=2, return_vars=None,
                     free_surface=True)                     fwi/forward.py:508

Results are torch tensors on the HIP device, in the precision set by
stfwi_amd.config.enable_x64 (float32 by default, as JAX without x64).
``block_size`` is accepted and ignored: it only steers the reference's
reverse-mode checkpointing and never changes values.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from ._common import (as_dev, as_int, as_int_array, compute_dtype, forward_fd_order, get_plan, device)

_REC = {"vx": 0, "vz": 1, "ex": 2, "ez": 3}
_FULL = {"vx_full": "vx", "vz_full": "vz", "ex_full": "ex", "ez_full": "ez", "es_full": "es"}


def ricker_jax(t, fc, t0, scaling_factor=1):
    """s (1 - 2a) exp(-a), a = (pi fc (t - t0))^2 (fwi/forward.py:25-28)."""
    dt = compute_dtype()
    t = as_dev(t, dt)
    arg = (math.pi * fc * (t - t0)) ** 2
    return scaling_factor * (1.0 - 2.0 * arg) * torch.exp(-arg)


def _run(Vs, Vp, rho, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt, fc, pad, src_wavelet,
         return_wavefields, return_vars, fd_order, free_surface, src_type, dtype=None):
    """Forward of one or more shots (src_x / src_z scalars or 1-D)."""
    dtype = compute_dtype(dtype)
    fd = forward_fd_order(fd_order)
    sx = as_int_array(src_x)
    sz = as_int_array(src_z)
    if sz.size == 1 and sx.size > 1:
        sz = np.full_like(sx, sz[0])
    rx = as_int_array(rec_x)
    rz = as_int_array(rec_z)
    n = sx.size
    if return_vars is not None:
        wanted = list(return_vars)
    elif return_wavefields:
        wanted = ["vx", "vz", "ex", "ez", "vx_full", "vz_full", "ex_full", "ez_full"]
    else:
        wanted = ["vx", "vz", "ex", "ez"]
    for k in wanted:
        if k not in _REC and k not in _FULL:
            raise KeyError(k)
    snap = [_FULL[k] for k in wanted if k in _FULL]
    chans = sorted({_REC[k] for k in wanted if k in _REC})
    p = get_plan(nz_dom=int(nz_dom), nx_dom=int(nx_dom), dx=float(dx), dz=float(dz), dt=float(dt), fc=float(fc),
                 pad=int(pad), fd_order=fd, free_surface=bool(free_surface), dtype=dtype, n_shots_max=n,
                 nt_max=int(nt), nrec_max=max(int(rx.size), 1), imaging="none", grad="vs", store_history=False,
                 device=device())
    p.set_model(Vs, Vp, rho)
    w = as_dev(src_wavelet, dtype)
    p.set_shots(sx, sz, src_type, int(nt), rx, rz, w)
    out = p.forward(channels=tuple(chans) if chans else (), snap=tuple(dict.fromkeys(snap)))
    res = []
    for k in wanted:
        if k in _REC:
            res.append(out["rec"][:n, _REC[k], :nt, :rx.size])
        else:
            res.append(out["snap"][:n, out["snap_order"].index(_FULL[k]), :nt])
    return tuple(res)


def forward_jax(Vs, Vp, rho, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt, fc, pad, src_wavelet,
                return_wavefields=False, block_size=100, return_vars=None, fd_order=2, free_surface=True,
                src_type='force'):
    """2-D P-SV elastic velocity-strain forward (fwi/forward.py:119-501).

    Returns the requested variables: receiver records (nt, nrec) for
    'vx','vz','ex','ez' and full padded fields (nt, nz_t, nx_t) for
    '*_full'; default (rec_vx, rec_vz, rec_exx, rec_ezz)."""
    if src_type not in ("force", "moment"):
        raise NotImplementedError(f"src_type={src_type!r} not supported (use 'force' or 'moment').")
    res = _run(Vs, Vp, rho, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt, fc, pad, src_wavelet,
               return_wavefields, return_vars, fd_order, free_surface, src_type)
    return tuple(r[0] for r in res)


class _RunShot:
    """The compiled single-shot forward that build_forward_fn returns; also
    callable on arrays of source positions (the reference's jax.vmap over
    shots, e.g. Main_strain_fwi.py:114-117) via ``run_shot.batched``."""

    def __init__(self, nz, nx, dx, dz, dt, nt, freq, pad, rec_x, rec_z, fd_order, return_vars, free_surface,
                 src_type="force"):
        self.nz_dom, self.nx_dom = int(nz), int(nx)
        self.dx, self.dz, self.dt, self.nt, self.fc, self.pad = float(dx), float(dz), float(dt), int(nt), float(freq), int(pad)
        self.rec_x, self.rec_z = as_int_array(rec_x), as_int_array(rec_z)
        self.fd_order, self.return_vars, self.free_surface = fd_order, return_vars, bool(free_surface)
        self.src_type = src_type

    def _go(self, Vs, Vp, rho, src_x, src_z, src_wavelet):
        return _run(Vs, Vp, rho, src_x, src_z, self.rec_x, self.rec_z, self.nx_dom, self.nz_dom, self.dx, self.dz,
                    self.dt, self.nt, self.fc, self.pad, src_wavelet, False, self.return_vars, self.fd_order,
                    self.free_surface, self.src_type)

    def __call__(self, Vs, Vp, rho, src_x, src_z, src_wavelet):
        return tuple(r[0] for r in self._go(Vs, Vp, rho, as_int(src_x), as_int(src_z), src_wavelet))

    def batched(self, Vs, Vp, rho, src_x, src_z, src_wavelet):
        """All shots in one launch; every output gains a leading shot axis."""
        return self._go(Vs, Vp, rho, src_x, src_z, src_wavelet)


def build_forward_fn(nz, nx, dx, dz, dt, nt, freq, pad, block_size, rec_x, rec_z, fd_order=2, return_vars=None,
                     free_surface=True):
    """run_shot(Vs, Vp, rho, src_x, src_z, src_wavelet) with the grid, time
    axis, PML and receivers fixed (fwi/forward.py:508-544).  As in the
    reference it always models a force source."""
    return _RunShot(nz, nx, dx, dz, dt, nt, freq, pad, rec_x, rec_z, fd_order, return_vars, free_surface)
