"""Forward modelling of a survey as a differentiable torch operation.

The reference differentiates its JAX forward with ``jax.value_and_grad``
(fwi/loss.py:134) and bounds memory with ``jax.checkpoint`` blocks
(fwi/forward.py:471-501).  Here ``SeismogramOperator`` runs all shots in one
batched HIP forward that keeps what the exact discrete adjoint needs: the
strain history, or checkpoints of the forward state every K steps with
segment recomputation.  Its backward is that adjoint with the exact imaging
condition fused into every reverse step, so any torch loss of the
seismograms gets exact gradients with respect to (Vs, Vp, rho) on the
physical grid.  Outside the physical domain the model is not differentiated,
as the reference's stop_gradient does (fwi/forward.py:217-219), and the C-PML
profile's Vp_max is a constant (fwi/forward.py:234-236).

    op = SeismogramOperator(nz, nx, dx, dz, dt, fc, pad, src_x, src_z, wavelets,
                            rec_x, rec_z, nt, fd_order=8, channels=(2,))
    rec = op(vs, vp, rho)        # (n_shots, len(channels), nt, nrec)
    loss = ((rec - obs) ** 2).mean(); loss.backward()
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .plan import Plan

_GRAD = {1: "vs", 2: "vs_vp", 3: "all"}


class _Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vs, vp, rho, op):
        rec, gp, gen = op._run_forward(vs, vp, rho)
        ctx.op, ctx.gp, ctx.gen = op, gp, gen
        ctx.save_for_backward(vs, vp, rho)
        return rec

    @staticmethod
    def backward(ctx, grad_rec):
        vs, vp, rho = ctx.saved_tensors
        g = ctx.op._run_adjoint(ctx.gp, ctx.gen, grad_rec, vs, vp, rho)
        return g[0], g[1], g[2], None


class SeismogramOperator:
    """Batched forward of one survey; call it on (Vs, Vp, rho) tensors of
    shape (nz, nx) on the HIP device.

    src_type: 'force' or 'moment' (one for all shots or one per shot);
    wavelets: (nt,) shared or (n_shots, nt); channels: receiver channels to
    return (0 vx, 1 vz, 2 exx, 3 ezz); ckpt_every: 0 keeps the strain
    history of every step, K > 0 keeps forward checkpoints every K steps
    (None picks K = ceil(sqrt(nt))).  Plans are built per gradient set
    (Vs; Vs, Vp; Vs, Vp, rho), chosen from the inputs that require grad.
    """

    def __init__(self, nz, nx, dx, dz, dt, fc, pad, src_x, src_z, wavelets, rec_x, rec_z, nt, src_type="force",
                 fd_order=8, free_surface=True, dtype=torch.float32, channels: Sequence[int] = (2,), ckpt_every=0,
                 groups=1, device=None):
        self.geom = dict(nz_dom=int(nz), nx_dom=int(nx), dx=float(dx), dz=float(dz), dt=float(dt), fc=float(fc),
                         pad=int(pad), fd_order=int(fd_order), free_surface=bool(free_surface), dtype=dtype)
        self.src_x = np.asarray(src_x, np.int32).reshape(-1)
        self.src_z = np.asarray(src_z, np.int32).reshape(-1)
        self.n_shots = self.src_x.size
        self.rec_x = np.asarray(rec_x, np.int32)
        self.rec_z = np.asarray(rec_z, np.int32)
        self.nt = int(nt)
        self.src_type = src_type
        self.channels = tuple(int(c) for c in channels)
        if not self.channels or any(c not in (0, 1, 2, 3) for c in self.channels):
            raise ValueError(f"channels={channels!r}: use a subset of (0, 1, 2, 3)")
        self.ckpt_every = int(np.ceil(np.sqrt(self.nt))) if ckpt_every is None else int(ckpt_every)
        self.groups = int(groups)
        self.dtype = dtype
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        w = torch.as_tensor(wavelets, dtype=dtype, device=self.device)
        self.wavelets = w.expand(self.n_shots, self.nt).contiguous() if w.dim() == 1 else w.contiguous()
        self._plans = {}
        self._gen = {}

    def _plan(self, gp):
        if gp not in self._plans:
            nrec = self.rec_x.size if self.rec_x.ndim == 1 else self.rec_x.shape[-1]
            p = Plan(**self.geom, n_shots_max=self.n_shots, nt_max=self.nt, nrec_max=nrec, imaging="exact",
                     grad=_GRAD[gp], store_history=True, device=self.device, ckpt_every=self.ckpt_every)
            p.set_groups(self.groups)
            self._plans[gp] = p
            self._gen[gp] = 0
        return self._plans[gp]

    @staticmethod
    def _grad_set(vs, vp, rho):
        return 3 if rho.requires_grad else (2 if vp.requires_grad else 1)

    def __call__(self, vs, vp, rho):
        return _Seismograms.apply(vs, vp, rho, self)

    def _run_forward(self, vs, vp, rho):
        gp = self._grad_set(vs, vp, rho)
        p = self._plan(gp)
        p.set_model(vs, vp, rho)
        p.set_shots(self.src_x, self.src_z, self.src_type, self.nt, self.rec_x, self.rec_z, self.wavelets)
        rec = p.forward(channels=self.channels)["rec"]
        self._gen[gp] += 1
        return rec[:, list(self.channels), :self.nt].clone(), gp, self._gen[gp]

    def _run_adjoint(self, gp, gen, grad_rec, vs, vp, rho):
        if self._gen[gp] != gen:
            raise RuntimeError("this SeismogramOperator ran another forward since the one being differentiated; "
                               "call backward before the next forward")
        p = self._plans[gp]
        radj = torch.zeros((self.n_shots, 4, p.nt_max, p.nrec_max), dtype=self.dtype, device=self.device)
        radj[:, list(self.channels), :self.nt, :grad_rec.shape[-1]] = grad_rec.to(self.dtype)
        p.adjoint(radj, want_src=False)
        g = p.finalize(vs.detach(), vp.detach(), rho.detach())
        return [g[i] if i < gp else None for i in range(3)]
