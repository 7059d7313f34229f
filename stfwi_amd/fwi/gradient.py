"""Drop-in for fwi/gradient.py: compute_gradient on the HIP device.

    compute_gradient(backend, *, Vs, Vp, rho, src_x, src_z, src_wavelet,
                     observed, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt,
                     fc, pad, component=2, invert='vs', scale=1.0, sos=None,
                     n_taper=0, fd_order=2, free_surface=False,
                     src_type='force', block_size=None,
                     grad_smooth_sigma=None, source_taper=None,
                     illumination=None)              fwi/gradient.py:48-99
    -> ((g_Vs, g_Vp, g_rho) as numpy (nz_dom, nx_dom), float loss)

Both backends run every shot in one batched launch: forward with the strain
history, residual prep (taper, DF-I Butterworth cascade, L2 misfit and its
transpose), and the adjoint with the imaging condition fused into each
reverse step.  When the strain history of all shots does not fit in the free
device memory, the plan checkpoints the forward state and recomputes the
history segment by segment (the library's planner picks the spacing; the
results do not depend on it).  backend='adjoint' uses the Zhou imaging of
fwi/imaging.py:83-169 (the reference's hand-coded gradient); backend='ad'
uses the exact discrete gradient, i.e. what jax.grad returns in
fwi/gradient.py:106-132.

Reference behaviour kept for drop-in fidelity:
  * the forward always models a force source (build_forward_fn has no
    src_type, fwi/forward.py:508-544); src_type only changes the adjoint's
    source extraction.  honour_src_type=True models src_type in the forward.
  * backend='ad' uses src_z[0] for every shot (fwi/loss.py:90).
  * the Zhou chain rule forms 2 rho Vs in the model's storage dtype and the
    Zhou buoyancy 1/rho in that dtype too (fwi/imaging.py:113-114,165-167).
  * backend='adjoint' returns float64 arrays, backend='ad' the model dtype.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from ._common import as_dev, as_int_array, compute_dtype, get_plan, storage_dtype, device, tune_launches
from .loss import gaussian_smooth_2d


def _postprocess(grads, grad_smooth_sigma, source_taper, illumination):
    """fwi/gradient.py:266-277: smooth, then source taper, then illumination."""
    out = list(grads)
    if grad_smooth_sigma is not None:
        out = [gaussian_smooth_2d(g, grad_smooth_sigma) for g in out]
    if source_taper is not None:
        st = as_dev(source_taper, out[0].dtype)
        out = [g * st for g in out]
    if illumination is not None:
        il = as_dev(illumination, out[0].dtype)
        out = [g * il for g in out]
    return out


def compute_gradient(backend: str, *, Vs, Vp, rho, src_x, src_z, src_wavelet, observed, rec_x, rec_z,
                     nx_dom, nz_dom, dx, dz, dt, nt, fc, pad, component: int = 2, invert: str = "vs",
                     scale: float = 1.0, sos=None, n_taper: int = 0, fd_order: int = 2,
                     free_surface: bool = False, src_type: str = "force", block_size: Optional[int] = None,
                     grad_smooth_sigma: Optional[float] = None, source_taper=None, illumination=None,
                     honour_src_type: bool = False, return_tensors: bool = False):
    if backend not in ("ad", "adjoint"):
        raise ValueError(f"unknown backend {backend!r}; use 'ad' or 'adjoint'")
    if invert not in ("vs", "vs_vp", "all"):
        raise ValueError(f"invert={invert!r}; use 'vs', 'vs_vp', or 'all'")
    if backend == "adjoint" and fd_order not in (2, 4, 8):
        raise NotImplementedError(f"fd_order={fd_order} not supported (use 2, 4, or 8).")
    if src_type not in ("force", "moment"):
        raise NotImplementedError(f"src_type={src_type!r} not supported (use 'force' or 'moment').")
    fd = 8 if fd_order == 8 else (4 if fd_order >= 4 else 2)
    dtype = compute_dtype()
    sx = as_int_array(src_x)
    sz = as_int_array(src_z)
    n = sx.size
    if backend == "ad":
        sz = np.full(n, sz[0], np.int32)
    rx, rz = as_int_array(rec_x), as_int_array(rec_z)
    nrec = rx.size
    nt = int(nt)
    p = get_plan(nz_dom=int(nz_dom), nx_dom=int(nx_dom), dx=float(dx), dz=float(dz), dt=float(dt), fc=float(fc),
                 pad=int(pad), fd_order=fd, free_surface=bool(free_surface), dtype=dtype, n_shots_max=n,
                 nt_max=nt, nrec_max=nrec, imaging="zhou" if backend == "adjoint" else "exact", grad=invert,
                 store_history=True, device=device(), ckpt_every="auto")
    sdt = storage_dtype(Vs)
    b_img = None
    if backend == "adjoint":
        b_img = (1.0 / as_dev(rho, storage_dtype(rho))).to(dtype)
    p.set_model(Vs, Vp, rho, b_img)
    p.set_shots(sx, sz, src_type if honour_src_type else "force", nt, rx, rz, as_dev(src_wavelet, dtype))
    tune_launches(p, n)
    rec = p.forward(channels=(int(component),))["rec"]
    obs = torch.zeros((n, nt, p.nrec_max), dtype=dtype, device=device())
    obs[:, :, :nrec] = as_dev(observed, dtype).reshape(n, nt, nrec)
    loss, radj = p.residual(rec, obs, int(component), sos, int(n_taper), float(scale), n)
    p.adjoint(radj, want_src=False)
    g = p.finalize(as_dev(Vs, sdt), as_dev(Vp, sdt), as_dev(rho, sdt))
    out_dtype = torch.float64 if backend == "adjoint" else sdt
    grads = [g[i].to(out_dtype) for i in range(3)]
    grads = _postprocess(grads, grad_smooth_sigma, source_taper, illumination)
    loss_total = float(loss.double().sum())
    if return_tensors:
        return tuple(grads), loss_total
    return tuple(x.cpu().numpy() for x in grads), loss_total
