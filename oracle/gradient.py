"""compute_gradient restated on the C oracle.  TEST INFRASTRUCTURE ONLY.

Mirrors fwi/gradient.py:48-277:
* backend='adjoint' (fwi/gradient.py:159-259): per shot, forward with the
  strain history, residual prep, hand-coded adjoint with Zhou imaging,
  float64 accumulation, then the ``invert`` masking.
* backend='ad' (fwi/gradient.py:106-132 via fwi/loss.py:33-152): the exact
  discrete gradient of the same misfit.  The reference gets it from jax.grad;
  here it is the adjoint with the exact imaging terms.  As in the reference,
  every shot uses src_z[0] (fwi/loss.py:90) and the gradient of frozen
  parameters is zero (stop_gradient, fwi/loss.py:106-110).
* Both backends model a force source in the forward, whatever ``src_type``
  says (build_forward_fn has no src_type parameter, fwi/forward.py:508-544);
  ``honour_src_type=True`` switches that off.
* _postprocess (fwi/gradient.py:266-277): smooth, source taper, illumination.
"""
from __future__ import annotations

import numpy as np

from .core import Geom, forward, adjoint, exact_chain, zhou_chain, padded_shape
from .signal import residual_prep
from .precond import gaussian_smooth_2d


def _embed_rho_dtype(rho, pad, free_surface):
    """fwi/imaging.py:56-80 extends rho in ITS OWN dtype; B = 1/rho there."""
    rho = np.asarray(rho)
    nzd, nxd = rho.shape
    nz, nx = padded_shape(nzd, nxd, pad, free_surface)
    z0 = 0 if free_surface else pad
    f = np.zeros((nz, nx), rho.dtype)
    f[z0:z0 + nzd, pad:pad + nxd] = rho
    f[z0:z0 + nzd, :pad] = f[z0:z0 + nzd, pad:pad + 1]
    f[z0:z0 + nzd, nx - pad:] = f[z0:z0 + nzd, pad + nxd - 1:pad + nxd]
    f[z0 + nzd:, :] = f[z0 + nzd - 1:z0 + nzd, :]
    if not free_surface:
        f[:pad, :] = f[pad:pad + 1, :]
    return 1.0 / f


def compute_gradient(backend, *, Vs, Vp, rho, src_x, src_z, src_wavelet, observed,
                     rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt, fc, pad,
                     component=2, invert="vs", scale=1.0, sos=None, n_taper=0,
                     fd_order=2, free_surface=False, src_type="force", block_size=None,
                     grad_smooth_sigma=None, source_taper=None, illumination=None,
                     honour_src_type=False, dtype=np.float64):
    if backend not in ("ad", "adjoint"):
        raise ValueError(f"unknown backend {backend!r}; use 'ad' or 'adjoint'")
    if invert not in ("vs", "vs_vp", "all"):
        raise ValueError(f"invert={invert!r}; use 'vs', 'vs_vp', or 'all'")
    g = Geom(nz_dom, nx_dom, pad, free_surface, fd_order, nt, dx, dz, dt, fc)
    src_x = np.asarray(src_x).astype(np.int32)
    src_z = np.asarray(src_z).astype(np.int32)
    n_shots = len(src_x)
    fwd_src = src_type if honour_src_type else "force"
    wav = np.asarray(src_wavelet, np.float64)
    acc = [np.zeros((nz_dom, nx_dom)) for _ in range(3)]
    loss_total = 0.0
    B_img = _embed_rho_dtype(rho, pad, free_surface) if backend == "adjoint" else None
    for k in range(n_shots):
        sx = int(src_x[k])
        sz = int(src_z[0] if backend == "ad" else src_z[k])
        tape = ("e", "uv") if backend == "adjoint" else ("e", "dxu0", "s")
        rec, t = forward(g, Vs, Vp, rho, sx, sz, wav, rec_x, rec_z, fwd_src, dtype, tape)
        pred = rec[component]
        lk, cot = residual_prep(pred, np.asarray(observed[k], np.float64), sos, n_taper, scale, n_shots)
        loss_total += lk
        ra = np.zeros_like(rec)
        ra[component] = cot
        if backend == "adjoint":
            a = adjoint(g, Vs, Vp, rho, sx, sz, ra, rec_x, rec_z, src_type, wav, "zhou", t, B_img, dtype)
            gk = zhou_chain(g, a["g"], Vs, Vp, rho)
        else:
            a = adjoint(g, Vs, Vp, rho, sx, sz, ra, rec_x, rec_z, fwd_src, wav, "exact", t, None, dtype)
            gk = exact_chain(g, a["g"], Vs, Vp, rho)
        for i in range(3):
            acc[i] += np.asarray(gk[i], np.float64)
    if invert == "vs":
        acc[1][:] = 0
        acc[2][:] = 0
    elif invert == "vs_vp":
        acc[2][:] = 0
    grads = acc
    if backend == "ad":
        grads = [a.astype(np.asarray(v).dtype) for a, v in zip(acc, (Vs, Vp, rho))]
    out = list(grads)
    if grad_smooth_sigma is not None:
        out = [gaussian_smooth_2d(x, grad_smooth_sigma) for x in out]
    if source_taper is not None:
        out = [x * np.asarray(source_taper) for x in out]
    if illumination is not None:
        out = [x * np.asarray(illumination) for x in out]
    return tuple(out), float(loss_total)
