"""Drop-in for fwi/loss.py.

    gaussian_smooth_2d(field, sigma)        fwi/loss.py:27-30 (HIP kernel)
    build_loss_fn(run_shot, observed, src_x_list, src_z, src_wavelet,
                  component=1, scale=1e20, invert='vs', sos=None, n_taper=0,
                  source_taper=None, grad_smooth_sigma=None,
                  tikhonov_alpha=None, illumination=None)  fwi/loss.py:33-152

build_loss_fn returns loss_fn(params=(Vs, Vp, rho)) -> (loss, (g_Vs, g_Vp,
g_rho)) as the reference's value_and_grad does: the exact discrete gradient
of the multi-shot misfit, every shot using src_z[0] (fwi/loss.py:90), frozen
parameters with zero gradient, optional Tikhonov term and post-processing.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _abi
from ._common import as_dev, compute_dtype, device


def gaussian_smooth_2d(field, sigma):
    """Normalised Gaussian of radius int(3 sigma + 0.5), 'same'-mode zero
    padding, on the HIP device; returns a tensor of the field's dtype."""
    f = field if isinstance(field, torch.Tensor) and field.is_cuda else as_dev(field, compute_dtype())
    if f.dtype not in (torch.float32, torch.float64):
        f = f.to(compute_dtype())
    f = f.contiguous()
    nz, nx = f.shape
    tmp = torch.empty_like(f)
    out = torch.empty_like(f)
    code = _abi.F64 if f.dtype == torch.float64 else _abi.F32
    _abi.check(_abi.lib().stfwi_smooth2d(C.c_void_p(f.data_ptr()), C.c_void_p(tmp.data_ptr()),
                                         C.c_void_p(out.data_ptr()), nz, nx, float(sigma), code,
                                         C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out


def _tikhonov(params, invert):
    """reg = sum of squared first differences along z and x of the inverted
    parameters, and its gradient (fwi/loss.py:123-128)."""
    active = {"vs": 1, "vs_vp": 2, "all": 3}[invert]
    reg = 0.0
    grads = []
    for i, m in enumerate(params):
        g = torch.zeros_like(m)
        if i < active:
            dz = m[1:, :] - m[:-1, :]
            dx = m[:, 1:] - m[:, :-1]
            reg = reg + float((dz.double() ** 2).sum() + (dx.double() ** 2).sum())
            g[1:, :] += 2 * dz
            g[:-1, :] -= 2 * dz
            g[:, 1:] += 2 * dx
            g[:, :-1] -= 2 * dx
        grads.append(g)
    return reg, grads


def build_loss_fn(run_shot, observed, src_x_list, src_z, src_wavelet, component=1, scale=1e20, invert='vs',
                  sos=None, n_taper=0, source_taper=None, grad_smooth_sigma=None, tikhonov_alpha=None,
                  illumination=None):
    from .gradient import compute_gradient, _postprocess
    rs = run_shot

    def loss_fn(params):
        Vs, Vp, rho = params
        grads, loss = compute_gradient(
            "ad", Vs=Vs, Vp=Vp, rho=rho, src_x=src_x_list, src_z=src_z, src_wavelet=src_wavelet,
            observed=observed, rec_x=rs.rec_x, rec_z=rs.rec_z, nx_dom=rs.nx_dom, nz_dom=rs.nz_dom, dx=rs.dx,
            dz=rs.dz, dt=rs.dt, nt=rs.nt, fc=rs.fc, pad=rs.pad, component=component, invert=invert,
            scale=scale, sos=sos, n_taper=n_taper, fd_order=rs.fd_order, free_surface=rs.free_surface,
            return_tensors=True)
        grads = list(grads)
        if tikhonov_alpha is not None:
            dt = grads[0].dtype
            reg, rg = _tikhonov([as_dev(x, dt) for x in params], invert)
            w = tikhonov_alpha * (loss / (reg + 1e-30))
            loss = loss + w * reg
            grads = [g + w * r for g, r in zip(grads, rg)]
        grads = _postprocess(grads, grad_smooth_sigma, source_taper, illumination)
        return loss, tuple(grads)

    return loss_fn
