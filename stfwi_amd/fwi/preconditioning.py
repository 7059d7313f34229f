"""Drop-ins for fwi/preconditioning.py and the eprecond3 driver
(fwi/optim_lbfgsb.py:46-100), on the HIP device.

    eprecond1(Ws, Wr, eps=5e-7) -> We                       preconditioning.py:37-62
    eprecond3(Ws, rec_x, rec_z, dx, dz, free_surface, pad, eps=5e-7) -> We   :65-139
    taper_grad(shape, dz, gradt1, gradt2, exp_taper=2.0, gradb1=None, gradb2=None)   :144-183
    eprecond3_preconditioner(...) -> normalised physical multiplier
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _abi
from ._common import as_dev, as_int_array, compute_dtype


def _call(ws, rec_x, rec_z, dx, dz, free_surface, pad, nz_dom, nx_dom, eps):
    ws = ws if isinstance(ws, torch.Tensor) and ws.is_cuda else as_dev(ws, compute_dtype())
    if ws.dim() == 2:
        ws = ws.unsqueeze(0)
    ws = ws.contiguous()
    S, nz_t, nx_t = ws.shape
    rx = as_int_array(rec_x)
    rz = as_int_array(rec_z)
    we = torch.empty((nz_t, nx_t), dtype=ws.dtype, device=ws.device)
    out = torch.empty((nz_dom, nx_dom), dtype=ws.dtype, device=ws.device)
    code = _abi.F64 if ws.dtype == torch.float64 else _abi.F32
    _abi.check(_abi.lib().stfwi_eprecond3(C.c_void_p(ws.data_ptr()), S, nz_t, nx_t, int(bool(free_surface)), int(pad),
                                          int(nz_dom), int(nx_dom), rx.ctypes.data_as(C.c_void_p),
                                          rz.ctypes.data_as(C.c_void_p), rx.size, float(dx), float(dz), float(eps),
                                          code, C.c_void_p(we.data_ptr()), C.c_void_p(out.data_ptr()),
                                          C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return we, out


def eprecond3(Ws, rec_x, rec_z, dx, dz, free_surface, pad, eps=5e-7):
    """Plessix-Mulder weight We on the padded grid of Ws (nz_t, nx_t)."""
    Ws = Ws if isinstance(Ws, torch.Tensor) else as_dev(Ws, compute_dtype())
    nz_t, nx_t = Ws.shape[-2:]
    nz_dom = nz_t - (pad if free_surface else 2 * pad)
    return _call(Ws, rec_x, rec_z, dx, dz, free_surface, pad, nz_dom, nx_t - 2 * pad, eps)[0]


def eprecond3_preconditioner(*, nz_dom, nx_dom, dx, dz, dt, nt, f_c, pad, rec_x, rec_z, fd_order, component, Vs, Vp,
                             rho, src_x, src_z, src_wavelet, free_surface, eps, src_type="force"):
    """[0, 1] gradient multiplier 1/We / max on the physical grid
    (_compute_eprecond3_preconditioner, fwi/optim_lbfgsb.py:46-100)."""
    from .illumination import shot_energy
    ws = shot_energy(nz_dom, nx_dom, dx, dz, dt, nt, f_c, pad, rec_x, rec_z, fd_order, component, Vs, Vp, rho,
                     src_x, src_z, src_wavelet, free_surface=free_surface, src_type=src_type)
    return _call(ws, rec_x, rec_z, dx, dz, free_surface, pad, int(nz_dom), int(nx_dom), eps)[1]


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def eprecond1(Ws, Wr, eps=5e-7):
    """Shin et al. energy weight We = sqrt(Ws Wr) + eps max(sqrt(Ws Wr)), in the
    promoted floating dtype of Ws and Wr (fwi/preconditioning.py:37-62)."""
    from ._common import storage_dtype
    dt = torch.float64 if torch.float64 in (storage_dtype(Ws), storage_dtype(Wr)) else torch.float32
    ws, wr = as_dev(Ws, dt), as_dev(Wr, dt)
    if ws.shape != wr.shape:
        raise ValueError(f"Ws {tuple(ws.shape)} and Wr {tuple(wr.shape)} differ in shape")
    we = torch.empty_like(ws)
    code = _abi.F64 if dt == torch.float64 else _abi.F32
    _abi.check(_abi.lib().stfwi_eprecond1(C.c_void_p(ws.data_ptr()), C.c_void_p(wr.data_ptr()), ws.numel(),
                                          float(eps), code, C.c_void_p(we.data_ptr()), _stream()))
    return we


def taper_grad(shape, dz, gradt1, gradt2, exp_taper=2.0, gradb1=None, gradb2=None):
    """DENISE depth taper (nz, nx), float64 (fwi/preconditioning.py:144-183)."""
    from ._common import device
    nz, nx = (int(v) for v in shape)
    out = torch.empty((nz, nx), dtype=torch.float64, device=device())
    bottom = gradb1 is not None and gradb2 is not None
    _abi.check(_abi.lib().stfwi_taper_grad(nz, nx, float(dz), int(gradt1), int(gradt2), float(exp_taper),
                                           int(bottom), int(gradb1) if bottom else 0, int(gradb2) if bottom else 0,
                                           C.c_void_p(out.data_ptr()), _stream()))
    return out
