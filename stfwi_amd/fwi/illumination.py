"""Drop-in for fwi/illumination.py.

    compute_illumination_for_stage(nz, nx, dx, dz, dt, nt, f_c, pad,
        block_size, rec_x, rec_z, fd_order, component, Vs, Vp, rho, src_x,
        src_z, src_wavelet, eps=0.001)              fwi/illumination.py:18-48

All shots run in one batched forward whose strain kernel accumulates the
energy sum_t u^2 of the chosen component on the fly (no wavefield cube is
stored); a HIP kernel sums the shots, crops and forms H = 1/(Ws + eps max Ws)
/ max.  As in the reference the forward uses a free surface and the crop
assumes one (illumination.py:29-31,36,44).
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _abi
from ._common import as_dev, as_int_array, compute_dtype, forward_fd_order, get_plan, device, tune_launches


def shot_energy(nz, nx, dx, dz, dt, nt, f_c, pad, rec_x, rec_z, fd_order, component, Vs, Vp, rho, src_x, src_z,
                src_wavelet, free_surface=True, src_type="force"):
    """(n_shots, nz_t, nx_t) energies sum_t u_comp^2 on the padded grid."""
    dtype = compute_dtype()
    sx, sz = as_int_array(src_x), as_int_array(src_z)
    rx = as_int_array(rec_x)
    p = get_plan(nz_dom=int(nz), nx_dom=int(nx), dx=float(dx), dz=float(dz), dt=float(dt), fc=float(f_c),
                 pad=int(pad), fd_order=forward_fd_order(fd_order), free_surface=bool(free_surface), dtype=dtype,
                 n_shots_max=sx.size, nt_max=int(nt), nrec_max=max(int(rx.size), 1), imaging="none", grad="vs",
                 store_history=False, device=device())
    p.set_model(Vs, Vp, rho)
    p.set_shots(sx, sz, src_type, int(nt), rx, as_int_array(rec_z), as_dev(src_wavelet, dtype))
    tune_launches(p, sx.size)
    return p.forward(channels=(), illum_comp=int(component))["illum"][:sx.size]


def compute_illumination_for_stage(nz, nx, dx, dz, dt, nt, f_c, pad, block_size, rec_x, rec_z, fd_order, component,
                                   Vs, Vp, rho, src_x, src_z, src_wavelet, eps=0.001):
    ws = shot_energy(nz, nx, dx, dz, dt, nt, f_c, pad, rec_x, rec_z, fd_order, component, Vs, Vp, rho, src_x,
                     src_z, src_wavelet, free_surface=True)
    S, nz_t, nx_t = ws.shape
    out = torch.empty((int(nz), int(nx)), dtype=ws.dtype, device=ws.device)
    code = _abi.F64 if ws.dtype == torch.float64 else _abi.F32
    _abi.check(_abi.lib().stfwi_illumination(C.c_void_p(ws.data_ptr()), S, nz_t, nx_t, 0, int(pad), int(nz),
                                             int(nx), float(eps), code, C.c_void_p(out.data_ptr()),
                                             C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out
