"""Drop-in for fwi/imaging.py: Zhou imaging of given history cubes.

    imaging_condition(Vs, Vp, rho, ex_hist, ez_hist, es_hist, U_a_hist,
                      V_a_hist, dx, dz, dt, pad, free_surface, nz_dom, nx_dom,
                      vx_hist=None, vz_hist=None, time_chunk=64)
        -> g_Vs, g_Vp, g_rho, g_lam, g_mu, g_rho_lame   fwi/imaging.py:83-169

The gradient path fuses this imaging into the adjoint kernel; this entry
takes cubes as the reference does (time_chunk is accepted and ignored).
The buoyancy and the chain-rule factors are formed in the model's own dtype,
as the reference does.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _abi
from ._common import as_dev, compute_dtype, storage_dtype


def _extend(field, pad, free_surface):
    """Nearest-neighbour extension into the PML (fwi/imaging.py:56-80)."""
    nz_dom, nx_dom = field.shape
    z0 = 0 if free_surface else pad
    nz_t = nz_dom + (pad if free_surface else 2 * pad)
    rr = (torch.arange(nz_t, device=field.device) - z0).clamp(0, nz_dom - 1)
    cc = (torch.arange(nx_dom + 2 * pad, device=field.device) - pad).clamp(0, nx_dom - 1)
    return field[rr][:, cc].contiguous()


def imaging_condition(Vs, Vp, rho, ex_hist, ez_hist, es_hist, U_a_hist, V_a_hist, dx, dz, dt, pad, free_surface,
                      nz_dom, nx_dom, vx_hist=None, vz_hist=None, time_chunk=64):
    dtype = ex_hist.dtype if isinstance(ex_hist, torch.Tensor) else compute_dtype()
    cubes = [as_dev(x, dtype) for x in (ex_hist, ez_hist, es_hist, U_a_hist, V_a_hist)]
    vel = [as_dev(x, dtype) for x in (vx_hist, vz_hist)] if vx_hist is not None and vz_hist is not None else [None, None]
    nt, nz_t, nx_t = cubes[0].shape
    sdt = storage_dtype(rho)
    rho_s = as_dev(rho, sdt)
    b_full = (1.0 / _extend(rho_s, int(pad), bool(free_surface))).to(dtype).contiguous()
    out = torch.empty((3, nz_t, nx_t), dtype=dtype, device=cubes[0].device)
    code = _abi.F64 if dtype == torch.float64 else _abi.F32
    P = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    _abi.check(_abi.lib().stfwi_zhou_imaging(*[P(t) for t in cubes], P(b_full), P(vel[0]), P(vel[1]), nt, nz_t, nx_t,
                                             float(dx), float(dz), float(dt), code, P(out),
                                             C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    z0 = 0 if free_surface else pad
    g_lam, g_mu, g_rl = (out[i, z0:z0 + nz_dom, pad:pad + nx_dom] for i in range(3))
    vs, vp = as_dev(Vs, sdt), as_dev(Vp, sdt)
    g_vs = 2.0 * rho_s * vs * (g_mu - 2.0 * g_lam)
    g_vp = 2.0 * rho_s * vp * g_lam
    g_rho = (vp ** 2 - 2.0 * vs ** 2) * g_lam + vs ** 2 * g_mu + g_rl
    return g_vs, g_vp, g_rho, g_lam, g_mu, g_rl
