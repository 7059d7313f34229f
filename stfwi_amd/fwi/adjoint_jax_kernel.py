"""Drop-in for fwi/adjoint_jax_kernel.py: the hand-coded discrete adjoint on
the HIP device.

    pml_apply_T(D_new_a, psi_new_a, k, b, a)            adjoint_jax_kernel.py:51
    adjoint_jax(Vs, Vp, rho, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx,
                dz, dt, nt, fc, pad, rec_vx_a, rec_vz_a, rec_ex_a, rec_ez_a,
                fd_order=2, free_surface=False, src_type='force',
                return_wavefields=False)                 adjoint_jax_kernel.py:65

adjoint_jax returns the source-wavelet adjoint (nt,); with
return_wavefields='velocity' also the adjoint velocity histories (nt, nz_t,
nx_t) U_a, V_a; with True also the three adjoint strain histories.  Frame
`it` is the adjoint state at the carry-out of reverse step `it`.
"""
from __future__ import annotations

import torch

from ._common import as_dev, as_int, as_int_array, compute_dtype, get_plan, device


def pml_apply_T(D_new_a, psi_new_a, k, b, a):
    """Transpose of the C-PML recursion psi' = b psi + a D, D' = D/k + psi'
    (elementwise; adjoint_jax_kernel.py:51-54)."""
    D_a = (1.0 / k + a) * D_new_a + a * psi_new_a
    psi_a = b * (D_new_a + psi_new_a)
    return D_a, psi_a


def adjoint_jax(Vs, Vp, rho, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt, fc, pad,
                rec_vx_a, rec_vz_a, rec_ex_a, rec_ez_a, fd_order=2, free_surface=False, src_type='force',
                return_wavefields=False):
    if fd_order not in (2, 4, 8):
        raise NotImplementedError(f"fd_order={fd_order} not supported (use 2, 4, or 8).")
    if src_type not in ("force", "moment"):
        raise NotImplementedError(f"src_type={src_type!r} not supported (use 'force' or 'moment').")
    dtype = compute_dtype()
    rx, rz = as_int_array(rec_x), as_int_array(rec_z)
    nt = int(nt)
    p = get_plan(nz_dom=int(nz_dom), nx_dom=int(nx_dom), dx=float(dx), dz=float(dz), dt=float(dt), fc=float(fc),
                 pad=int(pad), fd_order=int(fd_order), free_surface=bool(free_surface), dtype=dtype, n_shots_max=1,
                 nt_max=nt, nrec_max=max(int(rx.size), 1), imaging="none", grad="vs", store_history=False,
                 device=device())
    p.set_model(Vs, Vp, rho)
    p.set_shots([as_int(src_x)], [as_int(src_z)], src_type, nt, rx, rz, torch.zeros(nt, dtype=dtype, device=device()))
    ra = torch.zeros((1, 4, nt, p.nrec_max), dtype=dtype, device=device())
    for ch, y in enumerate((rec_vx_a, rec_vz_a, rec_ex_a, rec_ez_a)):
        ra[0, ch, :, :rx.size] = as_dev(y, dtype).reshape(nt, rx.size)
    if return_wavefields == "velocity":
        snap = ("vx", "vz")
    elif return_wavefields:
        snap = ("vx", "vz", "ex", "ez", "es")
    else:
        snap = ()
    out = p.adjoint(ra, want_src=True, snap=snap)
    src_a = out["src_adj"][0, :nt]
    if not snap:
        return src_a
    return (src_a,) + tuple(out["snap"][0, i, :nt] for i in range(len(snap)))
