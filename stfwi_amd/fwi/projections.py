"""Drop-in for fwi/projections.py: feasibility projections of the
velocity model applied before every gradient evaluation of the plain method.

    k_safe_ratio(K_margin=0.04)                      fwi/projections.py:36-48
    cfl_vp_cap(dx, dz, dt, cfl_coeff)               fwi/projections.py:51-62
    project_velocity(Vs, Vp, rho, K_margin=0.04, vp_cap=None,
                     vs_floor=1e-3, rho_floor=1e-3)  fwi/projections.py:69-117

Bulk modulus K = rho (Vp^2 - 4/3 Vs^2) > 0 needs Vp/Vs > 2/sqrt(3); the
projection keeps Vp/Vs >= (1 + K_margin) 2/sqrt(3) by lowering Vs, after
first capping Vp at the CFL limit, then applies floors to Vs and rho.
Host float64 arithmetic on (nz, nx) arrays, as in the reference.
"""
from __future__ import annotations

import math

import numpy as np


def k_safe_ratio(K_margin: float = 0.04) -> float:
    """Smallest allowed Vp/Vs: (1 + K_margin) * 2 / sqrt(3)."""
    return (1.0 + float(K_margin)) * 2.0 / math.sqrt(3.0)


def cfl_vp_cap(dx: float, dz: float, dt: float, cfl_coeff: float) -> float:
    """Largest stable Vp: cfl_coeff * min(dx, dz) / dt."""
    return float(cfl_coeff) * min(float(dx), float(dz)) / float(dt)


def project_velocity(Vs, Vp, rho, K_margin: float = 0.04, vp_cap=None, vs_floor: float = 1e-3,
                     rho_floor: float = 1e-3):
    """Projected float64 copies (Vs, Vp, rho)."""
    vp = np.array(Vp, dtype=np.float64, copy=True)
    vs = np.array(Vs, dtype=np.float64, copy=True)
    rh = np.array(rho, dtype=np.float64, copy=True)
    if vp_cap is not None:
        vp = np.minimum(vp, float(vp_cap))
    vs = np.minimum(vs, vp / k_safe_ratio(K_margin))
    vs = np.maximum(vs, float(vs_floor))
    rh = np.maximum(rh, float(rho_floor))
    return vs, vp, rh


def project_velocity_device(vs, vp, rho, K_margin: float = 0.04, vp_cap=None, vs_floor: float = 1e-3,
                            rho_floor: float = 1e-3):
    """project_velocity on float64 device tensors (new tensors; the same
    order of operations: Vp cap, then the Vs ceiling Vp / ratio, then the
    floors)."""
    import torch
    vp = vp.to(torch.float64).clone()
    if vp_cap is not None:
        vp.clamp_(max=float(vp_cap))
    vs = torch.minimum(vs.to(torch.float64), vp / k_safe_ratio(K_margin)).clamp_(min=float(vs_floor))
    rh = rho.to(torch.float64).clamp(min=float(rho_floor))
    return vs, vp, rh
