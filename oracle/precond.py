"""Gradient post-processing and preconditioners, restated in NumPy.
TEST INFRASTRUCTURE ONLY.

* gaussian_kernel_2d / gaussian_smooth_2d   fwi/loss.py:17-30
* illumination_from_energy                  fwi/illumination.py:38-48
* eprecond3                                 fwi/preconditioning.py:65-139
* eprecond3_multiplier                      fwi/optim_lbfgsb.py:90-100
"""
from __future__ import annotations

import numpy as np


def gaussian_kernel_2d(sigma, truncate=3.0):
    radius = int(truncate * sigma + 0.5)
    x = np.arange(2 * radius + 1) - radius
    g1 = np.exp(-0.5 * (x / sigma) ** 2)
    g2 = np.outer(g1, g1)
    return g2 / g2.sum()


def gaussian_smooth_2d(field, sigma):
    """'same'-mode zero-padded 2-D convolution with the normalised kernel."""
    k = gaussian_kernel_2d(sigma)
    r = k.shape[0] // 2
    f = np.asarray(field, np.float64)
    nz, nx = f.shape
    fp = np.zeros((nz + 2 * r, nx + 2 * r))
    fp[r:r + nz, r:r + nx] = f
    out = np.zeros_like(f)
    # symmetric kernel: convolution equals correlation
    for i in range(2 * r + 1):
        for j in range(2 * r + 1):
            out += k[i, j] * fp[i:i + nz, j:j + nx]
    return out


def illumination_from_energy(Ws_padded, nz, nx, pad, eps=0.001):
    """H_inv = 1/(Ws + eps*max Ws) normalised to max 1; the crop assumes a
    free surface exactly as the reference does (illumination.py:38-41)."""
    Ws = np.asarray(Ws_padded)[:nz, pad:pad + nx]
    lam = eps * np.max(Ws)
    H = 1.0 / (Ws + lam)
    return H / np.max(H)


def eprecond3(Ws, rec_x, rec_z, dx, dz, free_surface, pad, eps=5e-7):
    Ws = np.asarray(Ws)
    nz_t, nx_t = Ws.shape
    z0 = 0 if free_surface else pad
    cz, cx = np.meshgrid((np.arange(nz_t) - z0) * dz, (np.arange(nx_t) - pad) * dx, indexing="ij")
    rec_x = np.asarray(rec_x)
    rec_z = np.asarray(rec_z)
    if np.std(rec_x) >= np.std(rec_z):
        lmin, lmax, perp = float(np.min(rec_x) * dx), float(np.max(rec_x) * dx), float(rec_z[0] * dz)
        c_perp, c_line, sp = cz, cx, dz
    else:
        lmin, lmax, perp = float(np.min(rec_z) * dz), float(np.max(rec_z) * dz), float(rec_x[0] * dx)
        c_perp, c_line, sp = cx, cz, dx
    y = np.maximum(np.abs(c_perp - perp), sp)
    We = np.sqrt(Ws) * (np.arcsinh((lmax - c_line) / y) - np.arcsinh((lmin - c_line) / y))
    We = np.abs(We)
    return We + eps * float(np.max(We))


def eprecond3_multiplier(Ws, rec_x, rec_z, dx, dz, free_surface, pad, nz_dom, nx_dom, eps):
    We = eprecond3(Ws, rec_x, rec_z, dx, dz, free_surface, pad, eps)
    z0 = 0 if free_surface else pad
    p = 1.0 / We[z0:z0 + nz_dom, pad:pad + nx_dom]
    return p / np.max(p)
