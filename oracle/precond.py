"""Gradient post-processing and preconditioners, restated in NumPy.
TEST INFRASTRUCTURE ONLY.

* gaussian_kernel_2d / gaussian_smooth_2d   fwi/loss.py:17-30
* illumination_from_energy                  fwi/illumination.py:38-48
* eprecond3                                 fwi/preconditioning.py:65-139
* eprecond3_multiplier                      fwi/optim_lbfgsb.py:90-100
* eprecond1                                 fwi/preconditioning.py:37-62
* taper_grad                                fwi/preconditioning.py:144-183
* source_taper                              fwi/taper.py:16-84
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


def eprecond1(Ws, Wr, eps=5e-7):
    """sqrt(Ws Wr) plus eps times its maximum."""
    we = np.sqrt(np.asarray(Ws) * np.asarray(Wr))
    return we + eps * float(np.max(we))


def taper_grad(shape, dz, gradt1, gradt2, exp_taper=2.0, gradb1=None, gradb2=None):
    """Per-row depth taper broadcast along x (float64)."""
    nz, nx = shape
    t = np.empty(nz)
    for j in range(nz):
        if j < gradt1:
            t[j] = 0.0
        elif j <= gradt2:
            t[j] = np.exp(-float(gradt2 - j) ** 2)
        else:
            t[j] = (j * dz) ** exp_taper
    if gradb1 is not None and gradb2 is not None:
        den = max(float(gradb2 - gradb1), 1.0)
        for j in range(nz):
            if j > gradb2:
                t[j] = 0.0
            elif j >= gradb1:
                t[j] = t[j] * np.exp(-float(j - gradb1) ** 2 / den ** 2 * 4.0)
    return np.repeat(t[:, None], nx, axis=1)


def source_patch(dh, srt_radius):
    """The normalised erf patch: erf(r / (sqrt(2) R)) on a ps x ps grid of
    cell centres, shifted to minimum 0, divided by the largest edge-centre
    value, clipped to [0, 1]."""
    from math import erf
    ps = max(int(np.ceil(2.0 * srt_radius / dh)), 5)
    c = ps / 2.0
    mr = np.sqrt(2.0) * srt_radius
    p = np.empty((ps, ps))
    for jj in range(ps):
        for ii in range(ps):
            x = (ii - c + 0.5) * dh
            y = (jj - c + 0.5) * dh
            p[jj, ii] = erf(np.sqrt(x * x + y * y) / mr)
    p = p - p.min()
    m = ps // 2
    p = p / max(p[0, m], p[m, 0], p[m, -1], p[-1, m])
    return np.clip(p, 0.0, 1.0)


def source_taper(nz, nx, dh, src_positions, srt_radius, filt_size=2):
    """Minimum of the patches stamped at each source, normalised to [0, 1],
    then zero boxes of half-width filt_size at the sources."""
    p = source_patch(dh, srt_radius)
    ps = p.shape[0]
    h = ps // 2
    t = np.ones((nz, nx))
    for sx, sz in src_positions:
        for pj in range(ps):
            for pi in range(ps):
                gx, gz = sx + pi - h, sz + pj - h
                if 0 <= gx < nx and 0 <= gz < nz:
                    t[gz, gx] = min(t[gz, gx], p[pj, pi])
    t = t - t.min()
    if t.max() > 0:
        t = t / t.max()
    for sx, sz in src_positions:
        t[max(sz - filt_size, 0):min(sz + filt_size + 1, nz), max(sx - filt_size, 0):min(sx + filt_size + 1, nx)] = 0.0
    return t
