"""ctypes wrappers around the C oracle (liborc.so).  TEST INFRASTRUCTURE ONLY.

Shapes follow the reference: model (nz_dom, nx_dom) row-major (z = row),
receivers (nt, nrec), padded grid (nz_t, nx_t) with
nz_t = nz_dom + pad (free surface) or nz_dom + 2*pad, nx_t = nx_dom + 2*pad
(fwi/forward.py:185-191).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_DT = {np.dtype(np.float32): ("f32", C.c_float),
       np.dtype(np.float64): ("f64", C.c_double),
       np.dtype(np.complex128): ("c128", None)}


class _CGeom(C.Structure):
    _fields_ = [("nz_dom", C.c_int), ("nx_dom", C.c_int), ("pad", C.c_int),
                ("free_surface", C.c_int), ("fd_order", C.c_int), ("nt", C.c_int),
                ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double), ("fc", C.c_double)]


@dataclass
class Geom:
    nz_dom: int
    nx_dom: int
    pad: int
    free_surface: bool
    fd_order: int
    nt: int
    dx: float
    dz: float
    dt: float
    fc: float

    def c(self):
        return _CGeom(int(self.nz_dom), int(self.nx_dom), int(self.pad), int(bool(self.free_surface)),
                      int(self.fd_order), int(self.nt), float(self.dx), float(self.dz),
                      float(self.dt), float(self.fc))

    @property
    def shape(self):
        return padded_shape(self.nz_dom, self.nx_dom, self.pad, self.free_surface)

    @property
    def z0(self):
        return 0 if self.free_surface else self.pad


def padded_shape(nz_dom, nx_dom, pad, free_surface):
    return (nz_dom + (pad if free_surface else 2 * pad), nx_dom + 2 * pad)


def build(force=False):
    """Compile liborc.so with the Makefile beside this file."""
    so = os.path.join(_HERE, "liborc.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liborc.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return so


def _lib():
    global _LIB
    if _LIB is None:
        so = build()
        _LIB = C.CDLL(so)
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _fn(name, sfx):
    f = getattr(_lib(), f"{name}_{sfx}")
    f.restype = C.c_int
    return f


def _as(a, dt):
    return np.ascontiguousarray(np.asarray(a), dtype=dt)


def setup(g: Geom, vs, vp, rho, dtype=np.float64):
    """Padded moduli (L2M, L, M, B) and C-PML profiles (fwi/forward.py:193-238)."""
    dt = np.dtype(dtype)
    sfx = _DT[dt][0]
    nz, nx = g.shape
    out = [np.zeros((nz, nx), dt) for _ in range(4)]
    prof = np.zeros(3 * nx + 3 * nz)
    _fn("orc_setup", sfx)(C.byref(g.c()), _ptr(_as(vs, dt)), _ptr(_as(vp, dt)), _ptr(_as(rho, dt)),
                           *[_ptr(o) for o in out], _ptr(prof))
    kx, bx, ax = prof[:nx], prof[nx:2 * nx], prof[2 * nx:3 * nx]
    kz, bz, az = prof[3 * nx:3 * nx + nz], prof[3 * nx + nz:3 * nx + 2 * nz], prof[3 * nx + 2 * nz:]
    return dict(L2M=out[0], L=out[1], M=out[2], B=out[3], kx=kx, bx=bx, ax=ax, kz=kz, bz=bz, az=az)


def forward(g: Geom, vs, vp, rho, sx, sz, wavelet, rec_x, rec_z, src_type="force",
            dtype=np.float64, tape=()):
    """Run the forward (fwi/forward.py:119-501).

    Returns (rec, tape_dict): rec is (4, nt, nrec) = (vx, vz, exx, ezz) at the
    receivers after each step; tape_dict holds the requested histories among
    'uv' (nt,2,nz,nx), 'e' (nt,3,nz,nx), 's' (nt,2,nz,nx: dU/dB, dV/dB of the
    step) and 'dxu0' (nt,nx: C-PML'd D-x U on row 0).
    """
    dt = np.dtype(dtype)
    sfx = _DT[dt][0]
    nz, nx = g.shape
    nt = g.nt
    rx = _as(rec_x, np.int32)
    rz = _as(rec_z, np.int32)
    nrec = rx.size
    rec = np.zeros((4, nt, nrec), dt)
    t = {}
    if "uv" in tape:
        t["uv"] = np.zeros((nt, 2, nz, nx), dt)
    if "e" in tape:
        t["e"] = np.zeros((nt, 3, nz, nx), dt)
    if "s" in tape:
        t["s"] = np.zeros((nt, 2, nz, nx), dt)
    if "dxu0" in tape:
        t["dxu0"] = np.zeros((nt, nx), dt)
    st = {"force": 0, "moment": 1}[src_type]
    rc = _fn("orc_forward", sfx)(C.byref(g.c()), _ptr(_as(vs, dt)), _ptr(_as(vp, dt)), _ptr(_as(rho, dt)),
                                  C.c_int(int(sx)), C.c_int(int(sz)), C.c_int(st), _ptr(_as(wavelet, dt)),
                                  C.c_int(nrec), _ptr(rx), _ptr(rz), _ptr(rec),
                                  _ptr(t.get("uv")), _ptr(t.get("e")), _ptr(t.get("s")), _ptr(t.get("dxu0")))
    assert rc == 0
    return rec, t


def adjoint(g: Geom, vs, vp, rho, sx, sz, rec_adj, rec_x, rec_z, src_type="force",
            wavelet=None, imaging=None, tape=None, B_img=None, dtype=np.float64,
            return_wavefields=False):
    """Hand-coded adjoint (fwi/adjoint_jax_kernel.py:65-483).

    rec_adj: (4, nt, nrec) receiver adjoints (vx, vz, exx, ezz).
    imaging: None, 'zhou' (fwi/imaging.py) or 'exact'; needs ``tape`` from
    forward() ('e' always; 'uv' for the Zhou density term; 'dxu0' and 's' for
    the exact free-surface and buoyancy terms).
    Returns dict with 'src_a' (nt,), 'g' (4,nz,nx) padded accumulators when
    imaging is set, and 'ua'/'ea' histories when return_wavefields.
    """
    dt = np.dtype(dtype)
    sfx = _DT[dt][0]
    nz, nx = g.shape
    nt = g.nt
    rx = _as(rec_x, np.int32)
    rz = _as(rec_z, np.int32)
    ra = _as(rec_adj, dt)
    assert ra.shape == (4, nt, rx.size)
    src_a = np.zeros(nt, dt)
    tape = tape or {}
    mode = {None: 0, "zhou": 1, "exact": 2}[imaging]
    gacc = np.zeros((4, nz, nx), dt) if mode else None
    ua = np.zeros((nt, 2, nz, nx), dt) if return_wavefields else None
    ea = np.zeros((nt, 3, nz, nx), dt) if return_wavefields else None
    st = {"force": 0, "moment": 1}[src_type]
    wav = _as(wavelet if wavelet is not None else np.zeros(nt), dt)
    rc = _fn("orc_adjoint", sfx)(C.byref(g.c()), _ptr(_as(vs, dt)), _ptr(_as(vp, dt)), _ptr(_as(rho, dt)),
                                  C.c_int(int(sx)), C.c_int(int(sz)), C.c_int(st), _ptr(wav),
                                  C.c_int(rx.size), _ptr(rx), _ptr(rz), _ptr(ra), _ptr(src_a),
                                  C.c_int(mode), _ptr(tape.get("uv")), _ptr(tape.get("e")),
                                  _ptr(tape.get("s")), _ptr(tape.get("dxu0")),
                                  _ptr(None if B_img is None else _as(B_img, dt)),
                                  _ptr(gacc), _ptr(ua), _ptr(ea))
    assert rc == 0
    out = {"src_a": src_a}
    if mode:
        out["g"] = gacc
    if return_wavefields:
        out["ua"], out["ea"] = ua, ea
    return out


def _crop(g: Geom, a):
    z0 = g.z0
    return a[..., z0:z0 + g.nz_dom, g.pad:g.pad + g.nx_dom]


def exact_chain(g: Geom, gacc, vs, vp, rho):
    """Chain the exact Lame/buoyancy gradients to (Vs, Vp, rho) on the
    physical cells (the embedding is stop_gradient outside,
    fwi/forward.py:214-219).  M = rho Vs^2, L2M = rho Vp^2, L = L2M - 2M,
    B = 1/rho (fwi/forward.py:222-227)."""
    gL2M, gL, gM, gB = (_crop(g, gacc[i]) for i in range(4))
    vs = np.asarray(vs, np.float64)
    vp = np.asarray(vp, np.float64)
    rho = np.asarray(rho, np.float64)
    g_vs = 2.0 * rho * vs * (gM - 2.0 * gL)
    g_vp = 2.0 * rho * vp * (gL2M + gL)
    g_rho = vs ** 2 * gM + vp ** 2 * gL2M + (vp ** 2 - 2.0 * vs ** 2) * gL - gB / rho ** 2
    return g_vs, g_vp, g_rho


def zhou_chain(g: Geom, gacc, vs, vp, rho):
    """fwi/imaging.py:156-169: crop, then chain with the inputs' own dtype
    (the reference multiplies 2*rho*Vs in the model's storage dtype before
    promoting, imaging.py:165-167)."""
    g_lam, g_mu, g_rl = (_crop(g, gacc[i]) for i in range(3))
    vs = np.asarray(vs)
    vp = np.asarray(vp)
    rho = np.asarray(rho)
    g_vs = 2.0 * rho * vs * (g_mu - 2.0 * g_lam)
    g_vp = 2.0 * rho * vp * g_lam
    g_rho = (vp ** 2 - 2.0 * vs ** 2) * g_lam + vs ** 2 * g_mu + g_rl
    return g_vs, g_vp, g_rho


def zhou_cubes(ex, ez, es, ua, va, b_full, dx, dz, dt, vx=None, vz=None):
    """Zhou imaging from history cubes on the padded grid, before cropping
    (fwi/imaging.py:111-152): returns (g_lam, g_mu, g_rho_lame), each
    (nz_t, nx_t), with the reference's sign.

    Spatial derivatives are centred differences with periodic wrap
    (jnp.roll, fwi/imaging.py:47-53) of the buoyancy-weighted adjoint
    velocities; the density term pairs the forward velocity of step t with
    the centred time difference of the adjoint velocity over t-1 and t+1,
    for t = 1 .. nt-2 (fwi/imaging.py:145-152)."""
    ex, ez, es, ua, va = (np.asarray(a, np.float64) for a in (ex, ez, es, ua, va))
    b = np.asarray(b_full, np.float64)
    bu, bv = b * ua, b * va

    def ddx(f):
        return (np.roll(f, -1, axis=2) - np.roll(f, 1, axis=2)) / (2.0 * dx)

    def ddz(f):
        return (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) / (2.0 * dz)

    dxbu, dzbv, dzbu, dxbv = ddx(bu), ddz(bv), ddz(bu), ddx(bv)
    g_lam = -dt * np.sum((ex + ez) * (dxbu + dzbv), axis=0)
    g_mu = -dt * np.sum(2.0 * ex * dxbu + 2.0 * ez * dzbv + es * (dzbu + dxbv), axis=0)
    g_rl = np.zeros_like(g_lam)
    if vx is not None and vz is not None:
        vx, vz = np.asarray(vx, np.float64), np.asarray(vz, np.float64)
        dua = (ua[2:] - ua[:-2]) / (2.0 * dt)
        dva = (va[2:] - va[:-2]) / (2.0 * dt)
        g_rl = -dt * np.sum(vx[1:-1] * dua + vz[1:-1] * dva, axis=0)
    return g_lam, g_mu, g_rl
