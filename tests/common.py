"""Shared test inputs: the reference's dot-product model
(test_adjoint_kernel.py:33-58), fixture loaders and tolerances."""
import hashlib
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# relative-L2 bars
TOL_F64 = 1e-12
TOL_F32 = 1e-5
TOL_DOT = 1e-14


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dp_model():
    """24x32 smooth heterogeneous model of test_adjoint_kernel.py:36-46."""
    nz, nx = 24, 32
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    Vp = 2500.0 + 150.0 * np.sin(2 * np.pi * zz / nz) * np.cos(2 * np.pi * xx / nx)
    Vs = 1400.0 + 80.0 * np.cos(2 * np.pi * zz / nz) * np.sin(2 * np.pi * xx / nx)
    rho = 2000.0 + 100.0 * np.sin(np.pi * zz / nz) * np.cos(np.pi * xx / nx)
    return dict(nz=nz, nx=nx, dx=10.0, dz=10.0, nt=60, dt=1e-3, fc=15.0, pad=8, Vs=Vs, Vp=Vp, rho=rho)


def dp_vectors(nrec, nt=60):
    """Seeded wavelet and receiver adjoints (test_adjoint_kernel.py:33,54-58)."""
    rng = np.random.default_rng(2026)
    s = rng.standard_normal(nt)
    y = np.stack([rng.standard_normal((nt, nrec)) for _ in range(4)])
    return s, y


def dp_geometry(nz, nx, free_surface, src_type):
    """Source/receiver placement of test_adjoint_kernel.py:50,64-74."""
    src_x = nx // 3
    rec_x = np.arange(2, nx - 2, 4, dtype=np.int32)
    if src_type == "moment":
        src_z = nz // 4
    elif free_surface:
        src_z = 0
    else:
        src_z = nz // 4
    rec_z = np.zeros_like(rec_x) if free_surface else np.full_like(rec_x, nz // 5)
    return src_x, src_z, rec_x, rec_z


DP_CONFIGS = [(fd, fs, st) for fd in (2, 4, 8) for fs in (False, True) for st in ("force", "moment")]


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def marmousi_init():
    """True models and the drivers' initial models (gaussian_filter, sigma 20,
    on slowness for velocities; compute_gradient_handcoded.py:69-76),
    verified against the SHA-256 stored with the fixtures."""
    from scipy.ndimage import gaussian_filter
    t = load_golden("marmousi_true.npz")
    vs_t, vp_t, rho_t = t["vs"], t["vp"], t["rho"]
    vs_i = 1.0 / gaussian_filter(1.0 / vs_t, 20)
    vp_i = 1.0 / gaussian_filter(1.0 / vp_t, 20)
    rho_i = gaussian_filter(rho_t, 20)
    g = load_golden("gradient_compare.npz")
    for a, k in ((vs_i, "sha_Vs_init"), (vp_i, "sha_Vp_init"), (rho_i, "sha_rho_init")):
        assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == str(g[k]), k
    return dict(vs_true=vs_t, vp_true=vp_t, rho_true=rho_t, vs_init=vs_i, vp_init=vp_i, rho_init=rho_i)


def zhou_constructed(nt=9, nz=6, nx=8, dx=10.0, dz=20.0, dt=1e-3):
    """History cubes whose Zhou images are known in closed form.

    Strains a, c, e are constant in time.  The adjoint velocities are
    periodic in space, U_a = sin(kx j) and V_a = cos(kz i), so the wrapped
    centred differences are exact: sin(kx) cos(kx j) / dx and
    -sin(kz) sin(kz i) / dz.  For the density term the adjoint velocities
    are t^2 h(z, x), t^2 h2(z, x) and the forward velocities t f, t f2, so
    vx(t) (U_a(t+1) - U_a(t-1)) / (2 dt) = 2 t^2 f h / dt.  The buoyancy b is
    constant."""
    rng = np.random.default_rng(11)
    t = np.arange(nt, dtype=np.float64)[:, None, None]
    i = np.arange(nz)[None, :, None]
    j = np.arange(nx)[None, None, :]
    kx, kz = 2 * np.pi / nx, 2 * np.pi / nz
    a, c, e = (rng.uniform(0.5, 1.5, (nz, nx)) for _ in range(3))
    f, f2, h, h2 = (rng.standard_normal((nz, nx)) for _ in range(4))
    shape = (nt, nz, nx)
    cube = lambda x: np.broadcast_to(x, shape).copy()
    return dict(nt=nt, dx=dx, dz=dz, dt=dt, b=4e-4, a=a, c=c, e=e, f=f, f2=f2, h=h, h2=h2,
                ex=cube(a), ez=cube(c), es=cube(e), zero=np.zeros(shape),
                sin_x=cube(np.sin(kx * j)), cos_z=cube(np.cos(kz * i)),
                d_sin_x=np.sin(kx) * np.cos(kx * j[0]) / dx,      # d/dx sin(kx j), (1, nx)
                d_cos_z=-np.sin(kz) * np.sin(kz * i[0]) / dz,     # d/dz cos(kz i), (nz, 1)
                ua_t=t ** 2 * h, va_t=t ** 2 * h2, vx_t=t * f, vz_t=t * f2)
