"""Pins for the CPU oracle (oracle/): the reference's own golden outputs, its
dot-product test, complex-step differentiation, library routines and closed
forms.  Each check is chosen so that a dropped term, wrong sign, wrong index
or transposed operand in the restatement fails it."""
import numpy as np
import pytest

import oracle
from oracle import signal as osig
from oracle import precond as opre
from tests.common import (DP_CONFIGS, TOL_DOT, dp_geometry, dp_model, dp_vectors, load_golden,
                          marmousi_init, rel_l2, zhou_constructed)


@pytest.mark.parametrize("fd,fs,st", DP_CONFIGS)
def test_dot_product(fd, fs, st):
    """test_adjoint_kernel.py:61-124, tightened from 1e-10 to 1e-14."""
    m = dp_model()
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
    s, y = dp_vectors(rx.size, m["nt"])
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    rec, _ = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, s, rx, rz, st)
    a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, y, rx, rz, st)
    lhs, rhs = float(np.sum(rec * y)), float(np.sum(s * a["src_a"]))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < TOL_DOT


def _cs_case(fd, fs, st):
    rng = np.random.default_rng(100 * fd + 10 * fs + (st == "moment"))
    nz, nx, pad, nt = 14, 18, 5, 70
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    Vp = 2500 + 150 * np.sin(2 * np.pi * zz / nz) * np.cos(2 * np.pi * xx / nx)
    Vs = 1400 + 80 * np.cos(2 * np.pi * zz / nz) * np.sin(2 * np.pi * xx / nx)
    rho = 2000 + 100 * np.sin(np.pi * zz / nz) * np.cos(np.pi * xx / nx)
    g = oracle.Geom(nz, nx, pad, fs, fd, nt, 10.0, 10.0, 1e-3, 25.0)
    wav = osig.ricker(np.arange(nt) * 1e-3, 25.0, 1.5 / 25.0) * 1e6
    rx = np.arange(1, nx - 1, 3)
    rz = np.zeros_like(rx) if fs else np.full_like(rx, 2)
    sx, sz = nx // 2, (0 if (fs and st == "force") else 3)
    obs = rng.standard_normal((nt, rx.size)) * 1e-9
    sos = np.array([[0.1, 0.2, 0.1, 1.0, -1.2, 0.5], [1.0, 0.0, -1.0, 1.0, -0.3, 0.2]])
    return rng, g, Vs, Vp, rho, wav, rx, rz, sx, sz, obs, sos


@pytest.mark.parametrize("fd,fs,st", DP_CONFIGS)
def test_complex_step_exact_gradient(fd, fs, st):
    """The exact-mode gradient (what jax.grad returns for backend='ad') equals
    the complex-step derivative of the oracle's own misfit in random
    directions of Vs, Vp and rho, with filter + taper in the loss."""
    rng, g, Vs, Vp, rho, wav, rx, rz, sx, sz, obs, sos = _cs_case(fd, fs, st)
    comp, scale, ntap = 2, 1e15, 10
    rec, t = oracle.forward(g, Vs, Vp, rho, sx, sz, wav, rx, rz, st, np.float64, ("e", "dxu0", "s"))
    _, cot = osig.residual_prep(rec[comp], obs, sos, ntap, scale, 1)
    ra = np.zeros_like(rec)
    ra[comp] = cot
    a = oracle.adjoint(g, Vs, Vp, rho, sx, sz, ra, rx, rz, st, wav, "exact", t)
    grads = oracle.exact_chain(g, a["g"], Vs, Vp, rho)
    h = 1e-20
    for which in range(3):
        d = rng.standard_normal(Vs.shape) * 10.0
        mc = [Vs.astype(complex), Vp.astype(complex), rho.astype(complex)]
        mc[which] = mc[which] + 1j * h * d
        recc, _ = oracle.forward(g, *mc, sx, sz, wav, rx, rz, st, np.complex128)
        est = osig.preprocess(recc[comp], sos, ntap)
        ob = osig.preprocess(obs.astype(complex), sos, ntap)
        dd = (scale * np.mean((est - ob) ** 2)).imag / h
        an = float(np.sum(grads[which] * d))
        assert abs(dd - an) / abs(dd) < 1e-12, (which, dd, an)


def test_wavelet_linearity():
    m = dp_model()
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], True, "moment")
    s, _ = dp_vectors(rx.size, m["nt"])
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], True, 8, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    r1, _ = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, s, rx, rz, "moment")
    r2, _ = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, 2.5 * s, rx, rz, "moment")
    assert rel_l2(r2, 2.5 * r1) < 1e-14


def test_pml_profiles_invariants():
    """fwi/forward.py:35-104: interior cells are an exact no-op (k=1, b=1,
    a=0), the outermost cell keeps alpha=0 but d>0, profiles are mirrored."""
    nzd, nxd, pad = 10, 12, 6
    for fs in (False, True):
        g = oracle.Geom(nzd, nxd, pad, fs, 4, 10, 10.0, 10.0, 1e-3, 10.0)
        vp = np.full((nzd, nxd), 3000.0)
        s = oracle.setup(g, vp * 0.5, vp, vp * 0 + 2000.0)
        nz, nx = g.shape
        assert np.all(s["kx"][pad:nx - pad] == 1) and np.all(s["bx"][pad:nx - pad] == 1)
        assert np.all(s["ax"][pad:nx - pad] == 0)
        np.testing.assert_array_equal(s["kx"][:pad], s["kx"][nx - pad:][::-1])
        # outermost: alpha = 0 so b = exp(-d/k dt) with d = d0, k = 2
        d0 = -3 * 3000.0 * np.log(1e-3) / (2 * pad * 10.0)
        assert np.isclose(s["bx"][0], np.exp(-d0 / 2.0 * 1e-3), rtol=1e-15)
        assert s["kx"][0] == 2.0
        top_rows = s["kz"][:pad]
        assert np.all(top_rows == 1) if fs else np.all(top_rows > 1)
        assert np.all(s["kz"][nz - pad:] > 1)
        # moduli
        np.testing.assert_allclose(s["M"], 2000.0 * 1500.0 ** 2)
        np.testing.assert_allclose(s["L"], 2000.0 * 3000.0 ** 2 - 2 * 2000.0 * 1500.0 ** 2)


def test_filter_and_taper_transposes():
    rng = np.random.default_rng(3)
    sos = np.array([[0.1, 0.2, 0.1, 1.0, -1.2, 0.5], [1.0, 0.0, -1.0, 2.0, -0.3, 0.2]])
    x = rng.standard_normal((200, 5))
    y = rng.standard_normal((200, 5))
    for n_taper in (0, 17):
        lhs = np.sum(osig.preprocess(x, sos, n_taper) * y)
        rhs = np.sum(x * osig.preprocess_T(y, sos, n_taper))
        assert abs(lhs - rhs) / abs(lhs) < 1e-13


def test_taper_window_closed_form():
    w = osig.taper_window(100, 8)
    assert w[-1] == 0.0
    assert np.isclose(w[0], (np.cos(np.pi / 8) + 1) / 2)
    assert np.all(np.diff(w) < 0)
    assert osig.taper_window(5, 0) is None
    assert osig.taper_window(5, 9).size == 5


def test_sosfilt_matches_scipy():
    from scipy.signal import butter, sosfilt
    dt = 1.5e-3
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3000, 4))
    for cutoff in (8.0, 5.0, 3.0):
        sos = butter(6, cutoff, fs=1.0 / dt, output="sos")
        ours = osig.sosfilt(x, sos)
        ref = sosfilt(sos, x, axis=0)
        assert rel_l2(ours, ref) < 1e-8, cutoff


def test_residual_prep_gradient_fd():
    rng = np.random.default_rng(4)
    pred = rng.standard_normal((300, 3))
    obs = rng.standard_normal((300, 3))
    sos = np.array([[0.2, 0.4, 0.2, 1.0, -0.8, 0.3]])
    l0, cot = osig.residual_prep(pred, obs, sos, 30, 3.0, 2)
    d = rng.standard_normal(pred.shape)
    h = 1e-6
    lp, _ = osig.residual_prep(pred + h * d, obs, sos, 30, 3.0, 2)
    lm, _ = osig.residual_prep(pred - h * d, obs, sos, 30, 3.0, 2)
    assert abs((lp - lm) / (2 * h) - np.sum(cot * d)) < 1e-7 * abs(np.sum(cot * d))


def test_gaussian_smoothing_matches_scipy():
    from scipy.signal import convolve2d
    rng = np.random.default_rng(1)
    f = rng.standard_normal((30, 41))
    for sigma in (1.0, 2.5):
        k = opre.gaussian_kernel_2d(sigma)
        assert k.shape == (2 * int(3 * sigma + 0.5) + 1,) * 2
        ref = convolve2d(f, k, mode="same")
        assert rel_l2(opre.gaussian_smooth_2d(f, sigma), ref) < 1e-14


def test_eprecond3_line_integral_closed_form():
    """Below the centre of a horizontal receiver line of half-length a at
    perpendicular distance y, the weight is sqrt(Ws) * 2 asinh(a / y)."""
    pad, dx, dz = 4, 10.0, 10.0
    nz_t, nx_t = 30, 48
    Ws = np.ones((nz_t, nx_t))
    rx = np.arange(0, 41, 1)
    rz = np.zeros_like(rx)
    We = opre.eprecond3(Ws, rx, rz, dx, dz, True, pad, eps=0.0)
    a = 20 * dx
    for depth in (3, 11, 25):
        assert np.isclose(We[depth, pad + 20], 2 * np.arcsinh(a / (depth * dz)), rtol=1e-13)
    # row 0 is clamped to one grid spacing
    assert np.isclose(We[0, pad + 20], 2 * np.arcsinh(a / dz), rtol=1e-13)
    # vertical line: same geometry transposed
    We2 = opre.eprecond3(Ws.T.copy(), rz + 5, rx, dx, dz, False, pad, eps=0.0)
    assert np.isclose(We2[pad + 20, pad + 5 + 7], 2 * np.arcsinh(a / (7 * dx)), rtol=1e-13)


def test_illumination_normalisation():
    rng = np.random.default_rng(2)
    Ws = rng.random((20, 30)) + 0.1
    H = opre.illumination_from_energy(Ws, 12, 18, 6, eps=0.01)
    crop = Ws[:12, 6:24]
    ref = 1.0 / (crop + 0.01 * crop.max())
    np.testing.assert_allclose(H, ref / ref.max(), rtol=1e-15)
    assert H.max() == 1.0


@pytest.mark.slow
def test_oracle_matches_reference_gradient_compare():
    """Results_gradient/gradient_compare.npz: FP64, Marmousi, 5 shots,
    nt=400, fd8, pad 30, free surface, force, exx.  Pins the true-model
    forward (auto_scale), forward+misfit (loss), the Zhou gradient (g_Vs_HC)
    and the exact gradient (g_Vs_AD, stored as float32)."""
    from oracle.gradient import compute_gradient
    fx = load_golden("gradient_compare.npz")
    mm = marmousi_init()
    nt, dt, fc, pad = int(fx["nt"]), float(fx["dt"]), float(fx["f_c"]), int(fx["pad"])
    nz, nx = int(fx["nz"]), int(fx["nx"])
    wav = osig.ricker(np.arange(nt) * dt, fc, 1.5 / fc)
    g = oracle.Geom(nz, nx, pad, True, int(fx["fd_order"]), nt, 20.0, 20.0, dt, fc)
    obs = np.array([oracle.forward(g, mm["vs_true"], mm["vp_true"], mm["rho_true"], sx, sz, wav,
                                   fx["rec_x"], fx["rec_z"])[0][2]
                    for sx, sz in zip(fx["src_x"], fx["src_z"])])
    scale = float(1.0 / (np.mean(obs ** 2) + 1e-45))
    assert abs(scale / float(fx["auto_scale"]) - 1) < 1e-11
    common = dict(Vs=mm["vs_init"], Vp=mm["vp_init"], rho=mm["rho_init"], src_x=fx["src_x"],
                  src_z=fx["src_z"], src_wavelet=wav, observed=obs, rec_x=fx["rec_x"], rec_z=fx["rec_z"],
                  nx_dom=nx, nz_dom=nz, dx=20.0, dz=20.0, dt=dt, nt=nt, fc=fc, pad=pad, component=2,
                  invert="vs", scale=scale, fd_order=8, free_surface=True, src_type="force")
    (g_hc, g_vp, g_rho), l_hc = compute_gradient("adjoint", **common)
    assert abs(l_hc / float(fx["loss_HC"]) - 1) < 1e-11
    assert rel_l2(g_hc, fx["g_Vs_HC"]) < 1e-9
    assert not np.any(g_vp) and not np.any(g_rho)
    (g_ad, _, _), l_ad = compute_gradient("ad", **common)
    assert g_ad.dtype == np.float32
    assert abs(l_ad / float(fx["loss_AD"]) - 1) < 1e-11
    # float32 storage: agreement to the storage rounding
    assert rel_l2(g_ad, fx["g_Vs_AD"]) < 1e-7
    # the Zhou and exact gradients differ (the reference's 0.73)
    assert 0.6 < rel_l2(g_hc, fx["g_Vs_AD"]) < 0.85


def test_oracle_eprecond1_closed_forms_and_reference_checks():
    """eprecond1 (fwi/preconditioning.py:37-62): equal energies reduce to
    Ws + eps max Ws; the reference's own checks (test_phase1_modules.py:61-68):
    shape kept, all entries positive."""
    from oracle import precond as op
    rng = np.random.default_rng(3)
    ws = rng.random((30, 44)) ** 2 + 1e-6
    wr = rng.random((30, 44)) ** 2 + 1e-6
    we = op.eprecond1(ws, wr, eps=5e-7)
    assert we.shape == ws.shape and np.all(we > 0)
    same = op.eprecond1(ws, ws, eps=1e-3)
    assert np.allclose(same, ws + 1e-3 * ws.max(), rtol=1e-15, atol=0)
    i, j = 7, 19
    assert we[i, j] == np.sqrt(ws[i, j] * wr[i, j]) + 5e-7 * np.sqrt(ws * wr).max()


def test_oracle_taper_grad_values():
    """taper_grad rows: 0 above gradt1, exp(-(gradt2-j)^2) up to gradt2 (1 at
    gradt2), (j dz)^p below; the reference's checks
    (test_phase1_modules.py:79-82); the bottom roll-off ends at 0."""
    from oracle import precond as op
    t = op.taper_grad((20, 7), dz=5.0, gradt1=4, gradt2=8, exp_taper=2.0)
    assert np.all(t[:4] == 0.0) and np.all(t[8] > t[7])
    assert np.all(t[8] == 1.0) and np.allclose(t[6], np.exp(-4.0), rtol=1e-15)
    assert np.allclose(t[12], (12 * 5.0) ** 2, rtol=1e-15)
    assert np.all(t == t[:, :1])
    b = op.taper_grad((20, 3), dz=5.0, gradt1=2, gradt2=3, exp_taper=1.0, gradb1=10, gradb2=14)
    assert np.all(b[15:] == 0.0) and np.allclose(b[10], 50.0) and np.allclose(b[12], 60.0 * np.exp(-1.0))


def test_oracle_source_taper_against_scipy_erf_and_geometry():
    """The erf patch against scipy.special.erf; the taper is 0 in the box at a
    source, 1 far from every source, mirror-symmetric about an isolated
    source and non-decreasing along a ray from it."""
    from scipy.special import erf
    from oracle import precond as op
    dh, R = 10.0, 45.0
    p = op.source_patch(dh, R)
    ps = p.shape[0]
    c = ps / 2.0
    jj, ii = np.mgrid[0:ps, 0:ps]
    q = erf(np.hypot((ii - c + 0.5) * dh, (jj - c + 0.5) * dh) / (np.sqrt(2.0) * R))
    q -= q.min()
    m = ps // 2
    q = np.clip(q / max(q[0, m], q[m, 0], q[m, -1], q[-1, m]), 0, 1)
    assert np.allclose(p, q, rtol=0, atol=1e-15)
    t = op.source_taper(40, 60, dh, [(30, 20)], R, filt_size=2)
    assert np.all(t[18:23, 28:33] == 0.0) and t[0, 0] == 1.0 and t.min() == 0.0 and t.max() == 1.0
    assert np.allclose(t[20, 30 - 5:30], t[20, 31:36][::-1]) or ps % 2 == 0
    row = t[20, 30:]
    assert np.all(np.diff(row) >= -1e-15)
    two = op.source_taper(40, 60, dh, [(10, 5), (50, 30)], R, filt_size=1)
    assert two[5, 10] == 0.0 and two[30, 50] == 0.0 and np.all((two >= 0) & (two <= 1))


def test_zhou_cubes_closed_forms():
    """oracle.zhou_cubes on constructed cubes against closed forms: the
    periodic centred differences of the buoyancy-weighted adjoint velocities
    (g_lambda, g_mu, the shear strain paired with dz(B U_a) + dx(B V_a)), and
    the density term's sum over t = 1 .. nt-2 of vx(t) times the centred
    time difference of U_a.  A sum over all t, a pairing shifted by one step
    or the opposite sign each miss the closed form by far more than
    rounding."""
    k = zhou_constructed()
    nt, b, dt = k["nt"], k["b"], k["dt"]
    z = k["zero"]
    # normal strains: g_lambda and g_mu
    g_lam, g_mu, g_rl = oracle.zhou_cubes(k["ex"], k["ez"], z, k["sin_x"], k["cos_z"], b, k["dx"], k["dz"], dt)
    div = b * (k["d_sin_x"] + k["d_cos_z"])
    assert rel_l2(g_lam, -dt * nt * (k["a"] + k["c"]) * div) < 1e-13
    assert rel_l2(g_mu, -dt * nt * 2 * b * (k["a"] * k["d_sin_x"] + k["c"] * k["d_cos_z"])) < 1e-13
    assert not np.any(g_rl)
    # shear strain alone: es (dz(B U_a) + dx(B V_a)), U_a periodic in z and
    # V_a periodic in x
    g_lam, g_mu, _ = oracle.zhou_cubes(z, z, k["es"], k["cos_z"], k["sin_x"], b, k["dx"], k["dz"], dt)
    assert not np.any(g_lam)
    assert rel_l2(g_mu, -dt * nt * k["e"] * b * (k["d_cos_z"] + k["d_sin_x"])) < 1e-13
    # density term: strains zero, so only g_rho_lame
    g_lam, g_mu, g_rl = oracle.zhou_cubes(z, z, z, k["ua_t"], k["va_t"], b, k["dx"], k["dz"], dt,
                                          vx=k["vx_t"], vz=k["vz_t"])
    assert not np.any(g_lam) and not np.any(g_mu)
    fh = k["f"] * k["h"] + k["f2"] * k["h2"]
    closed = -2.0 * float(np.sum(np.arange(1, nt - 1) ** 2)) * fh
    assert rel_l2(g_rl, closed) < 1e-13
    # the closed form separates the plausible mistakes
    all_t = -2.0 * float(np.sum(np.arange(0, nt) ** 2)) * fh
    shifted = -2.0 * float(np.sum(np.arange(1, nt - 1) * np.arange(2, nt))) * fh   # vx(t+1) with dU_a(t)
    for wrong in (all_t, shifted, -closed):
        assert rel_l2(g_rl, wrong) > 1e-2


def test_zhou_fused_oracle_matches_cube_restatement():
    """The C oracle fuses Zhou imaging into its adjoint (a three-frame ring of
    carry-outs for the density term).  On its own forward and adjoint
    history cubes it must equal zhou_cubes, the cube form of
    fwi/imaging.py:111-152, for all three Lame images."""
    m = dp_model()
    for fd, fs, st in ((8, True, "moment"), (4, False, "force")):
        sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
        s, y = dp_vectors(rx.size, m["nt"])
        g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
        _, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, s, rx, rz, st, np.float64, ("uv", "e"))
        a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, y, rx, rz, st, s, "zhou", t,
                           return_wavefields=True)
        nz_t, nx_t = g.shape
        rr = np.clip(np.arange(nz_t) - g.z0, 0, m["nz"] - 1)
        cc = np.clip(np.arange(nx_t) - g.pad, 0, m["nx"] - 1)
        b_full = 1.0 / m["rho"][rr][:, cc]
        ref = oracle.zhou_cubes(t["e"][:, 0], t["e"][:, 1], t["e"][:, 2], a["ua"][:, 0], a["ua"][:, 1], b_full,
                                m["dx"], m["dz"], m["dt"], vx=t["uv"][:, 0], vz=t["uv"][:, 1])
        for i in range(3):
            assert rel_l2(a["g"][i], ref[i]) < 1e-13, (fd, fs, st, i)
