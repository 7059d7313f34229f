"""The drop-in modules (stfwi_amd.fwi, mirroring the reference's fwi package)
on an MI355X: against the reference's committed golden outputs and against
the oracle."""
import numpy as np
import pytest
import torch

import oracle
from oracle import signal as osig
from oracle import precond as opre
from tests.common import (TOL_F64, dp_geometry, dp_model, dp_vectors, load_golden, marmousi_init, rel_l2,
                          zhou_constructed)

pytestmark = pytest.mark.gpu


@pytest.fixture
def x64():
    from stfwi_amd import config
    config.enable_x64(True)
    yield
    config.enable_x64(False)


def _np(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def test_compute_gradient_matches_reference_golden_outputs(x64):
    """Results_gradient/gradient_compare.npz: FP64 Marmousi, 5 shots, nt=400,
    fd8, pad 30, free surface.  auto_scale pins the true-model forward, the
    losses the misfit, g_Vs_HC the Zhou gradient and g_Vs_AD (float32 storage)
    the exact gradient."""
    from stfwi_amd.fwi import build_forward_fn, compute_gradient, ricker_jax
    fx = load_golden("gradient_compare.npz")
    mm = marmousi_init()
    nt, dt, fc, pad = int(fx["nt"]), float(fx["dt"]), float(fx["f_c"]), int(fx["pad"])
    nz, nx = int(fx["nz"]), int(fx["nx"])
    wav = ricker_jax(torch.arange(nt, dtype=torch.float64) * dt, fc, 1.5 / fc)
    run = build_forward_fn(nz, nx, 20.0, 20.0, dt, nt, fc, pad, 200, fx["rec_x"], fx["rec_z"], fd_order=8,
                           free_surface=True)
    obs = run.batched(mm["vs_true"], mm["vp_true"], mm["rho_true"], fx["src_x"], fx["src_z"], wav)[2]
    scale = float(1.0 / (torch.mean(obs ** 2) + 1e-45))
    assert abs(scale / float(fx["auto_scale"]) - 1) < 1e-11
    common = dict(Vs=mm["vs_init"], Vp=mm["vp_init"], rho=mm["rho_init"], src_x=fx["src_x"], src_z=fx["src_z"],
                  src_wavelet=wav, observed=obs, rec_x=fx["rec_x"], rec_z=fx["rec_z"], nx_dom=nx, nz_dom=nz,
                  dx=20.0, dz=20.0, dt=dt, nt=nt, fc=fc, pad=pad, component=2, invert="vs", scale=scale,
                  fd_order=8, free_surface=True, src_type="force", block_size=200)
    (g_hc, g_vp, g_rho), l_hc = compute_gradient("adjoint", **common)
    assert g_hc.dtype == np.float64
    assert abs(l_hc / float(fx["loss_HC"]) - 1) < 1e-11
    assert rel_l2(g_hc, fx["g_Vs_HC"]) < 1e-9
    assert not np.any(g_vp) and not np.any(g_rho)
    (g_ad, _, _), l_ad = compute_gradient("ad", **common)
    assert g_ad.dtype == np.float32
    assert abs(l_ad / float(fx["loss_AD"]) - 1) < 1e-11
    assert rel_l2(g_ad, fx["g_Vs_AD"]) < 1e-7


def test_forward_and_adjoint_dropins_match_oracle(x64):
    from stfwi_amd.fwi import forward_jax, adjoint_jax, build_forward_fn
    m = dp_model()
    for fd, fs, st in ((8, True, "moment"), (4, False, "force")):
        sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
        s, y = dp_vectors(rx.size, m["nt"])
        args = (m["Vs"], m["Vp"], m["rho"], sx, sz, rx, rz, m["nx"], m["nz"], m["dx"], m["dz"], m["dt"], m["nt"],
                m["fc"], m["pad"])
        rec = forward_jax(*args, s, fd_order=fd, free_surface=fs, src_type=st)
        g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
        rec_o, tape = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, s, rx, rz, st, np.float64, ("uv", "e"))
        for i in range(4):
            assert rel_l2(_np(rec[i]), rec_o[i]) < TOL_F64
        full = forward_jax(*args, s, return_vars=["ex", "vz_full", "es_full"], fd_order=fd, free_surface=fs,
                           src_type=st)
        assert rel_l2(_np(full[0]), rec_o[2]) < TOL_F64
        assert rel_l2(_np(full[1]), tape["uv"][:, 1]) < TOL_F64
        assert rel_l2(_np(full[2]), tape["e"][:, 2]) < TOL_F64
        out = adjoint_jax(*args, *y, fd_order=fd, free_surface=fs, src_type=st, return_wavefields="velocity")
        a_o = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, y, rx, rz, st, return_wavefields=True)
        assert rel_l2(_np(out[0]), a_o["src_a"]) < TOL_F64
        assert rel_l2(_np(out[1]), a_o["ua"][:, 0]) < TOL_F64
        assert rel_l2(_np(out[2]), a_o["ua"][:, 1]) < TOL_F64
        lhs = sum(float(np.sum(_np(rec[i]) * y[i])) for i in range(4))
        rhs = float(np.sum(s * _np(out[0])))
        assert abs(lhs - rhs) / abs(lhs) < 1e-14
    # build_forward_fn models a force source and batches shots
    run = build_forward_fn(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["nt"], m["fc"], m["pad"], 100, rx, rz,
                           fd_order=4, free_surface=False)
    one = run(m["Vs"], m["Vp"], m["rho"], 5, 6, s)
    many = run.batched(m["Vs"], m["Vp"], m["rho"], np.array([5, 9]), np.array([6, 6]), s)
    assert torch.equal(many[2][0], one[2])
    with pytest.raises(NotImplementedError):
        adjoint_jax(*args, *y, fd_order=6)


def test_zhou_imaging_from_cubes_matches_oracle(x64):
    from stfwi_amd.fwi import imaging_condition
    m = dp_model()
    fs, st, fd = True, "moment", 8
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
    s, y = dp_vectors(rx.size, m["nt"])
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    _, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, s, rx, rz, st, np.float64, ("uv", "e"))
    a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, y, rx, rz, st, s, "zhou", t, return_wavefields=True)
    ref = oracle.zhou_chain(g, a["g"], m["Vs"], m["Vp"], m["rho"])
    out = imaging_condition(m["Vs"], m["Vp"], m["rho"], t["e"][:, 0], t["e"][:, 1], t["e"][:, 2], a["ua"][:, 0],
                            a["ua"][:, 1], m["dx"], m["dz"], m["dt"], m["pad"], fs, m["nz"], m["nx"],
                            vx_hist=t["uv"][:, 0], vz_hist=t["uv"][:, 1])
    for i in range(3):
        assert rel_l2(_np(out[i]), ref[i]) < TOL_F64


def test_postprocessing_kernels_match_oracle(x64):
    from stfwi_amd.fwi import gaussian_smooth_2d, jax_sosfilt, cosine_taper_end, sosfilt_T, eprecond3
    from stfwi_amd.fwi.preconditioning import _call as eprec_call
    from stfwi_amd.fwi.illumination import compute_illumination_for_stage, shot_energy
    from scipy.signal import butter
    rng = np.random.default_rng(5)
    f = rng.standard_normal((40, 57))
    for sigma in (1.0, 2.0):
        assert rel_l2(_np(gaussian_smooth_2d(f, sigma)), opre.gaussian_smooth_2d(f, sigma)) < 1e-13
    x = rng.standard_normal((900, 7))
    sos = butter(6, 5.0, fs=1.0 / 1.5e-3, output="sos")
    assert rel_l2(_np(jax_sosfilt(x, sos)), osig.sosfilt(x, sos)) < 1e-12
    assert rel_l2(_np(cosine_taper_end(x, 40)), osig.cosine_taper_end(x, 40)) < 1e-15
    assert rel_l2(_np(sosfilt_T(x, sos, 40)), osig.preprocess_T(x, sos, 40)) < 1e-12
    m = dp_model()
    rx = np.arange(0, m["nx"], 2)
    rz = np.zeros_like(rx)
    s = osig.ricker(np.arange(m["nt"]) * m["dt"], m["fc"], 1.5 / m["fc"])
    sx, sz = np.array([4, 15, 27]), np.array([0, 2, 5])
    H = compute_illumination_for_stage(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["nt"], m["fc"], m["pad"], 100,
                                       rx, rz, 8, 2, m["Vs"], m["Vp"], m["rho"], sx, sz, s, eps=1e-3)
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], True, 8, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    Ws = sum((oracle.forward(g, m["Vs"], m["Vp"], m["rho"], a, b, s, rx, rz, "force", np.float64, ("e",))[1]["e"][:, 0] ** 2).sum(0)
             for a, b in zip(sx, sz))
    assert rel_l2(_np(H), opre.illumination_from_energy(Ws, m["nz"], m["nx"], m["pad"], 1e-3)) < 1e-12
    ws = shot_energy(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["nt"], m["fc"], m["pad"], rx, rz, 8, 2, m["Vs"],
                     m["Vp"], m["rho"], sx, sz, s)
    assert rel_l2(_np(ws.sum(0)), Ws) < 1e-12
    We = eprecond3(ws.sum(0), rx, rz, m["dx"], m["dz"], True, m["pad"], eps=5e-7)
    assert rel_l2(_np(We), opre.eprecond3(Ws, rx, rz, m["dx"], m["dz"], True, m["pad"], 5e-7)) < 1e-12
    _, mult = eprec_call(ws, rx, rz, m["dx"], m["dz"], True, m["pad"], m["nz"], m["nx"], 5e-7)
    ref = opre.eprecond3_multiplier(Ws, rx, rz, m["dx"], m["dz"], True, m["pad"], m["nz"], m["nx"], 5e-7)
    assert rel_l2(_np(mult), ref) < 1e-12


def test_default_precision_is_float32():
    from stfwi_amd.fwi import forward_jax
    m = dp_model()
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], True, "force")
    s, _ = dp_vectors(rx.size, m["nt"])
    rec = forward_jax(m["Vs"], m["Vp"], m["rho"], sx, sz, rx, rz, m["nx"], m["nz"], m["dx"], m["dz"], m["dt"],
                      m["nt"], m["fc"], m["pad"], s, fd_order=8)
    assert rec[0].dtype == torch.float32 and rec[0].shape == (m["nt"], rx.size)


def test_eprecond1_taper_grad_and_source_taper_match_oracle():
    """eprecond1 (fwi/preconditioning.py:37-62) in both precisions, the depth
    taper (:144-183) with and without the bottom roll-off, and the source
    taper (fwi/taper.py:16-84) for sources inside, at the edge of and outside
    the model, against the oracle restatements."""
    from stfwi_amd.fwi import eprecond1, taper_grad, build_source_taper
    rng = np.random.default_rng(8)
    ws = rng.random((44, 61)) ** 2 + 1e-6
    wr = rng.random((44, 61)) ** 2 + 1e-6
    we = eprecond1(ws, wr, eps=5e-7)
    assert we.dtype == torch.float64
    assert rel_l2(_np(we), opre.eprecond1(ws, wr, 5e-7)) < 1e-15
    we32 = eprecond1(ws.astype(np.float32), wr.astype(np.float32), eps=5e-7)
    assert we32.dtype == torch.float32
    ref32 = opre.eprecond1(ws.astype(np.float32), wr.astype(np.float32), 5e-7)
    assert ref32.dtype == np.float32 and rel_l2(_np(we32), ref32) < 1e-7
    for kw in (dict(gradt1=4, gradt2=8, exp_taper=2.0), dict(gradt1=0, gradt2=3, exp_taper=1.5, gradb1=30, gradb2=40)):
        t = taper_grad((44, 61), 7.5, **kw)
        assert rel_l2(_np(t), opre.taper_grad((44, 61), 7.5, **kw)) < 1e-15
    for srcs, r, f in (([(30, 2), (5, 40)], 45.0, 2), ([(0, 0), (60, 43), (80, 10)], 30.0, 1), ([], 20.0, 2)):
        t = build_source_taper(44, 61, 10.0, srcs, r, filt_size=f)
        ref = opre.source_taper(44, 61, 10.0, srcs, r, filt_size=f)
        assert np.max(np.abs(_np(t) - ref)) < 1e-14, srcs


def _tiny_inversion(tmp_path):
    """Observed vz from a Vs/Vp anomaly on the 24x32 model, four force shots,
    and the drop-in loss_fn (build_loss_fn) on the smooth starting model."""
    from stfwi_amd.fwi import build_forward_fn, build_loss_fn, ricker_jax
    m = dp_model()
    nt = 160
    rx = np.arange(2, m["nx"] - 2, 2)
    rz = np.zeros_like(rx)
    run = build_forward_fn(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], nt, m["fc"], m["pad"], nt, rx, rz,
                           fd_order=4, free_surface=True)
    wav = _np(ricker_jax(np.arange(nt) * m["dt"], m["fc"], 1.5 / m["fc"])) * 1e6
    sx = np.array([4, 12, 20, 28])
    zz, xx = np.meshgrid(np.arange(m["nz"]), np.arange(m["nx"]), indexing="ij")
    bump = np.exp(-((zz - 12) ** 2 + (xx - 16) ** 2) / 18.0)
    vs_t, vp_t = m["Vs"] * (1 + 0.05 * bump), m["Vp"] * (1 + 0.03 * bump)
    obs = np.stack([_np(run(vs_t, vp_t, m["rho"], int(s), 2, wav)[1]) for s in sx])
    scale = 1.0 / np.mean(obs ** 2)
    return m, sx, wav, obs, scale, run


def test_run_lbfgs_dropins_decrease_misfit_and_save_models(tmp_path, x64):
    """run_lbfgs (fwi/optimizer.py:20-168) over Vs only and over all three
    parameters, and run_lbfgs_alternating (:171-284): the misfit decreases,
    frozen parameters stay put, and the inverted models and loss history are
    written as the reference names them."""
    from stfwi_amd.fwi import build_loss_fn, run_lbfgs, run_lbfgs_alternating
    m, sx, wav, obs, scale, run = _tiny_inversion(tmp_path)
    init = (m["Vs"].copy(), m["Vp"].copy(), m["rho"].copy())
    lf = build_loss_fn(run, obs, sx, np.full(sx.size, 2), wav, component=1, scale=scale, invert="vs")
    p, hist = run_lbfgs(lf, init, 4, save_dir=str(tmp_path / "a"), invert="vs")
    assert len(hist) == 4 and hist[-1] < hist[0]
    assert not np.array_equal(p[0], init[0]) and np.array_equal(p[1], init[1]) and np.array_equal(p[2], init[2])
    for n in ("Vs_inverted.npy", "Vp_inverted.npy", "rho_inverted.npy", "loss_history.npy"):
        assert (tmp_path / "a" / n).exists()
    assert np.array_equal(np.load(tmp_path / "a" / "Vs_inverted.npy"), p[0])
    lf2 = build_loss_fn(run, obs, sx, np.full(sx.size, 2), wav, component=1, scale=scale, invert="vs_vp")
    p2, hist2 = run_lbfgs_alternating(lf2, init, 3, save_dir=str(tmp_path / "b"), invert="vs_vp")
    assert hist2[-1] < hist2[0]
    assert not np.array_equal(p2[0], init[0]) and not np.array_equal(p2[1], init[1])
    assert np.array_equal(p2[2], init[2])
    p3, hist3 = run_lbfgs(lf2, init, 2, save_dir=str(tmp_path / "c"), invert="vs_vp", early_stop_ratio=2.0)
    assert len(hist3) == 1        # ratio 1 < 2 stops after the first epoch


def test_zhou_imaging_closed_forms(x64):
    """imaging_condition (HIP, from cubes) on the constructed cubes whose
    Zhou images are known in closed form: periodic centred differences of
    B U_a and B V_a for g_lambda and g_mu, and the density term's sum over
    t = 1 .. nt-2 (fwi/imaging.py:125-152).  The grid is the padded grid of
    a 4 x 4 model with pad 2 and a free surface; rho is constant (b = 1/rho)."""
    from stfwi_amd.fwi import imaging_condition
    k = zhou_constructed()
    nt, b, dt = k["nt"], k["b"], k["dt"]
    pad, nzd, nxd = 2, 4, 4
    rho = np.full((nzd, nxd), 1.0 / b)
    vs, vp = np.full((nzd, nxd), 1000.0), np.full((nzd, nxd), 2000.0)
    z = k["zero"]
    crop = (slice(0, nzd), slice(pad, pad + nxd))

    def run(ex, ez, es, ua, va, vx=None, vz=None):
        out = imaging_condition(vs, vp, rho, ex, ez, es, ua, va, k["dx"], k["dz"], dt, pad, True, nzd, nxd,
                                vx_hist=vx, vz_hist=vz)
        return [_np(x) for x in out[3:]]

    g_lam, g_mu, g_rl = run(k["ex"], k["ez"], z, k["sin_x"], k["cos_z"])
    div = b * (k["d_sin_x"] + k["d_cos_z"])
    assert rel_l2(g_lam, (-dt * nt * (k["a"] + k["c"]) * div)[crop]) < 1e-13
    assert rel_l2(g_mu, (-dt * nt * 2 * b * (k["a"] * k["d_sin_x"] + k["c"] * k["d_cos_z"]))[crop]) < 1e-13
    _, g_mu, _ = run(z, z, k["es"], k["cos_z"], k["sin_x"])
    assert rel_l2(g_mu, (-dt * nt * k["e"] * b * (k["d_cos_z"] + k["d_sin_x"]))[crop]) < 1e-13
    g_lam, g_mu, g_rl = run(z, z, z, k["ua_t"], k["va_t"], k["vx_t"], k["vz_t"])
    assert not np.any(g_lam) and not np.any(g_mu)
    closed = -2.0 * float(np.sum(np.arange(1, nt - 1) ** 2)) * (k["f"] * k["h"] + k["f2"] * k["h2"])
    assert rel_l2(g_rl, closed[crop]) < 1e-13


def test_build_loss_fn_tikhonov_composition(x64):
    """fwi/loss.py:123-130: total = misfit + alpha * stop_grad(misfit / reg) * reg.
    The value is exactly (1 + alpha) times the misfit, and the gradient adds
    alpha * (misfit / reg) times the regulariser's gradient, the ratio held
    fixed."""
    from stfwi_amd.fwi import build_forward_fn, build_loss_fn
    from stfwi_amd.fwi.loss import _tikhonov
    m = dp_model()
    fs, fd = True, 8
    rx = np.arange(2, m["nx"] - 2, 3)
    rz = np.zeros_like(rx)
    s, _ = dp_vectors(rx.size, m["nt"])
    run = build_forward_fn(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["nt"], m["fc"], m["pad"], 60, rx, rz,
                           fd_order=fd, free_surface=fs)
    sx, sz = np.array([6, 20]), np.array([0, 0])
    obs = run.batched(m["Vs"], m["Vp"], m["rho"], sx, sz, s)[2]
    params = (m["Vs"] * 1.02, m["Vp"], m["rho"])
    alpha = 0.01
    for invert in ("vs", "all"):
        plain = build_loss_fn(run, obs, sx, sz, s, component=2, scale=1e15, invert=invert)
        tik = build_loss_fn(run, obs, sx, sz, s, component=2, scale=1e15, invert=invert, tikhonov_alpha=alpha)
        l0, g0 = plain(params)
        l1, g1 = tik(params)
        assert abs(l1 / ((1 + alpha) * l0) - 1) < 1e-14
        reg, rg = _tikhonov([torch.as_tensor(p, dtype=torch.float64).cuda() for p in params], invert)
        w = alpha * l0 / (reg + 1e-30)
        for i in range(3):
            expect = _np(g0[i]) + w * _np(rg[i])
            assert rel_l2(_np(g1[i]), expect) < 1e-13 or not np.any(expect), (invert, i)
