"""The HIP path on the workload configurations beyond the 25-shot benchmark:

* the 64x128 FP64 anomaly survey (one moment source, 32 receivers, nt=1000,
  fd4, C-PML on every side) against the oracle, and its dot-product test
  (test_adjoint_kernel.py:61-124 at this size);
* 64 uneven shots in one batched launch per fd order (record lengths
  1500-4000 steps, 120-500 receivers, force and moment sources), each
  shot's history sized by its own record: sampled shots run alone give the
  same seismograms and source adjoints bit for bit;
* the 608x2000, nt=8000, 8-shot survey through SeismogramOperator with
  forward checkpoints every 90 steps: the gradient against central
  differences of the misfit (h = 1e-3 relative) along 16 random Vs
  directions.

Measured figures are printed (pytest -s shows them)."""
import time

import numpy as np
import pytest
import torch

import oracle
from oracle import signal as osig
from tests.common import TOL_DOT, TOL_F64, rel_l2

pytestmark = pytest.mark.gpu


def _plan(s, dtype, n_shots, nt_max, nrec_max, imaging="exact", grad="vs", hist=True, fd=None, **kw):
    from stfwi_amd import Plan
    return Plan(s["nz"], s["nx"], s["dx"], s["dz"], s["dt"], s["fc"], s["pad"],
                fd_order=fd if fd is not None else s["fd_order"], free_surface=s["free_surface"], dtype=dtype,
                n_shots_max=n_shots, nt_max=nt_max, nrec_max=nrec_max, imaging=imaging, grad=grad,
                store_history=hist, device=torch.device("cuda", 0), **kw)


@pytest.mark.parametrize("imaging", ["exact", "zhou"])
def test_anomaly_survey_fp64_matches_oracle(imaging):
    """Observed data on the anomaly model, then misfit and (Vs, Vp, rho)
    gradient at the background model, HIP against the oracle."""
    from stfwi_amd.surveys import anomaly_survey
    s = anomaly_survey()
    f64 = torch.float64
    nrec, nt = s["rec_x"].size, s["nt"]
    p = _plan(s, f64, 1, nt, nrec, imaging=imaging, grad="all" if imaging == "exact" else "vs_vp")
    p.set_model(*s["true"])
    p.set_shots(s["src_x"], s["src_z"], "moment", nt, s["rec_x"], s["rec_z"], s["wavelet"])
    obs = p.forward()["rec"]
    g = oracle.Geom(s["nz"], s["nx"], s["pad"], False, 4, nt, s["dx"], s["dz"], s["dt"], s["fc"])
    obs_o, _ = oracle.forward(g, *s["true"], 64, 3, s["wavelet"], s["rec_x"], s["rec_z"], "moment", np.float64)
    assert rel_l2(obs[0].cpu().numpy(), obs_o) < TOL_F64
    scale = float(1.0 / (np.mean(obs_o[2] ** 2) + 1e-45))
    def gradient(model):
        p.set_model(*model)
        rec = p.forward()["rec"]
        loss, radj = p.residual(rec, obs[:, 2], 2, None, 0, scale, 1)
        p.adjoint(radj, want_src=False)
        return loss, p.finalize(*model).cpu().numpy()

    loss, gg = gradient(s["init"])
    # conditioning: the misfit at the background is the small scattered
    # field of the anomaly, so the gradient amplifies rounding; `sens` is its
    # change for a 1e-15 relative perturbation of the model
    rng = np.random.default_rng(1)
    pert = tuple(a * (1 + 1e-15 * rng.standard_normal(a.shape)) for a in s["init"])
    gp = gradient(pert)[1]
    tape = ("e", "dxu0", "s") if imaging == "exact" else ("e", "uv")
    rec_o, t = oracle.forward(g, *s["init"], 64, 3, s["wavelet"], s["rec_x"], s["rec_z"], "moment", np.float64, tape)
    l_o, cot = osig.residual_prep(rec_o[2], obs_o[2], None, 0, scale, 1)
    ra = np.zeros_like(rec_o)
    ra[2] = cot
    a = oracle.adjoint(g, *s["init"], 64, 3, ra, s["rec_x"], s["rec_z"], "moment", s["wavelet"], imaging, t)
    chain = oracle.exact_chain if imaging == "exact" else oracle.zhou_chain
    go = chain(g, a["g"], *s["init"])
    assert abs(float(loss[0]) / l_o - 1) < TOL_F64
    n_par = 3 if imaging == "exact" else 2
    errs = [rel_l2(gg[i], go[i]) for i in range(n_par)]
    sens = [rel_l2(gp[i], gg[i]) for i in range(n_par)]
    print(f"\nanomaly survey {imaging}: loss {abs(float(loss[0]) / l_o - 1):.1e}, gradients "
          + ", ".join(f"{e:.1e}" for e in errs) + " (sensitivity " + ", ".join(f"{e:.1e}" for e in sens) + ")")
    for e, d in zip(errs, sens):
        assert e < max(TOL_F64, 10 * d)


def test_anomaly_survey_dot_product():
    """<L s, y> = <s, L^T y> over all four receiver channels, FP64."""
    from stfwi_amd.surveys import anomaly_survey
    s = anomaly_survey()
    nrec, nt = s["rec_x"].size, s["nt"]
    rng = np.random.default_rng(3)
    w = rng.standard_normal(nt)
    y = rng.standard_normal((1, 4, nt, nrec))
    p = _plan(s, torch.float64, 1, nt, nrec, imaging="none", hist=False)
    p.set_model(*s["true"])
    p.set_shots(s["src_x"], s["src_z"], "moment", nt, s["rec_x"], s["rec_z"], w)
    rec = p.forward()["rec"].cpu().numpy()
    sa = p.adjoint(torch.as_tensor(y))["src_adj"].cpu().numpy()
    lhs, rhs = float(np.sum(rec * y)), float(np.sum(w * sa[0]))
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"\nanomaly survey dot product: {err:.1e}")
    assert err < TOL_DOT


def _batched_group(s, fd):
    idx = [k for k in range(len(s["src_x"])) if s["fd_order"][k] == fd]
    return idx, int(sum(s["nt"][k] for k in idx))


def _set_shots(p, s, idx):
    p.set_shots([s["src_x"][k] for k in idx], [s["src_z"][k] for k in idx], [s["src_type"][k] for k in idx],
                [int(s["nt"][k]) for k in idx], [s["rec_x"][k] for k in idx], [s["rec_z"][k] for k in idx],
                s["wavelet"])


@pytest.mark.parametrize("fd", [4, 8])
def test_batched_uneven_shots_equal_single_shots(fd):
    """One batched FP32 launch over the survey's shots of one fd order, in 3
    stream groups: four sampled shots (the longest, the shortest, a force
    and a moment shot) run alone give bit-identical seismograms and source
    adjoints; the Vs gradient of the batch equals the sum of the single-shot
    gradients to FP32 summation order; the history holds only the shots'
    own frames."""
    from stfwi_amd.surveys import batched_survey
    s = batched_survey()
    idx, frames = _batched_group(s, fd)
    f32 = torch.float32
    nt_max, nrec_max = int(max(s["nt"][k] for k in idx)), 500
    p = _plan(s, f32, len(idx), nt_max, nrec_max, fd=fd, hist_frames=frames)
    full_hist = len(idx) * nt_max * 3 * s["nz"] * s["nx"] * 4
    assert p.nbytes < full_hist            # per-shot frames, not shots x nt_max
    p.set_groups(3)
    p.set_model(*s["init"])
    _set_shots(p, s, idx)
    t0 = time.perf_counter()
    rec = p.forward()["rec"]
    rng = np.random.default_rng(fd)
    radj = torch.as_tensor(rng.standard_normal(tuple(rec.shape)), dtype=f32, device=rec.device) * 1e-3
    for i, k in enumerate(idx):     # receiver adjoints only inside each shot's record and spread
        radj[i, :, int(s["nt"][k]):] = 0
        radj[i, :, :, len(s["rec_x"][k]):] = 0
    sa = p.adjoint(radj)["src_adj"]
    gb = p.finalize(*s["init"])[0]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nts = [int(s["nt"][k]) for k in idx]
    pick = sorted({int(np.argmax(nts)), int(np.argmin(nts)), [s["src_type"][k] for k in idx].index("force"),
                   [s["src_type"][k] for k in idx].index("moment")})
    gsum = torch.zeros_like(gb, dtype=torch.float64)
    q = _plan(s, f32, 1, nt_max, nrec_max, fd=fd)
    q.set_model(*s["init"])
    for i in range(len(idx)):
        _set_shots(q, s, [idx[i]])
        r1 = q.forward()["rec"]
        sa1 = q.adjoint(radj[i:i + 1, :, :, :q.nrec_max].contiguous())["src_adj"]
        gsum += q.finalize(*s["init"])[0].double()
        if i in pick:
            nr = len(s["rec_x"][idx[i]])
            assert torch.equal(r1[0, :, :, :nr], rec[i, :, :, :nr]), i
            assert torch.equal(sa1[0], sa[i]), i
    err = rel_l2(gb.double().cpu().numpy(), gsum.cpu().numpy())
    print(f"\nbatched fd{fd}: {len(idx)} shots, {frames} history frames, plan {p.nbytes / 1e9:.1f} GB "
          f"(shots x nt_max would hold {full_hist / 1e9:.1f} GB of history), forward+adjoint {el * 1e3:.0f} ms, "
          f"batched vs summed single-shot gradient {err:.1e}")
    assert err < 1e-5


def test_large_survey_checkpointed_gradient_against_central_differences():
    """608x2000, nt=8000, 8 shots, 2000 receivers, FP32, forward checkpoints
    every 90 steps: the exact gradient of the misfit with respect to Vs
    against central differences at h = 1e-3 relative along 16 random
    directions (fwi/forward.py:471-501 checkpoints the same recursion)."""
    from stfwi_amd.autograd import SeismogramOperator
    from stfwi_amd.plan import choose_ckpt
    from stfwi_amd.surveys import large_survey
    from stfwi_amd import _abi
    s = large_survey()
    dev = torch.device("cuda", 0)
    f32 = torch.float32
    op = SeismogramOperator(s["nz"], s["nx"], s["dx"], s["dz"], s["dt"], s["fc"], s["pad"], s["src_x"], s["src_z"],
                            s["wavelet"], s["rec_x"], s["rec_z"], s["nt"], src_type="moment", fd_order=8,
                            free_surface=True, dtype=f32, channels=(2,), ckpt_every=90, groups=2, device=dev)
    vs_t, vp_t, rho_t = (torch.as_tensor(a, dtype=f32, device=dev) for a in s["true"])
    with torch.no_grad():
        obs = op(vs_t, vp_t, rho_t)
    scale = float(1.0 / (obs.double().pow(2).mean() + 1e-45))
    vs0, vp0, rho0 = (torch.as_tensor(a, dtype=f32, device=dev) for a in s["init"])

    def misfit(vs):
        r = op(vs, vp0, rho0) - obs
        return 0.5 * scale * (r.double() ** 2).sum()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vs = vs0.clone().requires_grad_(True)
    loss = misfit(vs)
    loss.backward()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    g = vs.grad.double()
    peak = torch.cuda.max_memory_allocated(dev)
    plan_bytes = op._plans[1].nbytes
    cfg = _abi.Config(s["nz"], s["nx"], s["pad"], 1, 8, s["dx"], s["dz"], s["dt"], s["fc"], _abi.F32, 8, s["nt"],
                      2000, 2, 1, 1, 0, 0)
    from stfwi_amd.plan import footprint
    full_bytes = footprint(cfg)
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(7)
    errs = []
    t0 = time.perf_counter()
    with torch.no_grad():
        for _ in range(16):
            # random directions smoothed over 10 cells (50 m, below the
            # 15 Hz wavelengths), so that the FP32 misfit difference is
            # well above rounding
            d = gaussian_filter(rng.standard_normal((s["nz"], s["nx"])), 10.0)
            d = torch.as_tensor(d / np.abs(d).max(), dtype=torch.float64, device=dev)
            d[:, 0] = 0
            d[:, -1] = 0
            d[-1, :] = 0                 # cells the embedding copies into the C-PML (held fixed)
            dv = 1e-3 * vs0.double() * d
            fd = (misfit((vs0.double() + dv).float()) - misfit((vs0.double() - dv).float())) / 2.0
            an = float((g * dv).sum())
            errs.append(abs(float(fd) - an) / abs(an))
    t_fd = time.perf_counter() - t0
    print(f"\nlarge survey: gradient {t_grad:.1f} s, plan {plan_bytes / 1e9:.1f} GB with checkpoints every 90 "
          f"steps (full history {full_bytes / 1e9:.0f} GB; planner picks K={choose_ckpt(cfg, plan_bytes * 2)} "
          f"for twice the plan), torch peak {peak / 1e9:.1f} GB, 32 misfits {t_fd:.1f} s, "
          f"central-difference errors median {np.median(errs):.1e} max {max(errs):.1e}")
    assert full_bytes > 288e9
    assert np.median(errs) < 1e-2 and max(errs) < 5e-2
