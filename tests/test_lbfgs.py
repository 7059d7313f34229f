"""The plain method's L-BFGS step (float64 device kernels) and the
multiscale driver: the two-loop direction against the dense BFGS inverse
Hessian, convergence against scipy's L-BFGS-B on convex problems, the
projections and the driver's stage bookkeeping."""
import os

import numpy as np
import pytest

from oracle import lbfgs as olb


def _pairs(n, m, seed):
    """m curvature pairs of an SPD quadratic (y = A s, so y.s > 0)."""
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.linspace(0.5, 20.0, n)) @ Q.T
    S = [rng.standard_normal(n) for _ in range(m)]
    Y = [A @ s for s in S]
    g = rng.standard_normal(n)
    gamma = float(np.dot(Y[-1], S[-1]) / np.dot(Y[-1], Y[-1]))
    return A, S, Y, g, gamma


@pytest.mark.parametrize("m", [0, 1, 5, 10])
def test_oracle_two_loop_equals_dense_bfgs(m):
    n = 60
    _, S, Y, g, gamma = _pairs(n, max(m, 1), 1)
    S, Y = S[:m], Y[:m]
    gamma = gamma if m else 1.0
    d = olb.two_loop(g, S, Y, gamma)
    H = olb.dense_inverse_hessian(S, Y, gamma, n)
    np.testing.assert_allclose(d, -H @ g, rtol=1e-11, atol=1e-12 * np.abs(H @ g).max())
    # the dense update satisfies the secant equation of the newest pair
    if m:
        np.testing.assert_allclose(H @ Y[-1], S[-1], rtol=1e-9, atol=1e-10)


def test_projections_closed_form():
    from stfwi_amd.fwi.projections import project_velocity, k_safe_ratio, cfl_vp_cap
    rng = np.random.default_rng(0)
    vp = rng.uniform(1500, 5000, (20, 30))
    vs = vp / rng.uniform(1.0, 2.0, vp.shape)
    rho = rng.uniform(-1, 2500, vp.shape)
    cap = cfl_vp_cap(10.0, 12.0, 1e-3, 0.5)
    assert cap == 0.5 * 10.0 / 1e-3
    vs2, vp2, rho2 = project_velocity(vs, vp, rho, K_margin=0.04, vp_cap=cap)
    r = k_safe_ratio(0.04)
    assert abs(r - 1.04 * 2 / np.sqrt(3)) < 1e-15
    assert np.all(vp2 <= cap) and np.all(vp2 == np.minimum(vp, cap))
    assert np.all(vs2 <= vp2 / r + 1e-12)
    assert np.all(vs2 == np.maximum(np.minimum(vs, vp2 / r), 1e-3))
    assert np.all(rho2 >= 1e-3)
    # K = rho (Vp^2 - 4/3 Vs^2) > 0 after projection
    assert np.all(vp2 ** 2 - 4.0 / 3.0 * vs2 ** 2 > 0)


def test_driver_plateau_stop_stage_parsing_and_coordinates():
    """The multistage driver's pieces on the host: the plateau stop (mean of
    the last `window` misfits not below (1 - tol) times the mean of the
    window before, fwi/optim_lbfgsb.py:117-134), stage parsing with the
    reference's errors, and the scaled coordinates with their bounds."""
    import torch
    from stfwi_amd.fwi.optim_lbfgsb import _Checkpoints, _Coordinates, _Objective, _Stalled, _StageSpec
    model = tuple(torch.full((2, 3), v, dtype=torch.float64) for v in (1000.0, 2000.0, 2300.0))
    coords = _Coordinates(model, (0,), (None, None, None))
    x = coords.vector(model)

    def objective(window, losses):
        it = iter(losses)
        return _Objective(coords, lambda m: m, lambda m: ((m[0] * 0, m[1] * 0, m[2] * 0), next(it)),
                          _Checkpoints(None, 1, None), window, 0.1, False)

    ob = objective(2, [10.0, 9.0, 5.0, 4.9, 4.9, 4.9])
    for _ in range(5):       # 4.95 < 0.9 * 9.5 and 4.9 < 0.9 * 7: keep going
        ob(x)
    with pytest.raises(_Stalled):
        ob(x)                # 4.9 >= 0.9 * 4.95
    ob = objective(None, [1.0] * 8)
    for _ in range(8):
        ob(x)
    assert ob.losses == [1.0] * 8
    assert _StageSpec.parse(dict(f_c=8.0, invert="vs_vp")).invert == "vs_vp"
    assert _StageSpec.parse(dict(f_c=8.0)).n_iters == 50
    with pytest.raises(ValueError):
        _StageSpec.parse(dict(f_c=8.0, invert="vp"))
    with pytest.raises(ValueError):
        _StageSpec.parse(dict(f_c=8.0, precon="other"))
    c2 = _Coordinates(model, (0, 1), ((500.0, 4500.0), None, None))
    x2 = c2.vector(model)
    assert x2.numel() == 12 and float(x2.abs().max()) == 1.0
    back = c2.model(x2)
    assert torch.equal(back[0], model[0]) and torch.equal(back[1], model[1]) and back[2] is model[2]
    assert torch.equal(c2.lo[:6], torch.full((6,), 0.5, dtype=torch.float64)) and torch.isinf(c2.lo[6:]).all()
    g = c2.gradient((torch.ones(2, 3), 2 * torch.ones(2, 3), torch.ones(2, 3)))
    assert torch.equal(g, torch.cat([torch.full((6,), 1000.0), torch.full((6,), 4000.0)]).double())


# ---------------------------------------------------------------- GPU ----

@pytest.mark.gpu
@pytest.mark.parametrize("m", [1, 4, 10])
def test_gpu_two_loop_matches_dense_bfgs(m):
    import torch
    from stfwi_amd.lbfgs import LBFGSMemory
    n = 300
    A, S, Y, g, gamma = _pairs(n, m, 2)
    dev = torch.device("cuda", 0)
    mem = LBFGSMemory(n, m=m, device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    gz = torch.zeros(n, dtype=torch.float64, device=dev)
    for s, y in zip(S, Y):
        assert mem.push(torch.as_tensor(s, device=dev), x, torch.as_tensor(y, device=dev), gz)
    assert abs(mem.gamma - gamma) < 1e-15 * gamma
    gt = torch.as_tensor(g, device=dev)
    d, gd = mem.direction(gt)
    d = d.cpu().numpy()
    ref = -olb.dense_inverse_hessian(S, Y, gamma, n) @ g
    assert np.linalg.norm(d - ref) / np.linalg.norm(ref) < 1e-13
    assert np.linalg.norm(d - olb.two_loop(g, S, Y, gamma)) / np.linalg.norm(ref) < 1e-14
    assert abs(gd - float(np.dot(g, d))) < 1e-12 * abs(gd)


@pytest.mark.gpu
def test_gpu_history_keeps_the_newest_pairs():
    """With more pairs than rows the oldest drop out, and a pair failing the
    curvature test leaves the stored ones untouched."""
    import torch
    from stfwi_amd.lbfgs import LBFGSMemory
    n, m = 200, 3
    A, S, Y, g, _ = _pairs(n, 6, 3)
    dev = torch.device("cuda", 0)
    mem = LBFGSMemory(n, m=m, device=dev)
    z = torch.zeros(n, dtype=torch.float64, device=dev)
    for s, y in zip(S, Y):
        mem.push(torch.as_tensor(s, device=dev), z, torch.as_tensor(y, device=dev), z)
    bad = torch.as_tensor(-S[0], device=dev)
    assert not mem.push(bad, z, torch.as_tensor(Y[0], device=dev), z)   # y.s < 0
    gamma = float(np.dot(Y[-1], S[-1]) / np.dot(Y[-1], Y[-1]))
    d, _ = mem.direction(torch.as_tensor(g, device=dev))
    ref = -olb.dense_inverse_hessian(S[-m:], Y[-m:], gamma, n) @ g
    assert np.linalg.norm(d.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-13


def _quadratic(n, seed, dev):
    import torch
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.geomspace(1.0, 100.0, n)) @ Q.T
    b = rng.standard_normal(n)
    At = torch.as_tensor(A, device=dev)
    bt = torch.as_tensor(b, device=dev)

    def fun(x):
        Ax = At @ x
        return float(0.5 * (x @ Ax) - bt @ x), Ax - bt
    return A, b, fun


@pytest.mark.gpu
def test_gpu_minimize_quadratic_reaches_solution_like_scipy():
    import torch
    from scipy.optimize import minimize as sp_min
    from stfwi_amd.lbfgs import minimize
    dev = torch.device("cuda", 0)
    n = 400
    A, b, fun = _quadratic(n, 4, dev)
    xs = np.linalg.solve(A, b)
    res = minimize(fun, torch.zeros(n, dtype=torch.float64, device=dev), maxiter=300)
    x = res["x"].cpu().numpy()
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
    sp = sp_min(lambda v: (0.5 * v @ A @ v - b @ v, A @ v - b), np.zeros(n), jac=True, method="L-BFGS-B",
                options=dict(maxiter=300, gtol=0.0, ftol=1e-30, maxcor=10))
    assert np.linalg.norm(x - sp.x) / np.linalg.norm(xs) < 1e-5
    assert res["fun"] <= sp.fun + 1e-9 * abs(sp.fun)


@pytest.mark.gpu
def test_gpu_minimize_with_bounds_matches_scipy_minimiser():
    """Convex quadratic with box constraints: the minimiser is unique, so the
    projected L-BFGS and scipy's L-BFGS-B reach the same point."""
    import torch
    from scipy.optimize import minimize as sp_min
    from stfwi_amd.lbfgs import minimize
    dev = torch.device("cuda", 0)
    n = 200
    A, b, fun = _quadratic(n, 5, dev)
    lo, hi = -0.05, 0.05
    res = minimize(fun, torch.zeros(n, dtype=torch.float64, device=dev),
                   lo=torch.full((n,), lo, dtype=torch.float64, device=dev),
                   hi=torch.full((n,), hi, dtype=torch.float64, device=dev), maxiter=2000)
    sp = sp_min(lambda v: (0.5 * v @ A @ v - b @ v, A @ v - b), np.zeros(n), jac=True, method="L-BFGS-B",
                bounds=[(lo, hi)] * n, options=dict(maxiter=2000, gtol=1e-12, ftol=1e-30))
    x = res["x"].cpu().numpy()
    assert np.all(x >= lo - 1e-15) and np.all(x <= hi + 1e-15)
    assert abs(res["fun"] - sp.fun) < 1e-6 * abs(sp.fun)
    assert np.linalg.norm(x - sp.x) / np.linalg.norm(sp.x) < 1e-3


@pytest.mark.gpu
def test_gpu_minimize_rosenbrock():
    import torch
    from stfwi_amd.lbfgs import minimize
    dev = torch.device("cuda", 0)
    n = 50

    def fun(x):
        a, bb = x[:-1], x[1:]
        f = float((100.0 * (bb - a * a) ** 2 + (1 - a) ** 2).sum())
        g = torch.zeros_like(x)
        g[:-1] += -400.0 * a * (bb - a * a) - 2.0 * (1 - a)
        g[1:] += 200.0 * (bb - a * a)
        return f, g
    res = minimize(fun, torch.full((n,), -1.0, dtype=torch.float64, device=dev), maxiter=2000)
    assert res["fun"] < 1e-10
    assert torch.allclose(res["x"], torch.ones_like(res["x"]), atol=1e-5)


@pytest.mark.gpu
def test_gpu_multiscale_driver_decreases_misfit_and_writes_checkpoints(tmp_path):
    import torch
    from stfwi_amd.fwi import forward_jax, ricker_jax
    from stfwi_amd.fwi.optim_lbfgsb import run_multiscale_fwi
    from stfwi_amd.models import anomaly_model
    nz, nx, pad, nt, dt, dx = 40, 64, 10, 300, 1e-3, 10.0
    vs_t, vp_t, rho_t = anomaly_model(nz, nx, seed=1)
    vs0 = np.full_like(vs_t, 1730.0)
    sx, sz = np.array([10, 32, 54]), np.array([2, 2, 2])
    rx = np.arange(0, nx, 2)
    rz = np.full_like(rx, 1)
    wav = ricker_jax(torch.arange(nt, dtype=torch.float64) * dt, 15.0, 1.5 / 15.0)
    obs = np.stack([forward_jax(vs_t, vp_t, rho_t, int(a), int(b), rx, rz, nx, nz, dx, dx, dt, nt, 15.0, pad, wav,
                                fd_order=4, free_surface=True)[2].cpu().numpy() for a, b in zip(sx, sz)])
    stages = [dict(f_c=15.0, n_iters=3, invert="vs", illumination_eps=1e-3, grad_smooth_sigma=1.0)]
    params, hist = run_multiscale_fwi(
        Vs_init=vs0, Vp_init=vp_t, rho_init=rho_t, src_x=sx, src_z=sz, rec_x=rx, rec_z=rz, nx_dom=nx, nz_dom=nz,
        dx=dx, dz=dx, dt=dt, nt=nt, pad=pad, observed=obs, stages=stages, component=2, backend="ad", fd_order=4,
        free_surface=True, vs_bounds=(1000.0, 2500.0), save_dir=str(tmp_path), save_every=1, verbose=False)
    lh = hist[0]["loss_history"]
    assert hist[0]["n_evals"] >= 4 and lh.min() < 0.9 * lh[0]
    assert np.sqrt(np.mean((params[0] - vs_t) ** 2)) < np.sqrt(np.mean((vs0 - vs_t) ** 2))
    assert np.all(params[1] == vp_t)
    sd = tmp_path / "stage1"
    for f in ("epoch_0000.npz", "epoch_0001.npz", "eval_0001.npz", "params.npz"):
        assert (sd / f).exists(), f
    with np.load(sd / "params.npz", allow_pickle=True) as z:
        assert {"Vs", "Vp", "rho", "loss_history", "f_c", "cutoff", "invert", "backend"} <= set(z.files)
    with np.load(sd / "epoch_0001.npz") as z:
        assert {"Vs", "Vp", "rho", "iter", "loss"} <= set(z.files)


@pytest.mark.gpu
def test_gpu_failed_learned_proposals_fall_back_to_plain_steps():
    """A proposed direction that is not a descent direction, or that finds no
    decrease within its evaluation budget, is rejected: the iterates are then
    exactly the plain method's."""
    import torch
    from stfwi_amd.lbfgs import minimize
    dev = torch.device("cuda", 0)
    n = 120
    A, b, fun = _quadratic(n, 6, dev)
    x0 = torch.zeros(n, dtype=torch.float64, device=dev)
    plain = minimize(fun, x0, maxiter=15)
    ascent = minimize(fun, x0, maxiter=15, direction_fn=lambda x, g: g.clone())
    assert torch.equal(ascent["x"], plain["x"]) and ascent["n_fallback"] == 15 and ascent["n_learned"] == 0
    assert ascent["nfev"] == plain["nfev"]
    # a barely-descent direction (mostly orthogonal to g) with a one-evaluation
    # budget: at the L-BFGS step length its trial point increases f, so every
    # proposal falls back
    v = torch.as_tensor(np.random.default_rng(9).standard_normal(n), device=dev)

    def skew(x, g):
        orth = v - (v @ g) / (g @ g) * g
        return -1e-6 * g + orth * (g.norm() / orth.norm())
    bad = minimize(fun, x0, maxiter=15, maxls_learned=1, direction_fn=skew)
    assert torch.equal(bad["x"], plain["x"]) and bad["n_fallback"] == 15
    assert bad["nfev"] == plain["nfev"] + 15
    # a good proposal (the exact Newton step) is accepted and converges at once
    Ainv = torch.as_tensor(np.linalg.inv(A), device=dev)
    newton = minimize(fun, x0, maxiter=3, direction_fn=lambda x, g: -(Ainv @ g))
    xs = np.linalg.solve(A, b)
    assert newton["n_learned"] >= 1
    assert np.linalg.norm(newton["x"].cpu().numpy() - xs) / np.linalg.norm(xs) < 1e-8


@pytest.mark.gpu
def test_gpu_learned_driver_collects_pairs_and_keeps_descending(tmp_path):
    """Plain stage collecting training pairs, a short training, then a stage
    with the learned direction: the misfit never increases across accepted
    steps and every iteration is either a learned step or a fallback."""
    import torch
    from stfwi_amd.fwi import forward_jax, ricker_jax
    from stfwi_amd.fwi.optim_lbfgsb import run_multiscale_fwi
    from stfwi_amd.learned import LearnedPreconditioner, PairCollector
    from stfwi_amd.models import anomaly_model
    nz, nx, pad, nt, dt, dx = 40, 64, 10, 300, 1e-3, 10.0
    vs_t, vp_t, rho_t = anomaly_model(nz, nx, seed=1)
    vs0 = np.full_like(vs_t, 1730.0)
    sx, sz = np.array([10, 32, 54]), np.array([2, 2, 2])
    rx = np.arange(0, nx, 2)
    rz = np.full_like(rx, 1)
    wav = ricker_jax(torch.arange(nt, dtype=torch.float64) * dt, 15.0, 1.5 / 15.0)
    obs = np.stack([forward_jax(vs_t, vp_t, rho_t, int(a), int(b), rx, rz, nx, nz, dx, dx, dt, nt, 15.0, pad, wav,
                                fd_order=4, free_surface=True)[2].cpu().numpy() for a, b in zip(sx, sz)])
    common = dict(Vp_init=vp_t, rho_init=rho_t, src_x=sx, src_z=sz, rec_x=rx, rec_z=rz, nx_dom=nx, nz_dom=nz, dx=dx,
                  dz=dx, dt=dt, nt=nt, pad=pad, observed=obs, component=2, backend="ad", fd_order=4,
                  free_surface=True, verbose=False)
    stage = [dict(f_c=15.0, n_iters=4, invert="vs", illumination_eps=1e-3)]
    lp = LearnedPreconditioner(nz, nx, seed=0)
    col = PairCollector(lp)
    params, hist = run_multiscale_fwi(Vs_init=vs0, stages=stage, collect=col, **common)
    assert len(col) == 4 and len(col.pairs["ahead"]) == 4
    lp.train(epochs=60, lr=3e-3, batch=2, pairs=col.pairs["step"])
    params2, hist2 = run_multiscale_fwi(Vs_init=params[0], stages=stage, learned=lp, **common)
    h = hist2[0]
    assert h["n_learned"] + h["n_fallback"] == 4
    assert h["final_loss"] < hist[0]["final_loss"]


def test_unknown_learned_policy_is_refused():
    """run_multiscale_fwi checks learned.policy before any device work."""
    from stfwi_amd.fwi.optim_lbfgsb import run_multiscale_fwi

    class Bad:
        policy = "sometimes"
    with pytest.raises(ValueError, match="policy"):
        run_multiscale_fwi(Vs_init=None, Vp_init=None, rho_init=None, src_x=None, src_z=None, rec_x=None, rec_z=None,
                           nx_dom=4, nz_dom=4, dx=1.0, dz=1.0, dt=1e-3, nt=4, pad=2, observed=None,
                           stages=[dict(f_c=8.0)], learned=Bad())


class _StubProposer:
    """A learned-direction stand-in: steepest descent (accepted) or ascent
    (rejected without an evaluation), proposed per `policy`."""
    absolute = False

    def __init__(self, policy, sign=-1.0):
        self.policy, self.maxls, self.sign, self.calls = policy, 2, sign, 0

    def direction(self, grad, illum, vs):
        self.calls += 1
        return self.sign * grad


@pytest.mark.gpu
@pytest.mark.parametrize("policy,sign,want", [
    ("every", -1.0, (3, 3)), ("stage", -1.0, (1, 1)), ("run", -1.0, (1, 0)),
    ("until_fail", 1.0, (1, 1)), ("until_fail", -1.0, None)])
def test_gpu_learned_proposal_policies(policy, sign, want):
    """When the driver asks for a learned update: at every iteration, at the
    first iteration of each stage, once per run, or at every iteration
    until one proposal of the stage is rejected.  want = proposals in the
    first and second stage (3 iterations each); for 'until_fail' with a
    descent proposal, the proposals of a stage are its accepted learned
    steps plus at most the one rejection that ends them."""
    import torch
    from stfwi_amd.fwi import forward_jax, ricker_jax
    from stfwi_amd.fwi.optim_lbfgsb import run_multiscale_fwi
    from stfwi_amd.models import anomaly_model
    nz, nx, pad, nt, dt, dx = 40, 64, 10, 300, 1e-3, 10.0
    vs_t, vp_t, rho_t = anomaly_model(nz, nx, seed=1)
    sx, sz = np.array([10, 32, 54]), np.array([2, 2, 2])
    rx = np.arange(0, nx, 2)
    rz = np.full_like(rx, 1)
    wav = ricker_jax(torch.arange(nt, dtype=torch.float64) * dt, 15.0, 1.5 / 15.0)
    obs = np.stack([forward_jax(vs_t, vp_t, rho_t, int(a), int(b), rx, rz, nx, nz, dx, dx, dt, nt, 15.0, pad, wav,
                                fd_order=4, free_surface=True)[2].cpu().numpy() for a, b in zip(sx, sz)])
    stages = [dict(f_c=15.0, n_iters=3, invert="vs", illumination_eps=1e-3)] * 2
    prop = _StubProposer(policy, sign)
    _, hist = run_multiscale_fwi(Vs_init=np.full_like(vs_t, 1730.0), Vp_init=vp_t, rho_init=rho_t, src_x=sx,
                                 src_z=sz, rec_x=rx, rec_z=rz, nx_dom=nx, nz_dom=nz, dx=dx, dz=dx, dt=dt, nt=nt,
                                 pad=pad, observed=obs, stages=stages, component=2, backend="ad", fd_order=4,
                                 free_surface=True, verbose=False, learned=prop)
    got = tuple(h["n_learned"] + h["n_fallback"] for h in hist)
    if want is None:
        assert all(h["n_fallback"] <= 1 and 1 <= h["n_learned"] + h["n_fallback"] <= 3 for h in hist), hist
        assert all(h["n_fallback"] == 1 or h["n_learned"] == 3 for h in hist)
        want = got
    assert got == want, (got, want)
    assert prop.calls == sum(want)
    if sign > 0:
        assert all(h["n_learned"] == 0 for h in hist)
    assert all(h["n_iters_done"] == 3 for h in hist)


def test_tikhonov_term_gradient_by_central_differences():
    """The Tikhonov term of build_loss_fn (fwi/loss.py:123-128): the sum of
    squared first differences along z and x of the inverted parameters.
    Its gradient against central differences, exact for a quadratic; the
    frozen parameters get none."""
    import torch
    from stfwi_amd.fwi.loss import _tikhonov
    rng = np.random.default_rng(3)
    params = [torch.as_tensor(rng.standard_normal((5, 7)), dtype=torch.float64) for _ in range(3)]
    for invert, active in (("vs", 1), ("vs_vp", 2), ("all", 3)):
        reg, grads = _tikhonov(params, invert)
        ref = sum(float((np.diff(p.numpy(), axis=0) ** 2).sum() + (np.diff(p.numpy(), axis=1) ** 2).sum())
                  for p in params[:active])
        assert abs(reg / ref - 1) < 1e-14
        for i, p in enumerate(params):
            if i >= active:
                assert not torch.any(grads[i])
                continue
            for (r, c) in ((0, 0), (2, 3), (4, 6), (1, 6)):
                h = 1e-3
                pp = [q.clone() for q in params]
                pp[i][r, c] += h
                fp = _tikhonov(pp, invert)[0]
                pp[i][r, c] -= 2 * h
                fm = _tikhonov(pp, invert)[0]
                assert abs((fp - fm) / (2 * h) - float(grads[i][r, c])) < 1e-8
