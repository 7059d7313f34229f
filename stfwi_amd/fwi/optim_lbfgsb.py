"""Drop-in for fwi/optim_lbfgsb.py: the multistage plain-method FWI loop.

    run_multiscale_fwi(*, Vs_init, Vp_init, rho_init, src_x, src_z, rec_x,
        rec_z, nx_dom, nz_dom, dx, dz, dt, nt, pad, observed, stages,
        component=2, backend='ad', fd_order=8, free_surface=True,
        src_type='force', block_size=200, auto_scale=None, K_margin=0.04,
        cfl_coeff=None, vs_bounds=None, vp_bounds=None, rho_bounds=None,
        plateau_window=None, plateau_rel_tol=1e-3, save_dir=None,
        save_every=None, verbose=True)              fwi/optim_lbfgsb.py:137-473
    -> ((Vs, Vp, rho) float64 numpy, list of per-stage history dicts)

Same stage semantics as the reference: per-stage Ricker wavelet
(t0 = 1.5/f_c), 6th-order Butterworth low-pass, cosine end taper, optional
illumination ('illum' H_inv or 'eprecond3') computed once at the start of the
stage, per-block scaling x / max|x|, velocity projection before every
evaluation, per-cell bounds in scaled space, plateau stop, and the npz
checkpoints (stageN/epoch_NNNN.npz, eval_NNNN.npz, params.npz) with the
reference's key names.

The optimiser is this package's float64 device L-BFGS (stfwi_amd.lbfgs:
history 10, 20 line-search evaluations, gtol/ftol disabled as in the
reference), not scipy's L-BFGS-B; bounds are handled by projecting the
trial points.  Every gradient is one batched GPU evaluation of all shots.

Two keyword arguments beyond the reference's (Vs-only stages):
  learned  a stfwi_amd.learned.LearnedPreconditioner: its direction from the
           raw Vs gradient and the stage's illumination is tried before every
           L-BFGS step, which is taken instead whenever the learned direction
           does not decrease the misfit (history: n_learned, n_fallback);
  collect  a LearnedPreconditioner that receives (gradient, illumination,
           accepted update) training pairs from the run.
"""
from __future__ import annotations

import os
import time
from typing import Optional, Sequence

import numpy as np
import torch

from ..lbfgs import minimize
from ._common import device
from .forward import ricker_jax
from .gradient import compute_gradient
from .illumination import compute_illumination_for_stage
from .preconditioning import eprecond3_preconditioner
from .projections import cfl_vp_cap, project_velocity

_KEY_TO_IDX = {"Vs": 0, "Vp": 1, "rho": 2}


def _block_scales(params):
    return [max(float(np.abs(p).max()), 1e-12) for p in params]


def _active_keys(invert):
    keys = {"vs": ["Vs"], "vs_vp": ["Vs", "Vp"], "all": ["Vs", "Vp", "rho"]}
    if invert not in keys:
        raise ValueError(f"invert={invert!r}; use 'vs', 'vs_vp', or 'all'")
    return keys[invert]


class _Plateau:
    """Stop when the mean of the last `window` losses is not below
    (1 - rel_tol) times the mean of the `window` before (disabled for
    window None or 0) -- fwi/optim_lbfgsb.py:117-134."""

    def __init__(self, window: Optional[int], rel_tol: float):
        self.window = int(window or 0)
        self.rel_tol = float(rel_tol)
        self.history: list = []

    def push(self, loss: float) -> bool:
        self.history.append(float(loss))
        w = self.window
        if w <= 0 or len(self.history) < 2 * w:
            return False
        last = float(np.mean(self.history[-w:]))
        prev = float(np.mean(self.history[-2 * w:-w]))
        return last >= prev * (1.0 - self.rel_tol)


class _PlateauStop(Exception):
    pass


def run_multiscale_fwi(*, Vs_init, Vp_init, rho_init, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt,
                       pad, observed, stages: Sequence[dict], component: int = 2, backend: str = "ad",
                       fd_order: int = 8, free_surface: bool = True, src_type: str = "force", block_size: int = 200,
                       auto_scale: Optional[float] = None, K_margin: float = 0.04, cfl_coeff: Optional[float] = None,
                       vs_bounds=None, vp_bounds=None, rho_bounds=None, plateau_window: Optional[int] = None,
                       plateau_rel_tol: float = 1e-3, save_dir: Optional[str] = None,
                       save_every: Optional[int] = None, verbose: bool = True, learned=None, collect=None):
    if not stages:
        raise ValueError("`stages` must be a non-empty list")
    from scipy.signal import butter

    observed = np.asarray(observed)
    if auto_scale is None:
        auto_scale = float(1.0 / (np.mean(observed.astype(np.float64) ** 2) + 1e-45))
    if verbose:
        print(f"auto_scale = {auto_scale:.4e}")
    vp_cap = cfl_vp_cap(dx, dz, dt, cfl_coeff) if cfl_coeff is not None else None
    if vp_bounds is not None and vp_bounds[1] is not None:
        vp_cap = vp_bounds[1] if vp_cap is None else min(vp_cap, vp_bounds[1])

    params = tuple(np.array(p, dtype=np.float64, copy=True) for p in (Vs_init, Vp_init, rho_init))
    history = []
    dev = device()
    n_cells = int(nz_dom) * int(nx_dom)

    for istage, stage in enumerate(stages):
        f_c = float(stage["f_c"])
        cutoff = stage.get("cutoff", None)
        n_iters = int(stage.get("n_iters", 50))
        invert = stage.get("invert", "vs")
        n_taper = int(float(stage.get("n_taper_pct", 0.0)) * nt)
        illumination_eps = stage.get("illumination_eps", None)
        grad_smooth_sigma = stage.get("grad_smooth_sigma", None)
        source_taper = stage.get("source_taper", None)
        precon = stage.get("precon", "illum")
        if verbose:
            print(f"Stage {istage + 1}/{len(stages)}  f_c={f_c} Hz  cutoff={cutoff}  invert={invert}  "
                  f"n_iters={n_iters}  backend={backend}")

        wav = ricker_jax(torch.arange(nt, dtype=torch.float64) * dt, f_c, 1.5 / f_c)
        sos = butter(6, cutoff, fs=1.0 / dt, output="sos") if cutoff is not None else None

        illumination = None
        if illumination_eps is not None and precon != "none":
            if precon == "illum":
                illumination = compute_illumination_for_stage(
                    nz_dom, nx_dom, dx, dz, dt, nt, f_c, pad, block_size, rec_x, rec_z, fd_order, component,
                    params[0], params[1], params[2], src_x, src_z, wav, eps=illumination_eps)
            elif precon == "eprecond3":
                illumination = eprecond3_preconditioner(
                    nz_dom=nz_dom, nx_dom=nx_dom, dx=dx, dz=dz, dt=dt, nt=nt, f_c=f_c, pad=pad, rec_x=rec_x,
                    rec_z=rec_z, fd_order=fd_order, component=component, Vs=params[0], Vp=params[1],
                    rho=params[2], src_x=src_x, src_z=src_z, src_wavelet=wav, free_surface=free_surface,
                    eps=illumination_eps)
            else:
                raise ValueError(f"unknown precon={precon!r}; use 'illum', 'eprecond3', or 'none'")

        active = _active_keys(invert)
        scales = _block_scales(params)
        base = params

        def pack(p_tuple):
            return np.concatenate([(p_tuple[_KEY_TO_IDX[k]] / scales[_KEY_TO_IDX[k]]).ravel() for k in active])

        def unpack(x):
            full = list(base)
            for j, k in enumerate(active):
                i = _KEY_TO_IDX[k]
                full[i] = (x[j * n_cells:(j + 1) * n_cells] * scales[i]).reshape(int(nz_dom), int(nx_dom))
            return tuple(full)

        # per-cell bounds in scaled space (None = unbounded)
        lo_parts, hi_parts, have_lo, have_hi = [], [], False, False
        for k in active:
            lohi = {"Vs": vs_bounds, "Vp": vp_bounds, "rho": rho_bounds}[k]
            sc = scales[_KEY_TO_IDX[k]]
            lo, hi = (None, None) if lohi is None else lohi
            have_lo |= lo is not None
            have_hi |= hi is not None
            lo_parts.append(np.full(n_cells, -np.inf if lo is None else float(lo) / sc))
            hi_parts.append(np.full(n_cells, np.inf if hi is None else float(hi) / sc))
        lo_t = torch.as_tensor(np.concatenate(lo_parts), device=dev) if have_lo else None
        hi_t = torch.as_tensor(np.concatenate(hi_parts), device=dev) if have_hi else None

        x0 = torch.as_tensor(pack(params), dtype=torch.float64, device=dev)
        plateau = _Plateau(plateau_window, plateau_rel_tol)
        stage_dir = os.path.join(save_dir, f"stage{istage + 1}") if save_dir is not None else None
        if stage_dir is not None and save_every is not None:
            os.makedirs(stage_dir, exist_ok=True)
            np.savez_compressed(os.path.join(stage_dir, "epoch_0000.npz"), Vs=params[0], Vp=params[1],
                                rho=params[2], iter=0, loss=float("nan"))

        eval_count, iter_count = [0], [0]
        loss_history: list = []
        last_x = [x0.clone()]

        def f_and_g(x):
            eval_count[0] += 1
            last_x[0] = x.clone()
            p_curr = unpack(x.cpu().numpy())
            p_proj = project_velocity(*p_curr, K_margin=K_margin, vp_cap=vp_cap)
            (gvs, gvp, grho), loss = compute_gradient(
                backend, Vs=p_proj[0], Vp=p_proj[1], rho=p_proj[2], src_x=src_x, src_z=src_z, src_wavelet=wav,
                observed=observed, rec_x=rec_x, rec_z=rec_z, nx_dom=nx_dom, nz_dom=nz_dom, dx=dx, dz=dz, dt=dt,
                nt=nt, fc=f_c, pad=pad, component=component, invert=invert, scale=auto_scale, sos=sos,
                n_taper=n_taper, fd_order=fd_order, free_surface=free_surface, src_type=src_type,
                block_size=block_size, grad_smooth_sigma=grad_smooth_sigma, source_taper=source_taper,
                illumination=illumination, return_tensors=True)
            full = (gvs, gvp, grho)
            g = torch.cat([(torch.as_tensor(full[_KEY_TO_IDX[k]], device=dev).double() * scales[_KEY_TO_IDX[k]])
                           .reshape(-1) for k in active])
            loss = float(loss)
            loss_history.append(loss)
            if verbose:
                print(f"  iter {eval_count[0]:4d}  loss={loss:.4e}  |g|_inf={float(g.abs().max()):.3e}")
            if stage_dir is not None and save_every is not None and eval_count[0] % save_every == 0:
                np.savez_compressed(os.path.join(stage_dir, f"eval_{eval_count[0]:04d}.npz"), Vs=p_proj[0],
                                    Vp=p_proj[1], rho=p_proj[2], eval=eval_count[0], loss=loss)
            if plateau.push(loss):
                if verbose:
                    print("  plateau detected, stopping the stage")
                raise _PlateauStop()
            return loss, g

        def iter_callback(xk):
            iter_count[0] += 1
            if stage_dir is None or save_every is None or iter_count[0] % save_every != 0:
                return
            p_proj = project_velocity(*unpack(xk.cpu().numpy()), K_margin=K_margin, vp_cap=vp_cap)
            np.savez_compressed(os.path.join(stage_dir, f"epoch_{iter_count[0]:04d}.npz"), Vs=p_proj[0],
                                Vp=p_proj[1], rho=p_proj[2], iter=iter_count[0],
                                loss=float(loss_history[-1]) if loss_history else float("nan"))

        grid = (int(nz_dom), int(nx_dom))
        vs_only = active == ["Vs"]

        def learned_dir(x, g):
            # the network sees the raw Vs gradient (the scaled one differs by
            # a constant factor, which its input normalisation removes)
            return learned.direction(g.reshape(grid), illumination).reshape(-1)

        def on_step(x_old, g_old, x_new, g_new, was_learned):
            collect.add_pair(g_old.reshape(grid), illumination, (x_new - x_old).reshape(grid))

        t0 = time.time()
        n_learned = n_fallback = 0
        try:
            res = minimize(f_and_g, x0, lo=lo_t, hi=hi_t, maxiter=n_iters, m=10, maxls=20, callback=iter_callback,
                           direction_fn=learned_dir if (learned is not None and vs_only) else None,
                           on_step=on_step if (collect is not None and vs_only) else None)
            res_x, res_fun, ok = res["x"], float(res["fun"]), bool(res["success"])
            n_learned, n_fallback = res["n_learned"], res["n_fallback"]
        except _PlateauStop:
            res_x = last_x[0]
            res_fun = loss_history[-1] if loss_history else float("nan")
            ok = False
        t_stage = time.time() - t0
        params = project_velocity(*unpack(res_x.cpu().numpy()), K_margin=K_margin, vp_cap=vp_cap)
        history.append(dict(stage=istage, f_c=f_c, cutoff=cutoff, invert=invert, n_iters_max=n_iters,
                            n_evals=eval_count[0], loss_history=np.array(loss_history), final_loss=res_fun,
                            success=ok, time_s=t_stage, Vs=params[0].copy(), Vp=params[1].copy(),
                            rho=params[2].copy(), n_learned=n_learned, n_fallback=n_fallback))
        if verbose:
            print(f"  Stage {istage + 1} done in {t_stage:.1f}s  final_loss={res_fun:.4e}  success={ok}")
        if save_dir is not None:
            sd = os.path.join(save_dir, f"stage{istage + 1}")
            os.makedirs(sd, exist_ok=True)
            np.savez_compressed(os.path.join(sd, "params.npz"), Vs=params[0], Vp=params[1], rho=params[2],
                                loss_history=np.array(loss_history), f_c=f_c, cutoff=cutoff, invert=invert,
                                backend=backend)
    return params, history
