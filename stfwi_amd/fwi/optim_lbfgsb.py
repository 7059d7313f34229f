"""Drop-in for fwi/optim_lbfgsb.py: the multistage plain-method FWI loop.

    run_multiscale_fwi(*, Vs_init, Vp_init, rho_init, src_x, src_z, rec_x,
        rec_z, nx_dom, nz_dom, dx, dz, dt, nt, pad, observed, stages,
        component=2, backend='ad', fd_order=8, free_surface=True,
        src_type='force', block_size=200, auto_scale=None, K_margin=0.04,
        cfl_coeff=None, vs_bounds=None, vp_bounds=None, rho_bounds=None,
        plateau_window=None, plateau_rel_tol=1e-3, save_dir=None,
        save_every=None, verbose=True)              fwi/optim_lbfgsb.py:137-473
    -> ((Vs, Vp, rho) float64 numpy, list of per-stage history dicts)

The interface is the reference's: the keyword arguments and their defaults,
the stage keys (f_c, cutoff, n_iters, invert, n_taper_pct,
illumination_eps, grad_smooth_sigma, source_taper, precon), the history
dict keys, and the checkpoint files and their keys under save_dir
(stageN/epoch_NNNN.npz, stageN/eval_NNNN.npz, stageN/params.npz).  So are
the stage semantics: a Ricker wavelet per stage with t0 = 1.5 / f_c, a
6th-order Butterworth low-pass, a cosine end taper, an illumination
multiplier ('illum' = H_inv, or 'eprecond3') computed once at the stage
start, optimisation in coordinates where each inverted block is divided by
its largest |value| at the stage start, the velocity projection before every
evaluation, per-cell bounds, and a stop when the misfit plateaus.

The loop itself is built differently.  The model never leaves the device
during a stage: a ``_Coordinates`` object maps between the model tensors
and the optimiser's float64 device vector, an ``_Objective`` evaluates the
batched GPU gradient of all shots at a vector and keeps the stage's
bookkeeping, and a ``_Checkpoints`` object writes the npz files.  The
optimiser is this package's FP64 device L-BFGS (stfwi_amd.lbfgs: 10 pairs,
20 line-search evaluations, no gradient or function tolerance), not scipy's
L-BFGS-B; bounds are kept by projecting the trial points.

Keyword arguments beyond the reference's:
  learned  an object with ``direction(grad, illum, vs) -> (nz, nx)`` tensor
           (stfwi_amd.learned.LearnedPreconditioner): its direction is tried
           before the L-BFGS step of a Vs-only stage, and the plain step is
           taken whenever it does not decrease the misfit within
           ``learned.maxls`` evaluations (default 10; history keys
           n_learned, n_fallback).  ``learned.policy`` says when it is
           tried: 'every' iteration (default), every iteration of a stage
           'until_fail' (until one proposal of the stage is rejected), the
           first iteration of each Vs-only 'stage', or only the first
           iteration of the 'run'.  With
           ``learned.absolute`` true the direction is an update in m/s,
           tried as the step itself; otherwise it is scaled to the L-BFGS
           step;
  collect  an object with ``add_step(grad, illum, vs, step)``,
           ``end_stage(vs)`` and optionally ``end_run(vs)``
           (stfwi_amd.learned.PairCollector) that receives the accepted steps
           of Vs-only stages as training pairs;
  honour_src_type  model `src_type` in the gradient's forward (the
           reference's compute_gradient models a force source whatever
           src_type says, fwi/gradient.py:175-200; False keeps that);
  monitor  ``monitor(stage, iteration, vs, seconds) -> bool`` called at the
           start of every stage (iteration 0) and after every accepted
           iteration, with the current Vs (device tensor) and the seconds
           since the run began; returning True ends the run there (history
           key stopped_by_monitor).
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from ..lbfgs import minimize
from ._common import device
from .forward import ricker_jax
from .gradient import compute_gradient
from .illumination import compute_illumination_for_stage
from .preconditioning import eprecond3_preconditioner
from .projections import cfl_vp_cap, project_velocity_device

# inverted blocks of (Vs, Vp, rho) per `invert`
_BLOCKS = {"vs": (0,), "vs_vp": (0, 1), "all": (0, 1, 2)}


@dataclass
class _StageSpec:
    f_c: float
    cutoff: Optional[float]
    n_iters: int
    invert: str
    n_taper_pct: float
    illumination_eps: Optional[float]
    grad_smooth_sigma: Optional[float]
    source_taper: Optional[np.ndarray]
    precon: str

    @classmethod
    def parse(cls, d: dict) -> "_StageSpec":
        invert = d.get("invert", "vs")
        if invert not in _BLOCKS:
            raise ValueError(f"invert={invert!r}; use 'vs', 'vs_vp', or 'all'")
        precon = d.get("precon", "illum")
        if precon not in ("illum", "eprecond3", "none"):
            raise ValueError(f"unknown precon={precon!r}; use 'illum', 'eprecond3', or 'none'")
        return cls(f_c=float(d["f_c"]), cutoff=d.get("cutoff", None), n_iters=int(d.get("n_iters", 50)),
                   invert=invert, n_taper_pct=float(d.get("n_taper_pct", 0.0)),
                   illumination_eps=d.get("illumination_eps", None),
                   grad_smooth_sigma=d.get("grad_smooth_sigma", None), source_taper=d.get("source_taper", None),
                   precon=precon)


class _Coordinates:
    """The optimiser's vector for one stage: the inverted blocks of the model,
    each divided by its largest |value| at the stage start, concatenated; the
    other blocks stay as they are.  Bounds are given in model units and
    converted to this space."""

    def __init__(self, model, blocks, bounds):
        self.base = model
        self.blocks = blocks
        self.shape = tuple(model[0].shape)
        self.n = model[0].numel()
        self.scale = [max(float(m.abs().max()), 1e-12) for m in model]
        self.lo = self._box(bounds, 0)
        self.hi = self._box(bounds, 1)

    def _box(self, bounds, side):
        parts, any_set = [], False
        for b in self.blocks:
            v = None if bounds[b] is None else bounds[b][side]
            any_set |= v is not None
            fill = (-np.inf if side == 0 else np.inf) if v is None else float(v) / self.scale[b]
            parts.append(torch.full((self.n,), fill, dtype=torch.float64, device=self.base[0].device))
        return torch.cat(parts) if any_set else None

    def vector(self, model) -> torch.Tensor:
        return torch.cat([(model[b] / self.scale[b]).reshape(-1) for b in self.blocks])

    def model(self, x: torch.Tensor):
        out = list(self.base)
        for j, b in enumerate(self.blocks):
            out[b] = (x[j * self.n:(j + 1) * self.n] * self.scale[b]).reshape(self.shape)
        return tuple(out)

    def gradient(self, grads) -> torch.Tensor:
        return torch.cat([(grads[b].to(torch.float64) * self.scale[b]).reshape(-1) for b in self.blocks])


class _Checkpoints:
    """npz files of one stage under save_dir/stageN (the reference's names
    and keys); inactive without save_dir, per-iteration files only with
    save_every."""

    def __init__(self, save_dir, stage_no, save_every):
        self.dir = os.path.join(save_dir, f"stage{stage_no}") if save_dir is not None else None
        self.every = save_every

    def _periodic(self, k):
        return self.dir is not None and self.every is not None and k % self.every == 0

    @staticmethod
    def _np(model):
        return {n: m.detach().cpu().numpy() for n, m in zip(("Vs", "Vp", "rho"), model)}

    def _write(self, name, **kw):
        os.makedirs(self.dir, exist_ok=True)
        np.savez_compressed(os.path.join(self.dir, name), **kw)

    def start(self, model):
        if self._periodic(0):
            self._write("epoch_0000.npz", **self._np(model), iter=0, loss=float("nan"))

    def evaluation(self, k, model, loss):
        if self._periodic(k):
            self._write(f"eval_{k:04d}.npz", **self._np(model), eval=k, loss=loss)

    def iteration(self, k, model, loss):
        if self._periodic(k):
            self._write(f"epoch_{k:04d}.npz", **self._np(model), iter=k, loss=loss)

    def finish(self, model, losses, spec, backend):
        if self.dir is not None:
            self._write("params.npz", **self._np(model), loss_history=np.array(losses), f_c=spec.f_c,
                        cutoff=spec.cutoff, invert=spec.invert, backend=backend)


class _Stalled(Exception):
    pass


class _MonitorStop(Exception):
    pass


class _Objective:
    """Misfit and gradient of one stage at an optimiser vector: projection,
    one batched gradient evaluation of all shots, counts and losses, the
    eval checkpoints and the plateau test (mean of the last `window` losses
    not below (1 - rel_tol) times the mean of the `window` before,
    fwi/optim_lbfgsb.py:117-134)."""

    def __init__(self, coords, project, gradient_at, ckpt, window, rel_tol, verbose):
        self.coords, self.project, self.gradient_at, self.ckpt = coords, project, gradient_at, ckpt
        self.window, self.rel_tol, self.verbose = int(window or 0), float(rel_tol), verbose
        self.losses: list = []
        self.last_x = None

    def _plateau(self):
        w, h = self.window, self.losses
        if w <= 0 or len(h) < 2 * w:
            return False
        return float(np.mean(h[-w:])) >= float(np.mean(h[-2 * w:-w])) * (1.0 - self.rel_tol)

    def __call__(self, x):
        self.last_x = x.clone()
        model = self.project(self.coords.model(x))
        grads, loss = self.gradient_at(model)
        g = self.coords.gradient(grads)
        self.losses.append(float(loss))
        k = len(self.losses)
        if self.verbose:
            print(f"  evaluation {k:4d}: misfit {loss:.4e}, max |g| {float(g.abs().max()):.3e}")
        self.ckpt.evaluation(k, model, float(loss))
        if self._plateau():
            if self.verbose:
                print("  misfit plateau: the stage ends here")
            raise _Stalled()
        return float(loss), g


def run_multiscale_fwi(*, Vs_init, Vp_init, rho_init, src_x, src_z, rec_x, rec_z, nx_dom, nz_dom, dx, dz, dt, nt,
                       pad, observed, stages: Sequence[dict], component: int = 2, backend: str = "ad",
                       fd_order: int = 8, free_surface: bool = True, src_type: str = "force", block_size: int = 200,
                       auto_scale: Optional[float] = None, K_margin: float = 0.04, cfl_coeff: Optional[float] = None,
                       vs_bounds=None, vp_bounds=None, rho_bounds=None, plateau_window: Optional[int] = None,
                       plateau_rel_tol: float = 1e-3, save_dir: Optional[str] = None,
                       save_every: Optional[int] = None, verbose: bool = True, learned=None, collect=None,
                       monitor=None, honour_src_type: bool = False):
    if not stages:
        raise ValueError("`stages` must be a non-empty list")
    if learned is not None and getattr(learned, "policy", "every") not in ("every", "until_fail", "stage", "run"):
        raise ValueError(f"learned.policy={learned.policy!r}; use 'every', 'until_fail', 'stage' or 'run'")
    specs = [_StageSpec.parse(s) for s in stages]
    from scipy.signal import butter

    dev = device()
    obs = torch.as_tensor(np.asarray(observed) if not isinstance(observed, torch.Tensor) else observed, device=dev)
    if auto_scale is None:
        auto_scale = float(1.0 / (torch.mean(obs.double() ** 2) + 1e-45))
    if verbose:
        print(f"data scale {auto_scale:.4e}")
    vp_cap = cfl_vp_cap(dx, dz, dt, cfl_coeff) if cfl_coeff is not None else None
    if vp_bounds is not None and vp_bounds[1] is not None:
        vp_cap = vp_bounds[1] if vp_cap is None else min(vp_cap, vp_bounds[1])

    def project(m):
        return project_velocity_device(*m, K_margin=K_margin, vp_cap=vp_cap)

    model = tuple(torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev).clone()
                  for a in (Vs_init, Vp_init, rho_init))
    bounds = (vs_bounds, vp_bounds, rho_bounds)
    grid = (int(nz_dom), int(nx_dom))
    history = []
    t_run = time.time()
    stopped = False
    learned_used = [False]

    for k, spec in enumerate(specs):
        if verbose:
            print(f"stage {k + 1} of {len(specs)}: f_c {spec.f_c} Hz, low-pass {spec.cutoff}, "
                  f"inverting {spec.invert}, at most {spec.n_iters} iterations ({backend})")
        wav = ricker_jax(torch.arange(nt, dtype=torch.float64) * dt, spec.f_c, 1.5 / spec.f_c)
        sos = butter(6, spec.cutoff, fs=1.0 / dt, output="sos") if spec.cutoff is not None else None
        n_taper = int(spec.n_taper_pct * nt)
        illum = None
        if spec.illumination_eps is not None and spec.precon == "illum":
            illum = compute_illumination_for_stage(nz_dom, nx_dom, dx, dz, dt, nt, spec.f_c, pad, block_size, rec_x,
                                                   rec_z, fd_order, component, *model, src_x, src_z, wav,
                                                   eps=spec.illumination_eps)
        elif spec.illumination_eps is not None and spec.precon == "eprecond3":
            illum = eprecond3_preconditioner(nz_dom=nz_dom, nx_dom=nx_dom, dx=dx, dz=dz, dt=dt, nt=nt, f_c=spec.f_c,
                                             pad=pad, rec_x=rec_x, rec_z=rec_z, fd_order=fd_order,
                                             component=component, Vs=model[0], Vp=model[1], rho=model[2],
                                             src_x=src_x, src_z=src_z, src_wavelet=wav, free_surface=free_surface,
                                             eps=spec.illumination_eps)

        def gradient_at(m, spec=spec, wav=wav, sos=sos, n_taper=n_taper, illum=illum):
            return compute_gradient(backend, Vs=m[0], Vp=m[1], rho=m[2], src_x=src_x, src_z=src_z,
                                    src_wavelet=wav, observed=obs, rec_x=rec_x, rec_z=rec_z, nx_dom=nx_dom,
                                    nz_dom=nz_dom, dx=dx, dz=dz, dt=dt, nt=nt, fc=spec.f_c, pad=pad,
                                    component=component, invert=spec.invert, scale=auto_scale, sos=sos,
                                    n_taper=n_taper, fd_order=fd_order, free_surface=free_surface,
                                    src_type=src_type, block_size=block_size,
                                    grad_smooth_sigma=spec.grad_smooth_sigma, source_taper=spec.source_taper,
                                    illumination=illum, return_tensors=True, honour_src_type=honour_src_type)

        blocks = _BLOCKS[spec.invert]
        coords = _Coordinates(model, blocks, bounds)
        ckpt = _Checkpoints(save_dir, k + 1, save_every)
        ckpt.start(model)
        objective = _Objective(coords, project, gradient_at, ckpt, plateau_window, plateau_rel_tol, verbose)
        vs_only = blocks == (0,)
        n_it = [0]
        accepted = [None]
        if monitor is not None and monitor(k, 0, model[0], time.time() - t_run):
            stopped = True
            break

        def after_iteration(x):
            n_it[0] += 1
            accepted[0] = x
            m = project(coords.model(x))
            ckpt.iteration(n_it[0], m, objective.losses[-1] if objective.losses else float("nan"))
            if monitor is not None and monitor(k, n_it[0], m[0], time.time() - t_run):
                raise _MonitorStop()

        absolute = bool(getattr(learned, "absolute", False))
        policy = getattr(learned, "policy", "every")
        counts = {"proposed": 0, "learned": 0}

        def propose(x, g):
            # 'every': before every L-BFGS step; 'until_fail': before every
            # step of a stage until one proposal is rejected; 'stage': at the
            # first iteration of each Vs-only stage; 'run': at the first
            # iteration of the run's first Vs-only stage
            if (policy == "stage" and (n_it[0] > 0 or counts["proposed"] > 0)) or \
                    (policy == "run" and learned_used[0]) or \
                    (policy == "until_fail" and counts["proposed"] > counts["learned"]):
                return None
            counts["proposed"] += 1
            learned_used[0] = True
            # an update in model units becomes a step in the optimiser's
            # coordinates; a unit-free direction is scaled by the line search
            m = project(coords.model(x))
            d = learned.direction(g.reshape(grid), illum, m[0]).reshape(-1) * float(getattr(learned, "scale", 1.0))
            return d / coords.scale[0] if absolute else d

        def on_step(x_old, g_old, x_new, g_new, was_learned):
            counts["learned"] += int(bool(was_learned))
            if collect is not None:
                m = project(coords.model(x_old))
                collect.add_step(g_old.reshape(grid), illum, m[0], (x_new - x_old).reshape(grid) * coords.scale[0])

        t_stage = time.time()
        try:
            res = minimize(objective, coords.vector(model), lo=coords.lo, hi=coords.hi, maxiter=spec.n_iters, m=10,
                           maxls=20, callback=after_iteration,
                           direction_fn=propose if (learned is not None and vs_only) else None,
                           maxls_learned=int(getattr(learned, "maxls", 10)),
                           on_step=on_step if vs_only else None,
                           learned_step="own" if (learned is not None and absolute) else "lbfgs")
            x_end, f_end, ok = res["x"], float(res["fun"]), bool(res["success"])
        except _Stalled:
            x_end, ok = objective.last_x, False
            f_end = objective.losses[-1] if objective.losses else float("nan")
        except _MonitorStop:
            x_end, ok, stopped = accepted[0], True, True
            f_end = objective.losses[-1] if objective.losses else float("nan")
        model = project(coords.model(x_end))
        # a proposal is rejected when the accepted step of its iteration was
        # not the learned one
        n_learned = counts["learned"]
        n_fallback = counts["proposed"] - n_learned
        if collect is not None and vs_only:
            collect.end_stage(model[0])
        seconds = time.time() - t_stage
        history.append(dict(stage=k, f_c=spec.f_c, cutoff=spec.cutoff, invert=spec.invert,
                            n_iters_max=spec.n_iters, n_evals=len(objective.losses),
                            loss_history=np.array(objective.losses), final_loss=f_end, success=ok, time_s=seconds,
                            Vs=model[0].cpu().numpy(), Vp=model[1].cpu().numpy(), rho=model[2].cpu().numpy(),
                            n_iters_done=n_it[0], n_learned=n_learned, n_fallback=n_fallback,
                            stopped_by_monitor=stopped))
        if verbose:
            print(f"  stage {k + 1} finished after {seconds:.1f} s: misfit {f_end:.4e}, success {ok}")
        ckpt.finish(model, objective.losses, spec, backend)
        if stopped:
            break
    if collect is not None and hasattr(collect, "end_run"):
        collect.end_run(model[0])
    return tuple(m.cpu().numpy() for m in model), history
