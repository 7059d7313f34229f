"""Learned-preconditioned inversion against the plain method: training data,
training, and the time-to-RMSE evaluation.  No reference counterpart: the
reference has no learned part.  The plain method is the reference's
multistage driver (fwi/optim_lbfgsb.py, through run_multiscale_fwi) with the
stage settings of its production recipe (Main_strain_fwi.py:129-161:
H_inv illumination, eps 1e-4, Gaussian smoothing sigma = 1, 10 % end taper,
Vs inverted, Vs bounds 500-4500 m/s, the fd8 CFL cap on Vp).

Workload: the seeded 152x500 layered surveys of ``stfwi_amd.surveys``
(25 moment sources, 500 receivers, Ricker 8 Hz, nt = 3000, dt = 1.5 ms, FP32
gradients of backend 'ad' on the GPU), the observed data modelled on the
true model.  Vs starts from the true Vs smoothed with sigma = 10 cells; Vp
and rho are the true ones and stay fixed.  (With Vp and rho smoothed as well
and held fixed, a Vs-only inversion puts their errors into Vs and can raise
the Vs RMSE: 94 -> 172 m/s on seed 2000, as the reference's own demo run went
235.14 -> 235.38 m/s; the plain method's final RMSE is then reached before
the first iteration and measures nothing.)

* ``generate``: runs of the plain method on training seeds, 10 L-BFGS
  iterations each over the three bands (by default 2 at 3 Hz, 2 at 5 Hz and
  6 at 8 Hz: on this model family the plain method's 3 and 5 Hz bands leave
  the Vs RMSE where it was, e.g. 93.9 -> 93.6 m/s on seed 2000 after 40
  iterations, and its 8 Hz band carries the gain, 93.6 -> 82.0); every
  accepted step gives a pair (gradient, illumination, Vs) -> update, with
  the targets 'step' (the accepted step), 'ahead' (the update to the end of
  the band) and 'final' (the update to the end of the run).
* ``train``: Adam on the pairs (MSE over the physical cells), mirrored
  columns as augmentation.  Training may run in chunks of seeds, each
  continuing from the weights and Adam moments the previous chunk saved.
* ``evaluate``: per held-out seed, the plain method with 20 iterations in
  each of the bands 3, 5, 8 Hz, its final Vs RMSE and wall-clock; then the
  learned run over the same bands (the network's update tried before the
  L-BFGS step as its policy says, the plain step whenever it does not lower
  the misfit) stopped as soon as its Vs RMSE reaches the plain final RMSE.
  Wall-clock counts everything inside the driver: illumination, gradients,
  line searches, network inference.
"""
from __future__ import annotations

import time
from typing import Iterable, Optional

import numpy as np
import torch

from .fwi.optim_lbfgsb import run_multiscale_fwi
from .surveys import layered_survey

CUTOFFS = (3.0, 5.0, 8.0)
FD8_S = 1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168


def stages(iters=(20, 20, 20), f_c=8.0, cutoffs=CUTOFFS):
    return [dict(f_c=f_c, cutoff=c, n_iters=int(n), invert="vs", n_taper_pct=0.1, precon="illum",
                 illumination_eps=1e-4, grad_smooth_sigma=1.0) for c, n in zip(cutoffs, iters)]


def survey(seed):
    """layered_survey(seed) with Vp and rho of the initial model set to the
    true ones (only Vs is unknown)."""
    sv = layered_survey(seed)
    sv["init"] = (sv["init"][0], sv["true"][1], sv["true"][2])
    return sv


def observed_data(sv):
    """Component-2 (exx) records of the true model, FP32 on the device (on
    the cached forward-only plan that the illumination also uses)."""
    from .fwi._common import device, get_plan, tune_launches
    n = len(sv["src_x"])
    p = get_plan(nz_dom=int(sv["nz"]), nx_dom=int(sv["nx"]), dx=float(sv["dx"]), dz=float(sv["dz"]),
                 dt=float(sv["dt"]), fc=float(sv["fc"]), pad=int(sv["pad"]), fd_order=int(sv["fd_order"]),
                 free_surface=bool(sv["free_surface"]), dtype=torch.float32, n_shots_max=n, nt_max=int(sv["nt"]),
                 nrec_max=len(sv["rec_x"]), imaging="none", grad="vs", store_history=False, device=device())
    p.set_model(*sv["true"])
    p.set_shots(sv["src_x"], sv["src_z"], "moment", int(sv["nt"]), sv["rec_x"], sv["rec_z"], sv["wavelet"])
    tune_launches(p, n)
    return p.forward(channels=(2,))["rec"][:n, 2, :, :len(sv["rec_x"])].contiguous()


def vs_rmse(vs, vs_true_t):
    return float(torch.sqrt(torch.mean((vs.double() - vs_true_t) ** 2)))


def invert(sv, obs, stage_list, learned=None, collect=None, monitor=None):
    return run_multiscale_fwi(Vs_init=sv["init"][0], Vp_init=sv["init"][1], rho_init=sv["init"][2],
                              src_x=sv["src_x"], src_z=sv["src_z"], rec_x=sv["rec_x"], rec_z=sv["rec_z"],
                              nx_dom=sv["nx"], nz_dom=sv["nz"], dx=sv["dx"], dz=sv["dz"], dt=sv["dt"], nt=sv["nt"],
                              pad=sv["pad"], observed=obs, stages=stage_list, component=2, backend="ad",
                              fd_order=sv["fd_order"], free_surface=True, src_type="moment", K_margin=0.04,
                              cfl_coeff=FD8_S, vs_bounds=(500.0, 4500.0), verbose=False, learned=learned,
                              collect=collect, monitor=monitor, honour_src_type=True)


def generate(seeds: Iterable[int], collector, iters=(2, 2, 6), log=None, budget_s=None):
    """Plain runs on the training seeds; pairs go to `collector`.  With
    budget_s, no new run starts after that many seconds.  Returns the seeds
    used.  The next seed's model is built on a host thread while the
    current run holds the GPU."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    used = []
    seeds = [int(v) for v in seeds]
    if not seeds:
        return used
    with ThreadPoolExecutor(1) as pool:
        nxt = pool.submit(survey, seeds[0])
        for i, seed in enumerate(seeds):
            if budget_s is not None and time.time() - t0 > budget_s:
                break
            sv = nxt.result()
            if i + 1 < len(seeds):
                nxt = pool.submit(survey, seeds[i + 1])
            invert(sv, observed_data(sv), stages(iters), collect=collector)
            used.append(seed)
            if log is not None:
                log(f"training model {seed}: {len(collector)} pairs, {time.time() - t0:.0f} s")
    return used


def train(lp, pairs, epochs=8, lr=1e-3, batch=8, seed=0, log=None, flip=True):
    hist = []
    for e in range(epochs):
        h = lp.train(epochs=1, lr=lr, batch=batch, seed=seed + e, pairs=pairs, flip=flip)
        hist.append(h[0])
        if log is not None:
            log(f"epoch {e + 1}: loss {h[0]:.4e}")
    return hist


class _Track:
    """Vs RMSE after every iteration; optionally stops at a target RMSE."""

    def __init__(self, vs_true, target=None):
        self.vt = torch.as_tensor(vs_true, dtype=torch.float64, device=torch.device("cuda", torch.cuda.current_device()))
        self.target = target
        self.trace = []          # (stage, iteration, seconds, rmse)
        self.hit = None

    def __call__(self, stage, it, vs, seconds):
        r = vs_rmse(vs, self.vt)
        self.trace.append((stage, it, seconds, r))
        if self.target is not None and r <= self.target and self.hit is None:
            self.hit = (stage, it, seconds, r)
            return True
        return False


def evaluate_seed(seed, nets=None, plain_iters=(20, 20, 20), log=None, schedule=None, ablation=False):
    """Plain run, then one learned run per entry of `nets` ({name: lp}; each
    lp carries its own policy and maxls) stopped at the plain final Vs
    RMSE.  Every run keeps its RMSE trace: (stage, iteration, seconds,
    rmse) after each accepted iteration.

    schedule = (cutoffs, iters) gives the learned runs bands of their own
    (the plain run keeps the 3/5/8 Hz bands); ablation=True then also runs
    the plain method on that schedule, without the network ('plain on the
    schedule'), so that what the schedule alone gains is measured beside
    what the network adds."""
    sv = survey(seed)
    obs = observed_data(sv)
    rm0 = vs_rmse(torch.as_tensor(sv["init"][0]), torch.as_tensor(sv["true"][0], dtype=torch.float64))
    tr = _Track(sv["true"][0])
    _, hp = invert(sv, obs, stages(plain_iters), monitor=tr)
    # both runs are timed by the driver's own clock (seconds since the run
    # began, as the monitor sees them)
    t_plain = tr.trace[-1][2]
    rp = tr.trace[-1][3]
    out = dict(seed=int(seed), rmse_initial=rm0, plain_rmse_final=rp, plain_s=t_plain,
               plain_iters=sum(h["n_iters_done"] for h in hp), plain_evals=sum(h["n_evals"] for h in hp),
               plain_improves=rp < rm0, plain_trace=_rounded(tr.trace), learned={})
    if log is not None:
        log(f"seed {seed}: plain RMSE {rm0:.2f} -> {rp:.2f} m/s in {t_plain:.1f} s, {out['plain_evals']} evaluations")
    runs = dict(nets or {})
    if schedule is not None and ablation:
        runs["plain on the schedule"] = None
    st_l = stages(schedule[1], cutoffs=schedule[0]) if schedule is not None else stages(plain_iters)
    for name, lp in runs.items():
        tl = _Track(sv["true"][0], target=rp)
        _, hl = invert(sv, obs, st_l, learned=lp, monitor=tl)
        r = dict(policy=getattr(lp, "policy", None), maxls=getattr(lp, "maxls", None),
                 reached=tl.hit is not None, seconds=tl.hit[2] if tl.hit is not None else None,
                 run_s=tl.trace[-1][2], rmse_final=tl.trace[-1][3], iters=sum(h["n_iters_done"] for h in hl),
                 evals=sum(h["n_evals"] for h in hl), learned_steps=sum(h["n_learned"] for h in hl),
                 fallbacks=sum(h["n_fallback"] for h in hl),
                 time_ratio=(tl.hit[2] / t_plain) if tl.hit is not None else None, trace=_rounded(tl.trace))
        out["learned"][name] = r
        if log is not None:
            log(f"seed {seed} {name}: RMSE {r['rmse_final']:.2f} m/s after {r['run_s']:.1f} s, "
                f"{r['evals']} evaluations, {r['learned_steps']} learned steps, {r['fallbacks']} fallbacks, "
                f"reached {r['reached']}")
    return out


def _rounded(trace):
    return [(int(s), int(i), round(float(t), 3), round(float(r), 3)) for s, i, t, r in trace]


# ---------------------------------------------------------------------- CLI --

def _seeds(text):
    a, b = (int(v) for v in text.split("-"))
    return range(a, b + 1)


def variant(weights_dir, spec):
    """'mode:policy:maxls[:scale][@dir]' -> the network trained for target
    `mode` under `weights_dir` (or `dir`), set to be proposed per `policy`
    with at most `maxls` evaluations of backtracking, its update multiplied
    by `scale` (default 1)."""
    import os
    from .learned import LearnedPreconditioner
    if "@" in spec:
        spec, weights_dir = spec.split("@", 1)
    mode, policy, maxls, scale = (spec.split(":") + ["every", "10", "1"][len(spec.split(":")) - 1:])[:4]
    lp = LearnedPreconditioner.load(os.path.join(weights_dir, f"{mode}.pt"))
    lp.policy, lp.maxls, lp.scale = policy, int(maxls), float(scale)
    return lp


def main(argv=None):
    """python -m stfwi_amd.learned_eval generate-train --seeds 1000-1170
           --modes step,ahead,final --epochs 20 [--init DIR] --out DIR
       python -m stfwi_amd.learned_eval evaluate --seeds 2000-2015
           --variants final:run:3,ahead:stage:3 --weights DIR --out results.jsonl

    generate-train: plain runs on the seeds, then Adam epochs over their
    pairs for one network per target mode (continuing from DIR/<mode>.pt,
    weights and Adam moments, when --init is given), saved as
    <out>/<mode>.pt.  evaluate: one JSON object per seed (the plain run and
    one learned run per variant mode:policy:maxls), appended to --out as
    JSON lines."""
    import argparse
    import json
    import os
    from .learned import LearnedPreconditioner, PairCollector
    ap = argparse.ArgumentParser(prog="python -m stfwi_amd.learned_eval")
    ap.add_argument("command", choices=["generate-train", "evaluate"])
    ap.add_argument("--seeds", required=True)
    ap.add_argument("--modes", default="final")
    ap.add_argument("--variants", default="final:every:10")
    ap.add_argument("--schedule", default=None,
                    help="evaluate: bands of the learned runs, e.g. 8:40 or 5,8:10,30 (cutoffs:iterations)")
    ap.add_argument("--ablation", action="store_true",
                    help="evaluate: also run the plain method on --schedule, without the network")
    ap.add_argument("--gen-iters", default="2,2,6",
                    help="generate-train: plain iterations in the 3, 5 and 8 Hz bands of each training run")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--budget-s", type=float, default=None,
                    help="generate-train: start no new plain run after this many seconds")
    ap.add_argument("--init", default=None)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    t0 = time.time()

    def log(m):
        print(f"[{time.time() - t0:7.1f} s] {m}", flush=True)

    if args.command == "generate-train":
        col = PairCollector(LearnedPreconditioner(152, 500, seed=0))
        gi = tuple(int(v) for v in args.gen_iters.split(","))
        used = generate(_seeds(args.seeds), col, iters=gi, log=log, budget_s=args.budget_s)
        log(f"models {used[0]}-{used[-1]} ({len(used)}), {len(col)} pairs")
        for mode in col.MODES:
            rms = [float(torch.sqrt(torch.mean(t[0, 0, :col.lp.nz, :col.lp.nx] ** 2))) for _, t in col.pairs[mode]]
            log(f"{mode} targets: rms {np.mean(rms):.3f} (x {col.lp.step_ref:g} m/s), max {np.max(rms):.3f}")
        os.makedirs(args.out, exist_ok=True)
        with open(os.path.join(args.out, "seeds.txt"), "a") as f:
            f.write(f"{used[0]}-{used[-1]} iters {args.gen_iters}\n")
        for mode in args.modes.split(","):
            path = os.path.join(args.init, f"{mode}.pt") if args.init else None
            lp = LearnedPreconditioner.load(path) if path and os.path.exists(path) else \
                LearnedPreconditioner(152, 500, seed=0)
            h = train(lp, col.pairs[mode], epochs=args.epochs, lr=args.lr,
                      log=lambda m, mode=mode: log(f"{mode} {m}"))
            lp.save(os.path.join(args.out, f"{mode}.pt"))
            log(f"{mode}: {len(col.pairs[mode])} pairs, loss {h[0]:.4e} -> {h[-1]:.4e}")
    else:
        nets = {v: variant(args.weights, v) for v in args.variants.split(",")} if args.variants else {}
        sched = None
        if args.schedule:
            cs, its = args.schedule.split(":")
            sched = (tuple(float(c) for c in cs.split(",")), tuple(int(n) for n in its.split(",")))
        for seed in _seeds(args.seeds):
            r = evaluate_seed(seed, nets, log=log, schedule=sched, ablation=args.ablation)
            with open(args.out, "a") as f:
                f.write(json.dumps(r) + "\n")


if __name__ == "__main__":
    main()
