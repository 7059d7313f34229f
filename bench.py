"""Benchmark: FP32 shot-gradients/s of the elastic velocity-strain FWI
gradient path at 152x500, nt=3000 (25 moment sources, 500 receivers, fd8,
free surface, 20-cell C-PML), on one MI355X.

A step = one misfit+gradient evaluation of all 25 shots: batched forward
with strain history, residual prep, adjoint with the exact imaging condition
fused into each reverse step, shot reduction and chain rule to Vs.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NZ, NX, PAD, FD, NT, DT, FC = 152, 500, 20, 8, 3000, 1.5e-3, 8.0
DX = DZ = 20.0
NSHOT, NREC = 25, 500
HBM_PEAK = 8.0e12

# HBM bytes per 25-shot launch of the velocity-adjoint kernel (exact imaging)
# by tile height, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes over
# `bench.py --kernels-only --tiles 8,8,16,8` and `--tiles 8,4,8,4`, scaled by
# the counters' readings for a 256 MiB device copy (reads x 2.00, writes x 1.00).
ADJ_VEL_TRAFFIC = {4: (51545.2 * 2.0 + 38292.7) * 1024, 8: (51842.6 * 2.0 + 38380.1) * 1024}


def algorithmic_bytes_per_shot_gradient(s=4, k=1, nt=NT):
    """Minimal single-pass traffic: every array touched once per step.
    forward step: s (13 N + 10 N_pml + 3 N_p + 2 nrec)
    adjoint step: s (13 N + 10 N_pml + 3 N_p + 2 k N_p + 2 nrec)."""
    nz_t, nx_t = NZ + PAD, NX + 2 * PAD
    N = nz_t * nx_t
    Np = NZ * NX
    Npml = nz_t * 2 * PAD + PAD * nx_t     # cells of the x- and z-strips
    fwd = s * (13 * N + 10 * Npml + 3 * Np + 2 * NREC)
    adj = s * (13 * N + 10 * Npml + 3 * Np + 2 * k * Np + 2 * NREC)
    return nt * (fwd + adj), fwd, adj


def dominant_kernel_bytes(s=4, k=1):
    """Algorithmic bytes of one launch of the velocity-adjoint kernel for all
    shots: reads U_a, V_a, B, L2M, M (5 N), read+write exx_a, ezz_a, exz_a
    (6 N), read+write of its 4 C-PML strips (8 N_strip), the 3 history
    frames (3 N_p) and k gradient accumulators (2 k N_p)."""
    nz_t, nx_t = NZ + PAD, NX + 2 * PAD
    N = nz_t * nx_t
    Np = NZ * NX
    nxs = nz_t * 2 * PAD
    nzs = PAD * nx_t
    return NSHOT * s * (11 * N + 2 * 2 * (nxs + nzs) + 3 * Np + 2 * k * Np)


def survey(seed=0):
    from stfwi_amd.models import layered_model, smooth_initial, ricker
    vs_t, vp_t, rho_t = layered_model(NZ, NX, seed=seed)
    vs0, vp0, rho0 = smooth_initial(vs_t, vp_t, rho_t, 10.0)
    src_x = (10 + 20 * np.arange(NSHOT)) % NX
    src_z = np.full(NSHOT, 2)
    rec_x = np.arange(NREC)
    rec_z = np.zeros(NREC, int)
    wav = ricker(NT, DT, FC)
    return dict(true=(vs_t, vp_t, rho_t), init=(vs0, vp0, rho0), src_x=src_x, src_z=src_z,
                rec_x=rec_x, rec_z=rec_z, wav=wav)


def cpu_baseline(sv, budget_s=15.0):
    """The oracle (C restatement, FP32, OpenMP on the host cores) on a bounded
    sample of the workload: one shot's forward + exact adjoint for as many
    steps as fit in the budget, scaled to shot-gradients/s."""
    import oracle
    from oracle import signal as osig
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    cores = min(cores, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    vs0, vp0, rho0 = sv["init"]
    nt_s = 100
    g = oracle.Geom(NZ, NX, PAD, True, FD, nt_s, DX, DZ, DT, FC)
    t0 = time.perf_counter()
    reps = 0
    while True:
        rec, t = oracle.forward(g, vs0, vp0, rho0, sv["src_x"][0], sv["src_z"][0], sv["wav"][:nt_s],
                                sv["rec_x"], sv["rec_z"], "moment", np.float32, ("e", "dxu0"))
        ra = np.zeros_like(rec)
        ra[2] = rec[2]
        oracle.adjoint(g, vs0, vp0, rho0, sv["src_x"][0], sv["src_z"][0], ra, sv["rec_x"], sv["rec_z"],
                       "moment", sv["wav"][:nt_s], "exact", t, None, np.float32)
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    per_step = el / (reps * nt_s)
    return {"value": 1.0 / (per_step * NT), "unit": "shot-gradients/s", "cores": cores, "kind": "port",
            "sample": f"{reps}x one shot, {nt_s} forward+adjoint steps (exact imaging, FP32) of the "
                      f"152x500 fd8 workload, scaled to nt={NT}; {el:.1f} s"}


def learned_bench(reps=50):
    """The learned preconditioner's matrix-core kernels at the workload's
    grid: one 3x3 convolution layer (32 -> 32 channels, batch 4, 152x500) in
    its forward, data-gradient and weight-gradient forms, timed with HIP
    events on the launch stream; flops = 2 B H W Cin Cout 9 per form, against
    the 2.5 PFLOP/s dense BF16 peak.  Then one training epoch of the
    preconditioner on 8 pairs."""
    import torch
    from stfwi_amd import learned as L
    dev = torch.device("cuda", 0)
    g = torch.Generator().manual_seed(0)
    B, ci, co, h, w = 4, 32, 32, NZ, NX
    x = torch.randn((B, ci, h, w), generator=g).to(dev)
    wt = (torch.randn((co, ci, 3, 3), generator=g) * 0.1).to(dev)
    bias = torch.zeros(co, device=dev)
    gy = torch.randn((B, co, h, w), generator=g).to(dev)
    flops = 2.0 * B * h * w * ci * co * 9
    forms = {"conv_fwd": lambda: L.conv(x, wt, bias, relu=True),
             "conv_dgrad": lambda: L.conv_dgrad(gy, wt),
             "conv_wgrad": lambda: L.conv_wgrad(x, gy, wt.shape)}
    out = {}
    for name, f in forms.items():
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        out[name] = {"us": round(us, 2), "TFLOP/s": round(flops / us / 1e6, 1),
                     "frac_of_bf16_peak": round(flops / us / 1e6 / 2500.0, 4)}
    lp = L.LearnedPreconditioner(NZ, NX, seed=1)
    rng = np.random.default_rng(0)
    for _ in range(8):
        gr = rng.standard_normal((NZ, NX))
        lp.add_pair(gr, np.ones((NZ, NX)), np.full((NZ, NX), 2000.0), -gr)
    lp.train(epochs=1, lr=1e-3, batch=4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lp.train(epochs=5, lr=1e-3, batch=4)
    torch.cuda.synchronize()
    out["train_epoch_ms"] = round((time.perf_counter() - t0) / 5 * 1e3, 2)
    out["layer"] = f"batch {B}, {ci}->{co} channels, 3x3, {h}x{w}, BF16 operands, FP32 accumulation"
    out["unet_layers"] = unet_layer_table(lp)
    return out


def unet_layer_table(lp, batch=8, reps=20):
    """Every convolution of the preconditioner's U-Net at its training shape
    (batch 8 at its level's resolution of the padded 152x504 grid): time per
    form (forward, data gradient, weight gradient) with HIP events, the
    matrix-core fraction (2 B H W Cin Cout k^2 flop against 2.5 PFLOP/s
    dense BF16) and the memory fraction (FP32 input and output activations
    plus weights, once each, against 8 TB/s)."""
    import torch
    from stfwi_amd import learned as L
    net = lp.net
    dev = net.device
    g = torch.Generator().manual_seed(1)
    rows = []
    level = {}
    for l, c in enumerate(net.channels):
        level[f"enc{l}a"] = level[f"enc{l}b"] = level[f"dec{l}a"] = level[f"dec{l}b"] = l
    level["out"] = 0
    for name, cin, cout, k in net.specs:
        sc = 1 << level[name]
        h, w = lp.H // sc, lp.W // sc
        x = torch.randn((batch, cin, h, w), generator=g).to(dev)
        gy = torch.randn((batch, cout, h, w), generator=g).to(dev)
        wt, bias = net.params[name]
        flops = 2.0 * batch * h * w * cin * cout * k * k
        nbytes = 4.0 * (batch * h * w * (cin + cout) + wt.numel())
        row = {"layer": name, "shape": f"{cin}->{cout} k{k} {h}x{w}"}
        for form, f in (("fwd", lambda: L.conv(x, wt, bias, relu=name != "out")),
                        ("dgrad", lambda: L.conv_dgrad(gy, wt)),
                        ("wgrad", lambda: L.conv_wgrad(x, gy, wt.shape))):
            f()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                f()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / reps
            row[form] = {"us": round(us, 2), "mfma_frac": round(flops / (us * 1e-6) / 2.5e15, 4),
                         "hbm_frac": round(nbytes / (us * 1e-6) / HBM_PEAK, 3)}
        rows.append(row)
    return rows


def batched_bench(steps, warmup):
    """64 uneven shots (record lengths 1500-4000 steps, 120-500 receivers,
    force and moment sources, fd 4 and 8; stfwi_amd.surveys.batched_survey):
    one batched launch per fd order, each shot's history sized by its own
    record.  A step = the misfit gradient of all 64 shots (two plans, one per
    fd order, run back to back), exact imaging of Vs, FP32."""
    import torch
    from stfwi_amd import Plan
    from stfwi_amd.surveys import batched_survey, algorithmic_bytes
    sv = batched_survey()
    f32 = torch.float32
    dev = torch.device("cuda", 0)
    runs = []
    for fd in (4, 8):
        idx = [k for k in range(len(sv["src_x"])) if sv["fd_order"][k] == fd]
        nts = [int(sv["nt"][k]) for k in idx]
        p = Plan(sv["nz"], sv["nx"], sv["dx"], sv["dz"], sv["dt"], sv["fc"], sv["pad"], fd_order=fd,
                 free_surface=True, dtype=f32, n_shots_max=len(idx), nt_max=max(nts), nrec_max=500,
                 imaging="exact", grad="vs", store_history=True, hist_frames=sum(nts))
        p.set_groups(3)

        def shots(p=p, idx=idx, nts=nts):
            p.set_shots([sv["src_x"][k] for k in idx], [sv["src_z"][k] for k in idx],
                        [sv["src_type"][k] for k in idx], nts, [sv["rec_x"][k] for k in idx],
                        [sv["rec_z"][k] for k in idx], sv["wavelet"])
        p.set_model(*(torch.as_tensor(a, dtype=f32, device=dev) for a in sv["true"]))
        shots()
        obs = p.forward(channels=(2,))["rec"][:, 2].contiguous()
        scale = float(1.0 / (torch.mean(obs.double() ** 2) + 1e-45))
        init = tuple(torch.as_tensor(a, dtype=f32, device=dev) for a in sv["init"])
        p.set_model(*init)
        tiles, _ = p.autotune(10)
        shots()
        runs.append((p, obs, scale, init, len(idx), tiles))

    def step():
        for p, obs, scale, init, n, _ in runs:
            p.set_model(*init)
            rec = p.forward(channels=(2,))["rec"]
            loss, radj = p.residual(rec, obs, 2, None, 0, scale, n)
            p.adjoint(radj, want_src=False)
            p.finalize(*init)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    nshot = len(sv["src_x"])
    byt = sum(algorithmic_bytes(sv["nz"], sv["nx"], sv["pad"], True, int(sv["nt"][k]), len(sv["rec_x"][k]))
              for k in range(nshot))
    return {"metric": "shot-gradients/s, 64 uneven shots at 152x500 (%HBM roofline)", "value": nshot / el,
            "unit": "shot-gradients/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": el * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded layered model, Ricker 8 Hz)",
            "config": {"workload": "64-shot gradient, 152x500 free surface pad 20, nt 1500-4000 per shot, "
                                   "120-500 receivers per shot, force and moment sources, fd 4 and 8, "
                                   "exact imaging, Vs",
                       "shots": nshot, "record_steps": int(sum(sv["nt"])), "grid": [sv["nz"], sv["nx"]]},
            "path_roofline": {"algorithmic_GB_per_step": byt / 1e9, "achieved_GB_s": byt / el / 1e9,
                              "frac_of_hbm_peak": byt / el / HBM_PEAK},
            "tiles": {fd: r[5] for fd, r in zip((4, 8), runs)}}


def learned_vs_plain(seeds, weights, spec):
    """Time to the plain method's final Vs RMSE, learned-preconditioned
    against plain L-BFGS, on held-out seeded layered models
    (stfwi_amd.learned_eval; 25 moment shots, 152x500, nt=3000, FP32 'ad'
    gradients; bands 3/5/8 Hz x 20 iterations).  `spec` is the network's
    target and how the driver uses it, mode:policy:maxls.  Prints one JSON
    line; per-seed rows go to stderr as they finish."""
    from stfwi_amd import learned_eval as E
    name = spec
    lp = E.variant(weights, spec)
    rows = []
    for seed in seeds:
        r = E.evaluate_seed(seed, {name: lp})
        print(json.dumps({k: v for k, v in r.items() if k != "plain_trace"}), file=sys.stderr, flush=True)
        rows.append(r)
    ok = [r for r in rows if r["plain_improves"]]
    reached = [r for r in ok if r["learned"][name]["reached"]]
    ratios = [r["learned"][name]["time_ratio"] for r in reached]
    plain_s = sum(r["plain_s"] for r in ok)
    # a seed whose learned run never reaches the plain RMSE counts with its
    # whole run time plus a plain run (the method falls back to the plain
    # inversion); its run time alone is a lower bound on its time to the
    # target, reported beside
    L = lambda r: r["learned"][name]
    learned_s = sum(L(r)["seconds"] if L(r)["reached"] else L(r)["run_s"] + r["plain_s"] for r in ok)
    learned_lb = sum(L(r)["seconds"] if L(r)["reached"] else L(r)["run_s"] for r in ok)
    reached_l = sum(L(r)["seconds"] for r in reached)
    reached_p = sum(r["plain_s"] for r in reached)
    trained = None
    sfile = os.path.join(weights, "seeds.txt")
    if os.path.exists(sfile):
        with open(sfile) as f:
            trained = [ln.strip() for ln in f if ln.strip()]
    return {"metric": "time to the plain method's final Vs RMSE, learned vs plain (ratio)",
            "value": learned_s / plain_s if plain_s else None, "unit": "learned seconds / plain seconds",
            "higher_is_better": False, "n_gpus": 1, "dtype": "f32", "data": "synthetic (seeded layered models)",
            "config": {"workload": "25-shot 152x500 nt=3000 FWI, Vs only, bands 3/5/8 Hz x 20 L-BFGS iterations",
                       "held_out_seeds": [int(s) for s in seeds], "variant": spec, "weights": weights,
                       "training_runs": trained},
            "seeds_where_plain_lowers_rmse": len(ok), "reached": len(reached),
            "median_time_ratio_of_reached": float(np.median(ratios)) if ratios else None,
            "ratio_lower_bound": learned_lb / plain_s if plain_s else None,
            "ratio_of_reached_seeds": reached_l / reached_p if reached_p else None,
            "plain_seconds": plain_s, "learned_seconds": learned_s,
            "plain_evaluations": sum(r["plain_evals"] for r in ok),
            "learned_evaluations": sum(r["learned"][name]["evals"] for r in ok),
            "learned_steps": sum(r["learned"][name]["learned_steps"] for r in ok),
            "fallbacks": sum(r["learned"][name]["fallbacks"] for r in ok),
            "per_seed": [{"seed": r["seed"], "rmse0": round(r["rmse_initial"], 2),
                          "plain_rmse": round(r["plain_rmse_final"], 2), "plain_s": round(r["plain_s"], 2),
                          "learned_s": r["learned"][name]["seconds"], "learned_run_s": round(
                              r["learned"][name]["run_s"], 2), "learned_rmse": round(
                              r["learned"][name]["rmse_final"], 2)} for r in rows]}


def main():
    global NSHOT
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--imaging", default="exact", choices=["exact", "zhou"])
    ap.add_argument("--shots", type=int, default=NSHOT,
                    help="shots per gradient (the workload is 25; other values are for experiments)")
    ap.add_argument("--groups", type=int, default=3, help="concurrent stream groups of shots (1-4)")
    ap.add_argument("--tiles", default=None,
                    help="fixed tile heights of the four time-step kernels, e.g. 8,8,8,4 (default: autotune)")
    ap.add_argument("--workload", default="survey", choices=["survey", "batched"],
                    help="survey: the 25-shot 152x500 nt=3000 gradient (default); batched: 64 uneven shots")
    ap.add_argument("--learned-vs-plain", action="store_true",
                    help="time to the plain method's final Vs RMSE, learned vs plain, on held-out seeds")
    ap.add_argument("--seeds", default="2000-2015", help="held-out seeds for --learned-vs-plain (a-b)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "stfwi_amd", "weights"),
                    help="directory of the trained preconditioner (<mode>.pt)")
    ap.add_argument("--variant", default="final:stage:2",
                    help="network target and how the driver uses it, mode:policy:maxls (--learned-vs-plain)")
    ap.add_argument("--learned", action="store_true",
                    help="time the learned preconditioner's matrix-core convolution kernels and one "
                         "training epoch instead of the gradient path")
    ap.add_argument("--kernels-only", action="store_true",
                    help="after one forward, time each time-step kernel (20 launches) and a "
                         "256 MiB device copy; for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); more GPUs would split the "
              "shots, with one all-reduce of the gradient and misfit per evaluation, and that path "
              "is not built", file=sys.stderr)
        sys.exit(2)
    import torch
    from stfwi_amd import Plan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.learned:
        print(json.dumps({"learned": learned_bench()}))
        return
    if args.learned_vs_plain:
        a, b = (int(v) for v in args.seeds.split("-"))
        print(json.dumps(learned_vs_plain(range(a, b + 1), args.weights, args.variant)))
        return
    if args.workload == "batched":
        print(json.dumps(batched_bench(args.steps, args.warmup)))
        return
    NSHOT = args.shots
    sv = survey(0)
    f32 = torch.float32
    plan = Plan(NZ, NX, DX, DZ, DT, FC, PAD, fd_order=FD, free_surface=True, dtype=f32, n_shots_max=NSHOT,
                nt_max=NT, nrec_max=NREC, imaging=args.imaging, grad="vs", store_history=True)
    if args.kernels_only:
        # no observed data: time each time-step kernel on the plan's buffers
        plan.set_model(*(torch.as_tensor(a, dtype=f32, device=dev) for a in sv["init"]))
        plan.set_shots(sv["src_x"], sv["src_z"], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
        if args.tiles:
            tiles = [int(t) for t in args.tiles.split(",")]
            plan.set_tiles(tiles)
        else:
            tiles, _ = plan.autotune(10)
        res = {n: plan.time_kernel(i, 20) for i, n in enumerate(("fwd_vel", "fwd_strain", "adj_strain", "adj_vel"))}
        a = torch.zeros(64 << 20, dtype=f32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)                       # calibration: 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        print(json.dumps({"kernels_only": res, "tiles": tiles, "copy_bytes": a.numel() * 4}))
        return
    plan.set_groups(args.groups)
    # observed data on the true model
    plan.set_model(*sv["true"])
    plan.set_shots(sv["src_x"], sv["src_z"], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
    obs = plan.forward(channels=(2,))["rec"][:, 2].contiguous()
    scale = float(1.0 / (torch.mean(obs.double() ** 2) + 1e-45))
    vs0, vp0, rho0 = (torch.as_tensor(a, dtype=f32, device=dev) for a in sv["init"])
    plan.set_model(vs0, vp0, rho0)
    if args.tiles:
        tiles, tune = [int(t) for t in args.tiles.split(",")], None
        plan.set_tiles(tiles)
    else:
        tiles, tune = plan.autotune(20)

    def step():
        plan.set_model(vs0, vp0, rho0)
        rec = plan.forward(channels=(2,))["rec"]
        loss, radj = plan.residual(rec, obs, 2, None, 0, scale, NSHOT)
        plan.adjoint(radj, want_src=False)
        g = plan.finalize(vs0, vp0, rho0)
        return loss, g

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, g = step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = el / args.steps * 1e3
    value = NSHOT / (el / args.steps)
    ok = bool(torch.isfinite(g).all()) and bool(torch.isfinite(loss).all())

    # dominant kernel timed live, bracketed with HIP events on the plan's
    # stream: 200 steps of the velocity-adjoint kernel launched as the
    # reverse time loop launches it (one launch per shot group, the groups on
    # their own streams), and beside it one launch over all shots
    names = ("fwd_vel", "fwd_strain", "adj_strain", "adj_vel")
    per_kernel = {n: plan.time_kernel(i, 200, grouped=True) for i, n in enumerate(names)}
    single = {n: plan.time_kernel(i, 200) for i, n in enumerate(names)}
    step_s = per_kernel["adj_vel"] * 1e-6
    kb = dominant_kernel_bytes(4, 1)
    sizes = [NSHOT * k // args.groups - NSHOT * (k - 1) // args.groups for k in range(1, args.groups + 1)]
    kern = {"bound": "hbm", "achieved": kb / step_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": kb / step_s / HBM_PEAK, "traffic": ADJ_VEL_TRAFFIC.get(tiles[3]) if args.imaging == "exact" else None,
            "kernel": f"k_adj_vel<float,8,{tiles[3]}>",
            "launch_shape": f"{args.groups} concurrent stream groups of {'/'.join(map(str, sizes))} shots, "
                            "as the gradient runs it",
            "step_us": step_s * 1e6, "algorithmic_bytes_per_step": kb,
            "single_launch": {"shots": NSHOT, "avg_launch_us": single["adj_vel"],
                              "frac": kb / (single["adj_vel"] * 1e-6) / HBM_PEAK}}
    tot, _, _ = algorithmic_bytes_per_shot_gradient(4, 1)
    path = {"algorithmic_GB_per_shot_gradient": tot / 1e9, "achieved_GB_s": value * tot / 1e9,
            "frac_of_hbm_peak": value * tot / HBM_PEAK}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sv)
    out = {"metric": "shot-gradients/s at 152x500 nt=3000 (%HBM roofline)", "value": value,
           "unit": "shot-gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded layered model, Ricker 8 Hz)",
           "config": {"workload": f"{NSHOT}-shot gradient, 152x500 fd8 free surface pad 20, nt=3000, "
                                  f"500 receivers, moment sources, {args.imaging} imaging, Vs",
                      "shots": NSHOT, "nt": NT, "grid": [NZ, NX], "imaging": args.imaging},
           "roofline": kern, "kernel_step_us": per_kernel, "kernel_single_launch_us": single, "tiles": tiles, "groups": args.groups, "autotune_us": tune, "path_roofline": path, "cpu_baseline": cpu,
           "finite": ok, "loss": float(loss.double().sum())}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
