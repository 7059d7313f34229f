"""Drop-ins for fwi/optimizer.py: the epoch loops of the jaxopt L-BFGS
inversion, on the device float64 L-BFGS of ``stfwi_amd.lbfgs``.

    run_lbfgs(loss_fn, init_params, n_epochs, ...)              fwi/optimizer.py:20-168
    run_lbfgs_alternating(loss_fn, init_params, n_epochs, ...)  fwi/optimizer.py:171-284

Same signatures and semantics: one quasi-Newton iteration per epoch, the
step length capped by max_stepsize = step_pct * max|p| / max|g| (the most
restrictive of the inverted parameters, recomputed from the current
gradient every epoch; the first trial step of a search without curvature
pairs is that cap, i.e. a step_pct change of the model), at most 40
line-search evaluations, early stop when
loss / first loss falls below early_stop_ratio, and the inverted models and
the loss history saved as ``Vs_inverted.npy``, ``Vp_inverted.npy``,
``rho_inverted.npy`` and ``loss_history.npy`` in save_dir.

Differences, stated: jaxopt (absent here) is replaced by a strong-Wolfe
L-BFGS of its own, so the iterates are parity-unpinned step by step; the
history keeps at most 64 pairs (history_size is clipped); the figures the
reference saves every save_every epochs are not drawn (plotting is outside
this package), so true_models, run_shot, observed and the plot arguments are
accepted and unused.
"""
from __future__ import annotations

import os
import time

import numpy as np
import torch

from ..lbfgs import LBFGSMemory, minimize
from ._common import device


def _pack(params):
    return torch.cat([torch.as_tensor(np.asarray(p) if not isinstance(p, torch.Tensor) else p)
                      .to(device=device(), dtype=torch.float64).reshape(-1) for p in params])


def _unpack(x, shapes):
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(x[o:o + n].reshape(sh))
        o += n
    return tuple(out)


def _active(invert):
    act = [0]
    if invert in ("vs_vp", "all"):
        act.append(1)
    if invert == "all":
        act.append(2)
    return act


def _stepsize(params, grads, idx, step_pct):
    p_max = float(torch.max(torch.abs(params[idx])))
    g_max = float(torch.max(torch.abs(grads[idx])))
    return step_pct * p_max / (g_max + 1e-30)


def _objective(loss_fn, shapes, keep=None):
    """fun(x) -> (loss, packed float64 gradient); keep lists the parameters
    whose gradient is kept (the others are zeroed)."""
    def fun(x):
        loss, grads = loss_fn(_unpack(x, shapes))
        gs = [torch.as_tensor(g).to(device=device(), dtype=torch.float64) for g in grads]
        if keep is not None:
            gs = [g if i in keep else torch.zeros_like(g) for i, g in enumerate(gs)]
        return float(loss), torch.cat([g.reshape(-1) for g in gs])
    return fun


def _save(save_dir, params, loss_history):
    names = ("Vs_inverted.npy", "Vp_inverted.npy", "rho_inverted.npy")
    for n, p in zip(names, params):
        np.save(os.path.join(save_dir, n), p.detach().cpu().numpy())
    np.save(os.path.join(save_dir, "loss_history.npy"), np.array(loss_history))


def run_lbfgs(loss_fn, init_params, n_epochs, save_every=20, save_dir='results', true_models=None,
              history_size=100, dx=1.0, dz=1.0, run_shot=None, observed=None, src_x_list=None, src_z=None,
              src_wavelet=None, dt=None, component=1, invert='all', early_stop_ratio=None, step_pct=0.05):
    """L-BFGS inversion over (Vs, Vp, rho); returns (params, loss_history)
    with params float64 numpy arrays."""
    os.makedirs(save_dir, exist_ok=True)
    shapes = [tuple(np.shape(p)) for p in init_params]
    x = _pack(init_params)
    act = _active(invert)
    fun = _objective(loss_fn, shapes)
    f, g = fun(x)
    params, grads = _unpack(x, shapes), _unpack(g, shapes)
    max_step = min(_stepsize(params, grads, i, step_pct) for i in act)
    mem = LBFGSMemory(x.numel(), min(int(history_size), 64), x.device)
    loss_history, loss_first = [], None
    t_start = time.time()
    for epoch in range(1, n_epochs + 1):
        res = minimize(fun, x, maxiter=1, maxls=40, memory=mem, max_step=max_step, fg0=(f, g), t_init=max_step)
        x, f, g = res["x"], res["fun"], res["jac"]
        loss_history.append(f)
        if loss_first is None:
            loss_first = loss_history[0]
        params, grads = _unpack(x, shapes), _unpack(g, shapes)
        max_step = min(_stepsize(params, grads, i, step_pct) for i in act)
        ratio = loss_history[-1] / (loss_first + 1e-30)
        print(f"  Epoch {epoch:04d}/{n_epochs}  Loss: {loss_history[-1]:.4e}  ratio: {ratio:.6f}  step={max_step:.4e}")
        if early_stop_ratio is not None and ratio < early_stop_ratio:
            break
        if res["nit"] == 0:
            break
    mins, secs = divmod(time.time() - t_start, 60)
    print(f"\nTotal inversion time: {int(mins)}m {secs:.1f}s")
    params = _unpack(x, shapes)
    _save(save_dir, params, loss_history)
    return tuple(p.detach().cpu().numpy() for p in params), loss_history


def run_lbfgs_alternating(loss_fn, init_params, n_epochs, save_every=20, save_dir='results', true_models=None,
                          history_size=100, dx=1.0, dz=1.0, run_shot=None, observed=None, src_x_list=None,
                          src_z=None, src_wavelet=None, dt=None, component=1, invert='vs_vp',
                          early_stop_ratio=None, step_pct=0.05):
    """Alternating L-BFGS: each epoch one Vs iteration (Vp, rho gradients
    zeroed, its own history and step cap), then one Vp iteration likewise;
    the epoch's loss is the one after the Vp step."""
    os.makedirs(save_dir, exist_ok=True)
    shapes = [tuple(np.shape(p)) for p in init_params]
    x = _pack(init_params)
    fun_vs = _objective(loss_fn, shapes, keep=(0,))
    fun_vp = _objective(loss_fn, shapes, keep=(1,))
    full = _objective(loss_fn, shapes)
    _, g = full(x)
    params, grads = _unpack(x, shapes), _unpack(g, shapes)
    ss_vs, ss_vp = _stepsize(params, grads, 0, step_pct), _stepsize(params, grads, 1, step_pct)
    m = min(int(history_size), 64)
    mem_vs, mem_vp = LBFGSMemory(x.numel(), m, x.device), LBFGSMemory(x.numel(), m, x.device)
    loss_history, loss_first = [], None
    t_start = time.time()
    for epoch in range(1, n_epochs + 1):
        r1 = minimize(fun_vs, x, maxiter=1, maxls=40, memory=mem_vs, max_step=ss_vs, t_init=ss_vs)
        r2 = minimize(fun_vp, r1["x"], maxiter=1, maxls=40, memory=mem_vp, max_step=ss_vp, t_init=ss_vp)
        x = r2["x"]
        loss_val = r2["fun"]
        loss_history.append(loss_val)
        if loss_first is None:
            loss_first = loss_history[0]
        params = _unpack(x, shapes)
        ss_vs = _stepsize(params, _unpack(r1["jac"], shapes), 0, step_pct)
        ss_vp = _stepsize(params, _unpack(r2["jac"], shapes), 1, step_pct)
        ratio = loss_val / (loss_first + 1e-30)
        print(f"  Epoch {epoch:04d}/{n_epochs}  Loss: {loss_val:.4e}  ratio: {ratio:.6f}  "
              f"ss_vs={ss_vs:.4e}  ss_vp={ss_vp:.4e}")
        if early_stop_ratio is not None and ratio < early_stop_ratio:
            break
        if r1["nit"] == 0 and r2["nit"] == 0:
            break
    mins, secs = divmod(time.time() - t_start, 60)
    print(f"\nTotal inversion time: {int(mins)}m {secs:.1f}s")
    params = _unpack(x, shapes)
    _save(save_dir, params, loss_history)
    return tuple(p.detach().cpu().numpy() for p in params), loss_history
