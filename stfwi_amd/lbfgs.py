"""Limited-memory BFGS in float64 on the device.

The quasi-Newton step of the plain method.  The reference runs it on the host
with scipy's L-BFGS-B (``run_multiscale_fwi``, fwi/optim_lbfgsb.py:409-434:
maxcor 10, maxls 20, gtol 0, ftol 1e-30, per-cell bounds) or jaxopt.LBFGS
(fwi/optimizer.py:110-117).  Here the iterate, gradient, search direction and
the m stored pairs stay on the GPU in float64; the two-loop recursion
(Nocedal & Wright, Numerical Optimization, 2nd ed., Alg. 7.4), the pair
update, the dot products and the projected step are hand-written HIP kernels
(``stfwi_lbfgs_direction``, ``stfwi_lbfgs_pair``, ``stfwi_dot64``,
``stfwi_step_clip``).  The line search runs on host scalars: strong Wolfe
(N&W Alg. 3.5/3.6, cubic interpolation) without bounds, projected
backtracking Armijo with bounds, where the direction is computed in the
free variables (the binding set -- entries on a bound whose gradient points
out of the box -- is masked out of the gradient and of the direction).

scipy's L-BFGS-B handles bounds with a generalised Cauchy point and subspace
minimisation; this method projects the trial points instead, so iterates
differ from scipy's when bounds are active.  Without active bounds both are
L-BFGS with H0 = (y.s / y.y) I and a Wolfe line search.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, Optional

import torch

from . import _abi

_VP = C.c_void_p


def _p(t):
    return _VP(t.data_ptr()) if t is not None else None


def _stream():
    return _VP(torch.cuda.current_stream().cuda_stream)


class LBFGSMemory:
    """The m most recent (s, y) pairs of an n-vector problem, in float64
    device rows, with rho = 1/(y.s) and the H0 scale gamma = y.s / y.y of
    the newest pair."""

    def __init__(self, n: int, m: int = 10, device=None):
        if not 1 <= m <= 64:
            raise ValueError(f"history size m={m} outside [1, 64]")
        self.n, self.m = int(n), int(m)
        dev = device or torch.device("cuda", torch.cuda.current_device())
        # m + 1 rows: a new pair goes to a free row, so a rejected pair never
        # overwrites a stored one
        self.S = torch.zeros(self.m + 1, self.n, dtype=torch.float64, device=dev)
        self.Y = torch.zeros_like(self.S)
        self._dots = torch.zeros(3, dtype=torch.float64, device=dev)
        self._gd = torch.zeros(1, dtype=torch.float64, device=dev)
        self.slots: list[int] = []      # oldest first
        self.rho: list[float] = []
        self.gamma = 1.0

    def reset(self):
        self.slots.clear()
        self.rho.clear()
        self.gamma = 1.0

    def __len__(self):
        return len(self.slots)

    def direction(self, g: torch.Tensor, out: Optional[torch.Tensor] = None):
        """d = -H g and g . d."""
        d = out if out is not None else torch.empty_like(g)
        k = len(self.slots)
        slots = (C.c_int * max(k, 1))(*self.slots)
        rho = (C.c_double * max(k, 1))(*self.rho)
        _abi.check(_abi.lib().stfwi_lbfgs_direction(_p(g), _p(self.S), _p(self.Y), rho, slots, k, self.n, self.n,
                                                    float(self.gamma), _p(d), _p(self._gd), _stream()))
        return d, float(self._gd.item())

    def push(self, x_new, x_old, g_new, g_old, eps=2.2e-16) -> bool:
        """Store s = x_new - x_old, y = g_new - g_old unless y.s <= eps y.y
        (the curvature test of L-BFGS-B).  Returns whether it was stored."""
        slot = min(set(range(self.m + 1)) - set(self.slots))
        _abi.check(_abi.lib().stfwi_lbfgs_pair(_p(x_new), _p(x_old), _p(g_new), _p(g_old), self.n,
                                               _p(self.S[slot]), _p(self.Y[slot]), _p(self._dots), _stream()))
        ys, yy, _ = (float(v) for v in self._dots.tolist())
        if not (ys > eps * yy) or yy <= 0:
            return False
        if len(self.slots) == self.m:
            self.slots.pop(0)
            self.rho.pop(0)
        self.slots.append(slot)
        self.rho.append(1.0 / ys)
        self.gamma = ys / yy
        return True


def dot(a: torch.Tensor, b: torch.Tensor) -> float:
    out = torch.zeros(1, dtype=torch.float64, device=a.device)
    _abi.check(_abi.lib().stfwi_dot64(_p(a), _p(b), a.numel(), _p(out), _stream()))
    return float(out.item())


def step_clip(x, d, t, lo=None, hi=None, out=None):
    """clip(x + t d, lo, hi)."""
    out = out if out is not None else torch.empty_like(x)
    _abi.check(_abi.lib().stfwi_step_clip(_p(x), _p(d), float(t), _p(lo), _p(hi), x.numel(), _p(out), _stream()))
    return out


def bound_mask(x, g, lo, hi, v):
    """Zero v on the binding set of (x, g) for the box [lo, hi]."""
    _abi.check(_abi.lib().stfwi_bound_mask(_p(x), _p(g), _p(lo), _p(hi), x.numel(), _p(v), _stream()))
    return v


def _cubic_min(a, fa, ga, b, fb, gb):
    """Minimiser of the cubic interpolating (a, fa, ga), (b, fb, gb)
    (N&W eq. 3.59), or None."""
    d1 = ga + gb - 3.0 * (fa - fb) / (a - b)
    disc = d1 * d1 - ga * gb
    if disc < 0:
        return None
    d2 = math.copysign(math.sqrt(disc), b - a)
    den = gb - ga + 2.0 * d2
    if den == 0:
        return None
    return b - (b - a) * (gb + d2 - d1) / den


class _Eval:
    """f and g at clip(x + t d), remembering every trial."""

    def __init__(self, fun, x, d, lo, hi):
        self.fun, self.x, self.d, self.lo, self.hi = fun, x, d, lo, hi
        self.best = None
        self.nfev = 0

    def __call__(self, t):
        xt = step_clip(self.x, self.d, t, self.lo, self.hi)
        f, g = self.fun(xt)
        self.nfev += 1
        rec = (t, float(f), g, xt)
        if self.best is None or rec[1] < self.best[1]:
            self.best = rec
        return rec


def _wolfe(ev: _Eval, f0, gd0, t0, c1, c2, maxls, t_max=1e20):
    """Strong-Wolfe line search (N&W Alg. 3.5 with zoom Alg. 3.6).
    Returns (t, f, g, x) or None."""
    t_prev, f_prev, gd_prev = 0.0, f0, gd0
    t = t0
    for i in range(maxls):
        t, f, g, x = ev(t)
        gd = dot(g, ev.d)
        if f > f0 + c1 * t * gd0 or (i > 0 and f >= f_prev):
            return _zoom(ev, f0, gd0, t_prev, f_prev, gd_prev, t, f, gd, c1, c2, maxls - i - 1)
        if abs(gd) <= -c2 * gd0:
            return t, f, g, x
        if gd >= 0:
            return _zoom(ev, f0, gd0, t, f, gd, t_prev, f_prev, gd_prev, c1, c2, maxls - i - 1)
        t_prev, f_prev, gd_prev = t, f, gd
        t = min(2.0 * t, t_max)
    return None


def _zoom(ev, f0, gd0, lo, flo, gdlo, hi, fhi, gdhi, c1, c2, budget):
    for _ in range(max(budget, 0)):
        t = _cubic_min(lo, flo, gdlo, hi, fhi, gdhi)
        a, b = min(lo, hi), max(lo, hi)
        if t is None or not (a + 0.1 * (b - a) <= t <= b - 0.1 * (b - a)):
            t = 0.5 * (lo + hi)
        t, f, g, x = ev(t)
        gd = dot(g, ev.d)
        if f > f0 + c1 * t * gd0 or f >= flo:
            hi, fhi, gdhi = t, f, gd
        else:
            if abs(gd) <= -c2 * gd0:
                return t, f, g, x
            if gd * (hi - lo) >= 0:
                hi, fhi, gdhi = lo, flo, gdlo
            lo, flo, gdlo = t, f, gd
    return None


def _armijo_projected(ev: _Eval, f0, g0, x0, t0, c1, maxls):
    """Backtracking on the projected path clip(x + t d): accept the first t
    with f <= f0 + c1 g0 . (x_t - x0)."""
    t = t0
    for _ in range(maxls):
        t, f, g, x = ev(t)
        s = x - x0
        if f <= f0 + c1 * dot(g0, s):
            return t, f, g, x
        t *= 0.5
    return None


def minimize(fun: Callable, x0: torch.Tensor, lo: Optional[torch.Tensor] = None, hi: Optional[torch.Tensor] = None,
             maxiter: int = 50, m: int = 10, maxls: int = 20, c1: float = 1e-4, c2: float = 0.9,
             callback: Optional[Callable] = None, memory: Optional[LBFGSMemory] = None,
             direction_fn: Optional[Callable] = None, maxls_learned: int = 10, on_step: Optional[Callable] = None,
             max_step: Optional[float] = None, fg0: Optional[tuple] = None, t_init: Optional[float] = None,
             learned_step: str = "lbfgs"):
    """Minimise fun(x) -> (f, g) over float64 device vectors x, with optional
    per-entry bounds lo <= x <= hi (device tensors or None).

    One iteration = one accepted step; callback(x) runs after each.  Stops
    after maxiter iterations or when the line search finds no decrease.

    direction_fn(x, g) -> d (or None: no proposal at this iterate) proposes
    a search direction, e.g. a learned preconditioner's.  It is tried first,
    scaled to the L-BFGS
    direction's max |.| (learned_step='lbfgs') or taken as the step itself
    (learned_step='own': the first trial is x + d), with projected
    backtracking of at most maxls_learned evaluations; when it is not a
    descent direction or finds no decrease of f, the plain L-BFGS step is
    taken instead.  on_step(x_old, g_old, x_new, g_new, learned) sees every
    accepted step.

    max_step caps the step length t of the line search (jaxopt's
    max_stepsize); fg0 = (f, g) at x0 saves the first evaluation when the
    caller already has it; t_init is the first trial step while the memory
    holds no pair (default min(1, 1/|d|)).

    Returns a dict: x, fun, jac, nit, nfev, success, message, n_learned
    (accepted learned steps), n_fallback (learned proposals rejected)."""
    if x0.dtype != torch.float64 or not x0.is_cuda:
        raise TypeError("x0 must be a float64 device tensor")
    if learned_step not in ("lbfgs", "own"):
        raise ValueError(f"learned_step={learned_step!r}; use 'lbfgs' or 'own'")
    x = x0.clone()
    if lo is not None or hi is not None:
        x = step_clip(x, torch.zeros_like(x), 0.0, lo, hi)
    mem = memory or LBFGSMemory(x.numel(), m, x.device)
    if fg0 is not None:
        f, g = fg0
        nfev = 0
    else:
        f, g = fun(x)
        nfev = 1
    f = float(f)
    nit = 0
    t_cap = float(max_step) if max_step is not None else 1e20
    d = torch.empty_like(x)
    msg, ok = "maximum number of iterations reached", True
    bounded = lo is not None or hi is not None
    n_learned = n_fallback = 0
    while nit < maxiter:
        if bounded:
            # quasi-Newton step in the free variables: the binding set is
            # masked out of the gradient and of the direction
            gf = bound_mask(x, g, lo, hi, g.clone())
            d, _ = mem.direction(gf, out=d)
            bound_mask(x, g, lo, hi, d)
            gd = dot(g, d)
        else:
            d, gd = mem.direction(g, out=d)
        if not gd < 0:
            # not a descent direction: restart from (projected) steepest descent
            mem.reset()
            d, gd = mem.direction(gf if bounded else g, out=d)
            if bounded:
                bound_mask(x, g, lo, hi, d)
                gd = dot(g, d)
            if not gd < 0:
                msg, ok = "projected gradient is zero", True
                break
        t0 = 1.0
        if nit == 0 and len(mem) == 0:
            dn = math.sqrt(dot(d, d))
            t0 = t_init if t_init is not None else (min(1.0, 1.0 / dn) if dn > 0 else 1.0)
        t0 = min(t0, t_cap)
        res, learned = None, False
        if direction_fn is not None:
            dl = direction_fn(x, g)
            if dl is not None:
                dl = dl.to(torch.float64).contiguous()
                if bounded:
                    bound_mask(x, g, lo, hi, dl)
                amax = float(dl.abs().max())
                if dot(g, dl) < 0 and amax > 0:
                    ev = _Eval(fun, x, dl, lo, hi)
                    tl = 1.0 if learned_step == "own" else t0 * float(d.abs().max()) / amax
                    r = _armijo_projected(ev, f, g, x, tl, c1, maxls_learned)
                    nfev += ev.nfev
                    if r is not None and r[1] < f:
                        res, learned = r, True
                if learned:
                    n_learned += 1
                else:
                    n_fallback += 1
        ev = _Eval(fun, x, d, lo, hi)
        if res is not None:
            pass
        elif lo is None and hi is None:
            res = _wolfe(ev, f, gd, t0, c1, c2, maxls, t_max=t_cap)
        else:
            res = _armijo_projected(ev, f, g, x, t0, c1, maxls)
        nfev += ev.nfev
        if res is None:
            # accept the best trial if it decreased f, else stop
            if ev.best is not None and ev.best[1] < f:
                res = ev.best
            else:
                msg, ok = "line search found no decrease", False
                break
        _, f_new, g_new, x_new = res
        mem.push(x_new, x, g_new, g)
        if on_step is not None:
            on_step(x, g, x_new, g_new, learned)
        x, f, g = x_new, float(f_new), g_new
        nit += 1
        if callback is not None:
            callback(x)
    return {"x": x, "fun": f, "jac": g, "nit": nit, "nfev": nfev, "success": ok, "message": msg,
            "n_learned": n_learned, "n_fallback": n_fallback}
