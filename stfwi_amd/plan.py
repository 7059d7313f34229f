"""Python handle on a libstfwi plan: device buffers are torch tensors, every
call runs on torch's current HIP stream.

A Plan batches many shots of one model through one time loop: a forward of
all shots, the residual prep, and an adjoint with the imaging condition
fused into each reverse step (the hot path of fwi/gradient.py:159-259).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _abi
from ._abi import check, lib


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _code(dtype):
    if dtype == torch.float32:
        return _abi.F32
    if dtype == torch.float64:
        return _abi.F64
    raise TypeError(f"unsupported dtype {dtype}")


def padded_shape(nz_dom, nx_dom, pad, free_surface):
    return nz_dom + (pad if free_surface else 2 * pad), nx_dom + 2 * pad


def footprint(cfg) -> int:
    """Device bytes a plan of this _abi.Config would allocate (no allocation)."""
    b = C.c_longlong()
    check(lib().stfwi_plan_footprint(C.byref(cfg), C.byref(b)))
    return int(b.value)


def choose_ckpt(cfg, budget_bytes) -> int:
    """Checkpoint spacing for a plan of this _abi.Config within budget_bytes:
    0 when the whole history fits, else the K of the smallest footprint
    (raises if none fits)."""
    k = C.c_int()
    check(lib().stfwi_plan_choose_ckpt(C.byref(cfg), int(budget_bytes), C.byref(k)))
    return int(k.value)


class Plan:
    """Device workspace for shots of one grid.

    Parameters mirror forward_jax / adjoint_jax (fwi/forward.py:119-123) plus
    batch sizing: n_shots_max, nt_max, nrec_max; ``imaging`` ('none',
    'zhou' = backend 'adjoint', 'exact' = backend 'ad'); ``grad`` ('vs',
    'vs_vp', 'all' = the reference's ``invert``); ``store_history`` keeps the
    strain history for the imaging adjoint; ``ckpt_every`` = K > 0 keeps
    only forward-state checkpoints every K steps plus one K-step segment of
    history and recomputes each segment during the adjoint (same results);
    ``ckpt_every='auto'`` lets the library's planner pick K from the free
    device memory (0 when the whole history fits).  ``hist_frames`` is the
    history capacity in frames summed over the shots (each shot keeps its
    own nt frames); 0 means n_shots_max * nt_max.
    """

    def __init__(self, nz_dom, nx_dom, dx, dz, dt, fc, pad, fd_order=2, free_surface=True,
                 dtype=torch.float32, n_shots_max=1, nt_max=1, nrec_max=1, imaging="none",
                 grad="vs", store_history=False, device=None, ckpt_every=0, hist_frames=0):
        if fd_order not in (2, 4, 8):
            raise NotImplementedError(f"fd_order={fd_order} not supported (use 2, 4, or 8).")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise RuntimeError("stfwi runs on a HIP device only")
        self.dtype = dtype
        self.nz_dom, self.nx_dom, self.pad = int(nz_dom), int(nx_dom), int(pad)
        self.free_surface = bool(free_surface)
        self.fd_order = int(fd_order)
        self.dx, self.dz, self.dt, self.fc = float(dx), float(dz), float(dt), float(fc)
        self.nt_max, self.nrec_max, self.n_shots_max = int(nt_max), int(nrec_max), int(n_shots_max)
        self.imaging = imaging
        self.grad = grad
        self.shape = padded_shape(self.nz_dom, self.nx_dom, self.pad, self.free_surface)
        cfg = _abi.Config(self.nz_dom, self.nx_dom, self.pad, int(self.free_surface), self.fd_order,
                          self.dx, self.dz, self.dt, self.fc, _code(dtype), self.n_shots_max,
                          self.nt_max, self.nrec_max, _abi.IMAGING[imaging], _abi.GRAD[grad],
                          int(bool(store_history)), 0, int(hist_frames))
        if ckpt_every == "auto":
            with torch.cuda.device(self.device):
                free, _ = torch.cuda.mem_get_info()
            ckpt_every = choose_ckpt(cfg, int(free * 0.9))
        self.ckpt_every = int(ckpt_every)
        cfg.ckpt_every = self.ckpt_every
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().stfwi_plan_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._keep = {}
        self.n_shots = 0
        self.nt = None
        self._tuned = False             # launch shape chosen (fwi._common.tune_launches)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().stfwi_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def nbytes(self):
        return int(lib().stfwi_plan_bytes(self._h))

    def _t(self, x, dtype=None):
        dtype = dtype or self.dtype
        if isinstance(x, torch.Tensor):
            return x.to(device=self.device, dtype=dtype).contiguous()
        return torch.as_tensor(np.asarray(x), dtype=dtype, device=self.device).contiguous()

    # ------------------------------------------------------------------
    def set_model(self, vs, vp, rho, b_img=None):
        vs, vp, rho = self._t(vs), self._t(vp), self._t(rho)
        for a in (vs, vp, rho):
            if tuple(a.shape) != (self.nz_dom, self.nx_dom):
                raise ValueError(f"model shape {tuple(a.shape)} != {(self.nz_dom, self.nx_dom)}")
        bi = None if b_img is None else self._t(b_img)
        self._keep["model"] = (vs, vp, rho, bi)
        check(lib().stfwi_set_model(self._h, _ptr(vs), _ptr(vp), _ptr(rho), _ptr(bi), _stream()))

    def set_shots(self, src_x: Sequence[int], src_z: Sequence[int], src_type, nt, rec_x, rec_z, wavelets):
        """rec_x / rec_z: one index array shared by all shots, or a list per shot.
        wavelets: (n_shots, nt_max) or (nt,) shared.  src_type / nt: scalar or per shot."""
        n = len(src_x)
        if n < 1 or n > self.n_shots_max:
            raise ValueError(f"{n} shots for a plan of {self.n_shots_max}")
        st = [src_type] * n if isinstance(src_type, (str, int)) else list(src_type)
        st = [_abi.SRC_FORCE if s in ("force", 0) else _abi.SRC_MOMENT if s in ("moment", 1) else s for s in st]
        for s in st:
            if s not in (0, 1):
                raise NotImplementedError(f"src_type={s!r} not supported (use 'force' or 'moment').")
        nts = [int(nt)] * n if np.ndim(nt) == 0 else [int(v) for v in nt]
        if isinstance(rec_x, (list, tuple)) and len(rec_x) == n and np.ndim(rec_x[0]) == 1:
            rxl, rzl = [np.asarray(r, np.int32) for r in rec_x], [np.asarray(r, np.int32) for r in rec_z]
        else:
            rxl = [np.asarray(rec_x, np.int32)] * n
            rzl = [np.asarray(rec_z, np.int32)] * n
        shots = (_abi.Shot * n)()
        off = 0
        for i in range(n):
            shots[i] = _abi.Shot(int(src_x[i]), int(src_z[i]), st[i], nts[i], off, len(rxl[i]))
            off += len(rxl[i])
        rx = np.ascontiguousarray(np.concatenate(rxl), np.int32)
        rz = np.ascontiguousarray(np.concatenate(rzl), np.int32)
        w = self._t(wavelets)
        if w.dim() == 1:
            w = w.unsqueeze(0).expand(n, -1)
        if w.shape[1] < self.nt_max:
            w = torch.nn.functional.pad(w, (0, self.nt_max - w.shape[1]))
        w = w[:, :self.nt_max].contiguous()
        self._keep["wav"] = w
        check(lib().stfwi_set_shots(self._h, shots, n, rx.ctypes.data_as(C.c_void_p),
                                    rz.ctypes.data_as(C.c_void_p), _ptr(w), _stream()))
        self.n_shots = n
        self.nts = nts
        self.nrecs = [len(r) for r in rxl]

    # ------------------------------------------------------------------
    def forward(self, channels=(0, 1, 2, 3), illum_comp=None, snap=()):
        """Returns dict: 'rec' (n_shots, 4, nt_max, nrec_max), optional
        'illum' (n_shots, nz_t, nx_t), optional 'snap'
        (n_shots, len(snap), nt_max, nz_t, nx_t)."""
        S = self.n_shots
        out = {}
        rec = None
        mask = 0
        if channels:
            rec = torch.zeros((S, 4, self.nt_max, self.nrec_max), dtype=self.dtype, device=self.device)
            for ch in channels:
                mask |= 1 << ch
            out["rec"] = rec
        il = None
        if illum_comp is not None:
            il = torch.zeros((S,) + self.shape, dtype=self.dtype, device=self.device)
            out["illum"] = il
        sn, smask = None, 0
        if snap:
            for k in snap:
                smask |= _abi.SNAP[k]
            order = [k for k in ("vx", "vz", "ex", "ez", "es") if _abi.SNAP[k] & smask]
            sn = torch.empty((S, len(order), self.nt_max) + self.shape, dtype=self.dtype, device=self.device)
            out["snap"] = sn
            out["snap_order"] = order
        check(lib().stfwi_forward(self._h, _ptr(rec), mask, _ptr(il), -1 if illum_comp is None else int(illum_comp),
                                  _ptr(sn), smask, _stream()))
        return out

    def residual(self, pred, obs, component=2, sos=None, n_taper=0, scale=1.0, n_norm=None):
        """Misfit per shot and receiver cotangent (fwi/gradient.py:139-156)."""
        S = self.n_shots
        pred = self._t(pred)
        obs = self._t(obs)
        if tuple(obs.shape) != (S, self.nt_max, self.nrec_max):
            raise ValueError(f"obs shape {tuple(obs.shape)} != {(S, self.nt_max, self.nrec_max)}")
        loss = torch.empty(S, dtype=self.dtype, device=self.device)
        radj = torch.empty((S, 4, self.nt_max, self.nrec_max), dtype=self.dtype, device=self.device)
        if sos is not None:
            sos_np = np.ascontiguousarray(np.asarray(sos, np.float64).reshape(-1, 6))
            nsos = sos_np.shape[0]
            sp = sos_np.ctypes.data_as(C.c_void_p)
        else:
            sos_np, nsos, sp = None, 0, None
        check(lib().stfwi_residual(self._h, _ptr(pred), _ptr(obs), int(component), sp, nsos, int(n_taper),
                                   float(scale), int(n_norm or S), _ptr(loss), _ptr(radj), _stream()))
        return loss, radj

    def adjoint(self, rec_adj, want_src=True, snap=(), energy=None):
        """Reverse time loop for receiver adjoints rec_adj (n_shots, 4,
        nt_max, nrec_max).  Returns 'src_adj' (n_shots, nt_max), optional
        'snap' as forward(), and with energy = 'vx' | 'vz' | 'ex' | 'ez' | 'es'
        'energy' (n_shots, nz_t, nx_t): the sum over steps of that adjoint
        field squared (the Wr of eprecond1)."""
        S = self.n_shots
        ra = self._t(rec_adj)
        if tuple(ra.shape) != (S, 4, self.nt_max, self.nrec_max):
            raise ValueError(f"rec_adj shape {tuple(ra.shape)} != {(S, 4, self.nt_max, self.nrec_max)}")
        out = {}
        sa = None
        if want_src:
            sa = torch.zeros((S, self.nt_max), dtype=self.dtype, device=self.device)
            out["src_adj"] = sa
        sn, smask = None, 0
        if snap:
            for k in snap:
                smask |= _abi.SNAP[k]
            order = [k for k in ("vx", "vz", "ex", "ez", "es") if _abi.SNAP[k] & smask]
            sn = torch.empty((S, len(order), self.nt_max) + self.shape, dtype=self.dtype, device=self.device)
            out["snap"] = sn
            out["snap_order"] = order
        self._keep["radj"] = ra
        en = None
        if energy is not None:
            comps = ("vx", "vz", "ex", "ez", "es")
            if energy not in comps:
                raise ValueError(f"energy={energy!r} (use one of {comps})")
            en = torch.zeros((S,) + self.shape, dtype=self.dtype, device=self.device)
            check(lib().stfwi_set_adjoint_energy(self._h, _ptr(en), comps.index(energy)))
            out["energy"] = en
        try:
            check(lib().stfwi_adjoint(self._h, _ptr(ra), _ptr(sa), _ptr(sn), smask, _stream()))
        finally:
            if en is not None:
                check(lib().stfwi_set_adjoint_energy(self._h, None, 0))
        return out

    def autotune(self, reps=20):
        """Pick the tile height of each time-step kernel by timing every
        compiled variant on this plan's shots (run before forward()).
        Returns (tile heights, {kernel: {tz: microseconds}})."""
        tz = (C.c_int * 4)()
        us = (C.c_double * 16)()
        check(lib().stfwi_autotune(self._h, int(reps), _stream(), tz, us))
        names = ("fwd_vel", "fwd_strain", "adj_strain", "adj_vel")
        table = {n: {h: us[4 * i + j] for j, h in enumerate((4, 8, 16, 32)) if us[4 * i + j] > 0}
                 for i, n in enumerate(names)}
        return list(tz), table

    def set_graphs(self, enable=True):
        """Graph replay of the time loops is refused by the library
        (NotImplementedError for enable=True); set_graphs(False) keeps the
        eager launches."""
        check(lib().stfwi_set_graphs(self._h, int(bool(enable))))

    def set_groups(self, groups):
        """Run the shots as `groups` concurrent stream groups (1..4)."""
        check(lib().stfwi_set_groups(self._h, int(groups)))

    def set_lean(self, enable=True):
        """Allow (the default) or forbid the time-step kernels' specialised
        instantiations for runs without optional outputs; the results are
        the same bit for bit either way."""
        check(lib().stfwi_set_lean(self._h, int(bool(enable))))

    def set_tiles(self, tz):
        arr = (C.c_int * 4)(*[int(t) for t in tz])
        check(lib().stfwi_set_tiles(self._h, arr))

    def time_kernel(self, which, reps=100, grouped=False):
        """Average time in microseconds of one time-step kernel over all
        shots (0 forward velocity, 1 forward strain, 2 adjoint strain,
        3 adjoint velocity), measured with HIP events on the current stream.
        grouped=True launches the plan's shot groups on their streams, as the
        time loops do, and returns the time per step of all groups.  The
        launches overwrite the wavefield state: run forward() again before
        the next adjoint() or finalize()."""
        ms = C.c_double()
        check(lib().stfwi_time_kernel(self._h, int(which), int(reps), int(bool(grouped)), _stream(),
                                      C.byref(ms)))
        return ms.value * 1e3 / reps

    def finalize(self, vs, vp, rho):
        """(3, nz_dom, nx_dom) gradients w.r.t. (Vs, Vp, rho); the chain
        factors are formed in the dtype of the given model tensors."""
        vs = vs if isinstance(vs, torch.Tensor) else torch.as_tensor(np.asarray(vs))
        cdt = vs.dtype if vs.dtype in (torch.float32, torch.float64) else torch.float64
        vs, vp, rho = (self._t(a, cdt) for a in (vs, vp, rho))
        g = torch.empty((3, self.nz_dom, self.nx_dom), dtype=self.dtype, device=self.device)
        check(lib().stfwi_gradient_finalize(self._h, _ptr(vs), _ptr(vp), _ptr(rho), _code(cdt), _ptr(g), _stream()))
        return g
