"""Shared plumbing of the drop-in modules: tensor conversion, dtype policy,
plan cache."""
from __future__ import annotations

from collections import OrderedDict

import numpy as np
import torch

from .. import config
from ..plan import Plan

_CACHE: "OrderedDict[tuple, Plan]" = OrderedDict()
_CACHE_MAX = 3


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("stfwi_amd needs a HIP device (no CPU path)")
    return torch.device("cuda", torch.cuda.current_device())


def compute_dtype(dtype=None):
    return dtype if dtype is not None else config.default_dtype()


def as_dev(x, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device()).contiguous()


def as_int_array(x):
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.atleast_1d(np.asarray(x)).astype(np.int32)


def as_int(x):
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu()
    return int(np.asarray(x))


def storage_dtype(x):
    """Floating dtype the caller stored an array in (numpy or torch)."""
    if isinstance(x, torch.Tensor):
        return x.dtype if x.dtype in (torch.float32, torch.float64) else torch.float64
    a = np.asarray(x)
    return torch.float32 if a.dtype == np.float32 else torch.float64


def forward_fd_order(fd_order):
    """forward_jax accepts any fd_order: 8 is 8th order, >= 4 is 4th, anything
    else 2nd (fwi/forward.py:177-182)."""
    return 8 if fd_order == 8 else (4 if fd_order >= 4 else 2)


def get_plan(**kw):
    """A cached Plan with exactly these settings (n_shots_max, nt_max and
    nrec_max are capacities: a cached plan at least this large is reused)."""
    key = tuple(sorted((k, v) for k, v in kw.items() if k not in ("n_shots_max", "nt_max", "nrec_max")))
    p = _CACHE.get(key)
    if p is not None and p.n_shots_max >= kw["n_shots_max"] and p.nt_max == kw["nt_max"] \
            and p.nrec_max >= kw["nrec_max"]:
        _CACHE.move_to_end(key)
        return p
    if p is not None:
        del _CACHE[key]
    p = Plan(**kw)
    _CACHE[key] = p
    while len(_CACHE) > _CACHE_MAX:
        _CACHE.popitem(last=False)
    return p


# launch shapes: a plan runs its shots as up to 3 concurrent stream groups
# when a group still fills the chip (about 1200 tiles of 8 x 64 cells per
# launch), and picks its tile heights by timing them on its own shots once.
# Both choices change no result bit (the kernels are tested bitwise across
# groupings and tile heights); they only decide how fast a call runs.
_GROUP_TILES = 1200
_TUNE_CELLS = 1 << 20


def tune_launches(p, n_shots):
    """Stream groups and tile heights for plan p, once per plan, after its
    model and shots are set (small problems keep the defaults)."""
    if p._tuned:
        return
    p._tuned = True
    nz_t, nx_t = p.shape
    tiles = -(-nz_t // 8) * -(-nx_t // 64)
    groups = max(1, min(3, (n_shots * tiles) // _GROUP_TILES))
    if groups > 1:
        p.set_groups(groups)
    if n_shots * nz_t * nx_t >= _TUNE_CELLS:
        p.autotune(10)


def clear_plan_cache():
    _CACHE.clear()
