"""Drop-in for fwi/filters.py on the HIP device (time axis first).

    jax_biquad(x, b0, b1, b2, a1, a2)   direct-form-I biquad, zero state   fwi/filters.py:14-23
    cosine_taper_end(x, n_taper)        (cos(pi k/n) + 1)/2 on the last n   fwi/filters.py:26-32
    jax_sosfilt(x, sos)                 SOS cascade, normalised by a0       fwi/filters.py:35-42
    sosfilt_T(x, sos, n_taper=0)        exact transpose of taper-then-filter
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _abi
from ._common import as_dev, compute_dtype


def _run(x, sos, n_taper, transpose):
    t = x if isinstance(x, torch.Tensor) and x.is_cuda and x.dtype in (torch.float32, torch.float64) \
        else as_dev(x, compute_dtype())
    out = t.clone().contiguous()
    nt = out.shape[0]
    ntr = int(out.numel() // max(nt, 1))
    if sos is not None:
        s = np.ascontiguousarray(np.asarray(sos, np.float64).reshape(-1, 6))
        sp, ns = s.ctypes.data_as(C.c_void_p), s.shape[0]
    else:
        s, sp, ns = None, None, 0
    code = _abi.F64 if out.dtype == torch.float64 else _abi.F32
    _abi.check(_abi.lib().stfwi_sosfilt(C.c_void_p(out.data_ptr()), nt, ntr, sp, ns, int(n_taper), int(transpose),
                                        code, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out


def jax_biquad(x, b0, b1, b2, a1, a2):
    return _run(x, [[b0, b1, b2, 1.0, a1, a2]], 0, 0)


def cosine_taper_end(x, n_taper):
    if n_taper <= 0:
        return x if isinstance(x, torch.Tensor) else as_dev(x, compute_dtype())
    return _run(x, None, n_taper, 0)


def jax_sosfilt(x, sos):
    return _run(x, sos, 0, 0)


def sosfilt_T(x, sos, n_taper=0):
    """Adjoint of x -> jax_sosfilt(cosine_taper_end(x, n_taper), sos)."""
    return _run(x, sos, n_taper, 1)
