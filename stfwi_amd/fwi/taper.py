"""Drop-in for fwi/taper.py:16-84 on the HIP device.

    build_source_taper(nz, nx, dh, src_positions, srt_radius, filt_size=2) -> (nz, nx) float64

The plot helper of that file (save_source_taper_plot) is not provided:
plotting is outside this package.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _abi
from ._common import device


def build_source_taper(nz, nx, dh, src_positions, srt_radius, filt_size=2):
    """Gradient damping around sources: an erf radial patch of radius
    srt_radius (m) stamped at every (src_x, src_z) grid position (minimum over
    sources), normalised to [0, 1], with a (2 filt_size + 1)^2 zero box at each
    source."""
    pos = np.asarray([(int(x), int(z)) for x, z in src_positions], dtype=np.int32).reshape(-1, 2)
    out = torch.empty((int(nz), int(nx)), dtype=torch.float64, device=device())
    _abi.check(_abi.lib().stfwi_source_taper(int(nz), int(nx), float(dh), pos.ctypes.data_as(C.c_void_p), len(pos),
                                             float(srt_radius), int(filt_size), C.c_void_p(out.data_ptr()),
                                             C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out
