"""The allocation-free eprecond1 entry point against the allocating one,
through the C-ABI: bit-identical weights on two calls in a row, the second
with a smaller input whose maximum is smaller than the first's (a maximum
left in the shared scratch by the first call would show)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(lib, fn, ws, wr, eps, code):
    we = torch.empty_like(ws)
    rc = fn(C.c_void_p(ws.data_ptr()), C.c_void_p(wr.data_ptr()), ws.numel(), eps, code, C.c_void_p(we.data_ptr()),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, lib.stfwi_last_error().decode()
    return we


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_eprecond1_noalloc_matches_allocating_entry(dtype):
    from stfwi_amd import _abi
    lib = _abi.lib()
    code = _abi.F64 if dtype == torch.float64 else _abi.F32
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    eps = 1e-3
    for n, amp in ((3_000_017, 1000.0), (70_001, 1.0)):
        ws = torch.as_tensor(rng.random(n) * amp, dtype=dtype, device=dev)
        wr = torch.as_tensor(rng.random(n) * amp, dtype=dtype, device=dev)
        we_a = _run(lib, lib.stfwi_eprecond1, ws, wr, eps, code)
        we_b = _run(lib, lib.stfwi_eprecond1_noalloc, ws, wr, eps, code)
        torch.cuda.synchronize()
        assert torch.equal(we_a, we_b), n
        sq = np.sqrt(ws.double().cpu().numpy() * wr.double().cpu().numpy())
        ref = sq + eps * sq.max()
        tol = 1e-14 if dtype == torch.float64 else 1e-6
        assert np.max(np.abs(we_b.double().cpu().numpy() - ref) / ref) < tol, n


def test_eprecond1_noalloc_rejects_bad_arguments():
    from stfwi_amd import _abi
    lib = _abi.lib()
    x = torch.ones(8, dtype=torch.float64, device=torch.device("cuda", 0))
    p = C.c_void_p(x.data_ptr())
    assert lib.stfwi_eprecond1_noalloc(p, p, 0, 1e-3, _abi.F64, p, None) == -1
    assert lib.stfwi_eprecond1_noalloc(p, p, 8, 1e-3, 7, p, None) == -1
    assert lib.stfwi_eprecond1_noalloc(None, p, 8, 1e-3, _abi.F64, p, None) == -1
