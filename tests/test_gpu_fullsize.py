"""Size-independent properties of the HIP path at the full gradient workload:
152 x 500 cells, pad 20, dx = dz = 20 m, dt = 1.5 ms, nt = 3000, fd_order 8,
a free surface, 25 moment sources at depth 2 and 500 receivers at the
surface.  The CPU oracle is too slow at this size, so these tests check what
holds at any size: the adjoint is the transpose of the forward, one shot's
results do not depend on the other shots of the launch, and the FP32 gradient
stays close to the FP64 one."""
import numpy as np
import pytest
import torch

from tests.common import TOL_DOT, TOL_F32, rel_l2

pytestmark = pytest.mark.gpu

NZ, NX, PAD, FD, NT, DT, FC = 152, 500, 20, 8, 3000, 1.5e-3, 8.0
DX = DZ = 20.0
NSHOT, NREC = 25, 500


def _survey():
    from stfwi_amd.models import layered_model, ricker, smooth_initial
    true = layered_model(NZ, NX, seed=0)
    init = smooth_initial(*true, 10.0)
    src_x = (10 + 20 * np.arange(NSHOT)) % NX
    return dict(true=true, init=init, src_x=src_x, src_z=np.full(NSHOT, 2), rec_x=np.arange(NREC),
                rec_z=np.zeros(NREC, int), wav=ricker(NT, DT, FC))


def _plan(dtype, nshots, imaging="none", hist=False):
    from stfwi_amd import Plan
    return Plan(NZ, NX, DX, DZ, DT, FC, PAD, fd_order=FD, free_surface=True, dtype=dtype, n_shots_max=nshots,
                nt_max=NT, nrec_max=NREC, imaging=imaging, grad="vs", store_history=hist)


def test_gpu_fullsize_dot_product_f64():
    """<F s, y> = <s, F^T y> for one shot at the full size, in FP64, over
    all four recorded channels (test_adjoint_kernel.py:100-108)."""
    sv = _survey()
    rng = np.random.default_rng(5)
    s = rng.standard_normal(NT)
    y = rng.standard_normal((4, NT, NREC))
    print("case full-size dot product, shot 12", flush=True)
    p = _plan(torch.float64, 1)
    p.set_model(*sv["true"])
    p.set_shots([sv["src_x"][12]], [2], "moment", NT, sv["rec_x"], sv["rec_z"], s)
    rec = p.forward()["rec"][0].cpu().numpy()
    src_a = p.adjoint(torch.as_tensor(y[None]))["src_adj"][0].cpu().numpy()
    lhs, rhs = float(np.sum(rec * y)), float(np.sum(s * src_a))
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"dot product: lhs {lhs:.17e} rhs {rhs:.17e} rel {err:.3e}", flush=True)
    assert err < TOL_DOT


def test_gpu_fullsize_batched_shots_equal_single_f32():
    """Each shot of the 25-shot launch, run in the default 3 stream groups,
    records the same seismograms as the shot run alone (bitwise)."""
    sv = _survey()
    print("case full-size 25-shot forward", flush=True)
    pb = _plan(torch.float32, NSHOT)
    pb.set_model(*sv["true"])
    pb.set_shots(sv["src_x"], sv["src_z"], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
    rb = pb.forward()["rec"]
    assert torch.isfinite(rb).all()
    del pb
    for i in (0, 12, 24):
        print(f"case full-size single shot {i}", flush=True)
        p1 = _plan(torch.float32, 1)
        p1.set_model(*sv["true"])
        p1.set_shots([sv["src_x"][i]], [2], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
        assert torch.equal(p1.forward()["rec"][0], rb[i])


def _gradient(dtype, sv, obs, ishot, scale):
    p = _plan(dtype, 1, imaging="exact", hist=True)
    p.set_model(*sv["init"])
    p.set_shots([sv["src_x"][ishot]], [2], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
    rec = p.forward(channels=(2,))["rec"]
    loss, radj = p.residual(rec, obs.to(dtype), 2, None, 0, scale, 1)
    p.adjoint(radj, want_src=False)
    g = p.finalize(*sv["init"])[0].double().cpu().numpy()
    return float(loss[0]), g


def test_gpu_fullsize_f32_gradient_close_to_f64():
    """The FP32 Vs gradient of one shot at the full size, from the smoothed
    initial model against data of the layered model, agrees with the FP64
    gradient of the same kernels (misfit and gradient, relative L2).  The
    misfit is normalised by the mean square of the data, as in the
    benchmark: unscaled, a unit wavelet gives gradients near 1e-37, below
    the smallest normal FP32 value."""
    sv = _survey()
    print("case full-size observed data, FP64", flush=True)
    po = _plan(torch.float64, 1)
    po.set_model(*sv["true"])
    po.set_shots([sv["src_x"][12]], [2], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
    obs = po.forward(channels=(2,))["rec"][:, 2].contiguous()
    del po
    scale = float(1.0 / torch.mean(obs ** 2))
    print(f"case full-size gradient, FP64 and FP32, misfit scale {scale:.3e}", flush=True)
    l64, g64 = _gradient(torch.float64, sv, obs, 12, scale)
    l32, g32 = _gradient(torch.float32, sv, obs, 12, scale)
    el, eg = abs(l32 / l64 - 1), rel_l2(g32, g64)
    print(f"loss rel {el:.3e}  gradient rel-L2 {eg:.3e}", flush=True)
    assert np.isfinite(g64).all() and np.abs(g64).max() > 1e-20
    assert el < TOL_F32
    assert eg < TOL_F32


def test_gpu_fullsize_f64_gradient_matches_finite_difference():
    """The FP64 exact Vs gradient at the full size against a central
    difference of the misfit along a random direction d:
    (L(Vs + h d) - L(Vs - h d)) / 2h = <g, d> + O(h^2).  With h = 0.025 m/s
    (a 2e-5 relative change of a 1500 m/s model) the truncation term is far
    below the bar, and the rounding of the misfit (about 1e-16 L / h) too.
    The gradient is taken with the C-PML copies of the edge cells held fixed
    (stop_gradient on the embedding, fwi/forward.py:214-219), while
    perturbing a model moves those copies too, so d is zero on the cells the
    embedding copies: the first and last columns and, under a free surface,
    the bottom row."""
    sv = _survey()
    print("case full-size finite-difference check, FP64", flush=True)
    po = _plan(torch.float64, 1)
    po.set_model(*sv["true"])
    po.set_shots([sv["src_x"][12]], [2], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
    obs = po.forward(channels=(2,))["rec"][:, 2].contiguous()
    del po
    scale = float(1.0 / torch.mean(obs ** 2))
    l0, g = _gradient(torch.float64, sv, obs, 12, scale)
    vs0, vp0, rho0 = sv["init"]
    d = np.random.default_rng(9).standard_normal(np.shape(vs0))
    d[:, 0] = d[:, -1] = d[-1, :] = 0.0

    def misfit(vs):
        p = _plan(torch.float64, 1)
        p.set_model(vs, vp0, rho0)
        p.set_shots([sv["src_x"][12]], [2], "moment", NT, sv["rec_x"], sv["rec_z"], sv["wav"])
        loss, _ = p.residual(p.forward(channels=(2,))["rec"], obs, 2, None, 0, scale, 1)
        return float(loss[0])

    h = 0.025
    fd = (misfit(vs0 + h * d) - misfit(vs0 - h * d)) / (2 * h)
    an = float(np.sum(g * d))
    err = abs(fd / an - 1)
    print(f"misfit {l0:.6e}  finite difference {fd:.12e}  <g, d> {an:.12e}  rel {err:.3e}", flush=True)
    assert err < 1e-6
