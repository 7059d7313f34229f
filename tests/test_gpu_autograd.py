"""The forward/adjoint pair behind torch.autograd (stfwi_amd.autograd) and
checkpointed recomputation: gradients against central finite differences and
against compute_gradient('ad'), and independence of the checkpoint spacing."""
import numpy as np
import pytest
import torch

from tests.common import rel_l2

pytestmark = pytest.mark.gpu


def _model(nz=30, nx=44, seed=0):
    rng = np.random.default_rng(seed)
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    vp = 2400.0 + 300.0 * np.sin(zz / 7.0) * np.cos(xx / 9.0) + 20 * rng.standard_normal((nz, nx))
    vs = vp / 1.75
    rho = 2000.0 + 100.0 * np.cos(xx / 11.0)
    return vs, vp, rho


def _op(dtype, ckpt=0, channels=(1, 2), nt=150, src_type=("force", "moment"), src_z=(4, 6)):
    from stfwi_amd.autograd import SeismogramOperator
    from stfwi_amd.models import ricker
    nz, nx = 30, 44
    rx = np.arange(2, nx - 2, 3)
    rz = np.full_like(rx, 1)
    wav = ricker(nt, 1e-3, 25.0) * 1e6
    return SeismogramOperator(nz, nx, 10.0, 10.0, 1e-3, 25.0, 8, [12, 30], list(src_z), wav, rx, rz, nt,
                              src_type=list(src_type), fd_order=4, free_surface=True, dtype=dtype,
                              channels=channels, ckpt_every=ckpt)


def _tensors(dtype, grad=(True, True, True)):
    dev = torch.device("cuda", 0)
    return [torch.tensor(a, dtype=dtype, device=dev, requires_grad=r) for a, r in zip(_model(), grad)]


def test_gradient_matches_central_differences_f64():
    """<grad L, d> against (L(m + h d) - L(m - h d)) / 2h for a random linear
    functional of the seismograms, 16 random relative directions, h = 1e-5
    (the data's phase moves by about 45 rad per unit relative velocity here,
    so the central difference's O(h^2) error is ~3e-4 at h = 1e-3 and ~3e-8
    at h = 1e-5).
    The directions vanish where the reference does not differentiate: the
    edge columns and the bottom row, which the model embedding copies into
    the C-PML under stop_gradient (fwi/forward.py:199-219), and the cells
    near max(Vp), whose change would move the C-PML profile
    (fwi/forward.py:234-236)."""
    op = _op(torch.float64)
    m = _tensors(torch.float64)
    rng = np.random.default_rng(7)
    rec = op(*m)
    w = torch.as_tensor(rng.standard_normal(rec.shape), device=rec.device)
    (rec * w).sum().backward()
    g = [t.grad.clone() for t in m]
    base = [t.detach() for t in m]
    worst = 0.0
    for _ in range(16):
        u = rng.uniform(-1, 1, (3,) + tuple(base[0].shape))
        u[:, :, 0] = u[:, :, -1] = 0.0
        u[:, -1, :] = 0.0
        vp = base[1].cpu().numpy()
        u[1][vp > vp.max() * (1 - 3e-3)] = 0.0
        d = [torch.as_tensor(u[i], device=t.device) * t for i, t in enumerate(base)]
        h = 1e-5
        with torch.no_grad():
            lp = float((op(*[b + h * di for b, di in zip(base, d)]) * w).sum())
            lm = float((op(*[b - h * di for b, di in zip(base, d)]) * w).sum())
        fd = (lp - lm) / (2 * h)
        an = float(sum((gi * di).sum() for gi, di in zip(g, d)))
        worst = max(worst, abs(fd - an) / abs(an))
    assert worst < 1e-6, worst


def test_gradient_equals_compute_gradient_ad_f64():
    """A torch L2 misfit through the operator reproduces compute_gradient's
    exact ('ad') gradient for all three parameters."""
    from stfwi_amd import config
    from stfwi_amd.fwi import compute_gradient
    from stfwi_amd.models import ricker
    config.enable_x64(True)
    try:
        op = _op(torch.float64, channels=(2,), src_type=("force", "force"), src_z=(4, 4))
        vs, vp, rho = _model()
        rng = np.random.default_rng(3)
        with torch.no_grad():
            clean = op(*_tensors(torch.float64, (False, False, False)))[:, 0]
            obs = clean * (1 + 0.05 * torch.as_tensor(rng.standard_normal(tuple(clean.shape)), device="cuda"))
        scale = 2.0
        m = _tensors(torch.float64)
        pred = op(*m)[:, 0]
        loss = scale / 2 * ((pred - obs) ** 2).mean(dim=(1, 2)).sum()
        loss.backward()
        rx = np.arange(2, 44 - 2, 3)
        (gvs, gvp, grho), l_ref = compute_gradient(
            "ad", Vs=vs, Vp=vp, rho=rho, src_x=[12, 30], src_z=[4, 4], src_wavelet=ricker(150, 1e-3, 25.0) * 1e6,
            observed=obs.cpu().numpy(), rec_x=rx, rec_z=np.full_like(rx, 1), nx_dom=44, nz_dom=30, dx=10.0,
            dz=10.0, dt=1e-3, nt=150, fc=25.0, pad=8, component=2, invert="all", scale=scale, fd_order=4,
            free_surface=True)
        assert abs(float(loss.detach()) / l_ref - 1) < 1e-12
        for a, b in zip(m, (gvs, gvp, grho)):
            assert rel_l2(a.grad.cpu().numpy(), b) < 1e-12
    finally:
        config.enable_x64(False)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_checkpoint_spacing_does_not_change_gradients(dtype):
    grads = []
    for ck in (0, 7, 40, None):
        op = _op(dtype, ckpt=ck)
        m = _tensors(dtype)
        rec = op(*m)
        w = torch.linspace(-1, 1, rec.numel(), dtype=dtype, device=rec.device).reshape(rec.shape)
        (rec * w).sum().backward()
        grads.append((rec.detach(), [t.grad for t in m]))
    for rec, g in grads[1:]:
        assert torch.equal(rec, grads[0][0])
        for a, b in zip(g, grads[0][1]):
            assert torch.equal(a, b)


@pytest.mark.parametrize("imaging,grad", [("exact", "all"), ("zhou", "vs_vp")])
def test_checkpointing_with_uneven_record_lengths(imaging, grad):
    """Plan-level: shots of different nt, checkpoint spacings that do and do
    not divide them; gradient and source adjoint identical to full history."""
    from stfwi_amd import Plan
    from stfwi_amd.models import ricker
    vs, vp, rho = _model()
    rx = np.arange(1, 44, 2)
    rz = np.zeros_like(rx)
    out = []
    for ck in (0, 16, 33):
        p = Plan(30, 44, 10.0, 10.0, 1e-3, 25.0, 8, fd_order=8, free_surface=True, dtype=torch.float64,
                 n_shots_max=3, nt_max=140, nrec_max=rx.size, imaging=imaging, grad=grad, store_history=True,
                 ckpt_every=ck)
        p.set_model(vs, vp, rho)
        p.set_shots([5, 20, 38], [3, 10, 6], ["moment", "force", "moment"], [140, 97, 121], rx, rz,
                    ricker(140, 1e-3, 25.0) * 1e6)
        rec = p.forward()["rec"]
        radj = torch.flip(rec, dims=(2,)).contiguous()
        sa = p.adjoint(radj)["src_adj"]
        out.append((rec, sa, p.finalize(vs, vp, rho)))
    for r, s, g in out[1:]:
        assert torch.equal(r, out[0][0]) and torch.equal(s, out[0][1]) and torch.equal(g, out[0][2])


def test_second_forward_before_backward_raises():
    op = _op(torch.float32)
    m1 = _tensors(torch.float32)
    r1 = op(*m1)
    with torch.no_grad():
        op(*_tensors(torch.float32, (False, False, False)))
    m2 = _tensors(torch.float32)
    op(*m2)
    with pytest.raises(RuntimeError, match="another forward"):
        r1.sum().backward()
