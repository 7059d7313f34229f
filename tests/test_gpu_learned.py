"""Learned preconditioner kernels: BF16-MFMA convolutions (forward, data and
weight gradients) and the FP32 layer kernels against torch on the CPU, the
whole U-Net against a CPU autograd reference, and training."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def _bf(t):
    """Round to BF16 (nearest even) and back, on the CPU."""
    return t.to(torch.bfloat16).to(torch.float64)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("k,B,cin,cout,h,w", [(3, 2, 5, 19, 13, 21), (3, 1, 16, 32, 24, 40), (1, 2, 16, 1, 10, 14),
                                              (3, 1, 48, 16, 8, 6), (3, 2, 70, 40, 9, 70), (1, 1, 80, 33, 5, 130),
                                              (3, 2, 192, 64, 6, 40), (3, 1, 128, 130, 5, 70)])
def test_conv_forward_and_gradients_match_cpu(k, B, cin, cout, h, w):
    """Each kernel against the same convolution evaluated in float64 on the
    BF16-rounded operands it consumes: only the FP32 accumulation differs.
    Inputs wider than 32 channels run as several passes that add into the
    output (48, 70, 80, 128, 192 channels; 33, 40, 64, 130 in the data
    gradient)."""
    from stfwi_amd import learned as L
    g = torch.Generator().manual_seed(k * 100 + cin)
    x = torch.randn((B, cin, h, w), generator=g)
    wt = torch.randn((cout, cin, k, k), generator=g) * 0.3
    b = torch.randn(cout, generator=g)
    gy = torch.randn((B, cout, h, w), generator=g)
    dev = torch.device("cuda", 0)
    y = L.conv(x.to(dev), wt.to(dev), b.to(dev), relu=False).cpu()
    ref = F.conv2d(_bf(x), _bf(wt), b.double(), padding=k // 2)
    assert _rel(y, ref) < 1e-5
    yr = L.conv(x.to(dev), wt.to(dev), b.to(dev), relu=True).cpu()
    assert torch.equal(yr, torch.clamp(y, min=0))
    gx = L.conv_dgrad(gy.to(dev), wt.to(dev)).cpu()
    gx_ref = torch.nn.grad.conv2d_input(x.shape, _bf(wt), _bf(gy), padding=k // 2)
    assert _rel(gx, gx_ref) < 1e-5
    dw, db = L.conv_wgrad(x.to(dev), gy.to(dev), wt.shape)
    dw_ref = torch.nn.grad.conv2d_weight(_bf(x), wt.shape, _bf(gy), padding=k // 2)
    assert _rel(dw.cpu(), dw_ref) < 1e-5
    assert _rel(db.cpu(), gy.double().sum(dim=(0, 2, 3))) < 1e-6


def test_layer_kernels_match_cpu():
    from stfwi_amd import learned as L
    dev = torch.device("cuda", 0)
    g = torch.Generator().manual_seed(1)
    x = torch.randn((2, 3, 10, 14), generator=g)
    p = L.pool2(x.to(dev)).cpu()
    assert torch.allclose(p, F.avg_pool2d(x, 2), rtol=1e-6, atol=1e-7)
    gp = torch.randn(p.shape, generator=g)
    xr = x.clone().requires_grad_(True)
    F.avg_pool2d(xr, 2).backward(gp)
    assert torch.allclose(L.pool2_bwd(gp.to(dev), x.shape).cpu(), xr.grad, rtol=1e-6, atol=1e-7)
    big = torch.zeros((2, 5, 20, 28), device=dev)
    L.up2_into(x.to(dev), big, 2)
    assert torch.equal(big[:, 2:].cpu(), F.interpolate(x, scale_factor=2, mode="nearest"))
    assert not torch.any(big[:, :2])
    gb = torch.randn((2, 5, 20, 28), generator=g)
    xr = x.clone().requires_grad_(True)
    F.interpolate(xr, scale_factor=2, mode="nearest").backward(gb[:, 2:])
    assert torch.allclose(L.up2_bwd_from(gb.to(dev), 2, 3, 10, 14).cpu(), xr.grad, rtol=1e-6, atol=1e-6)
    y = torch.randn(1000, generator=g)
    gg = torch.randn(1000, generator=g)
    out = L.relu_bwd_(gg.clone().to(dev), y.to(dev)).cpu()
    assert torch.equal(out, torch.where(y > 0, gg, torch.zeros_like(gg)))
    t = torch.randn(1000, generator=g)
    wgt = (torch.rand(1000, generator=g) > 0.3).float()
    loss, gl = L.mse(y.to(dev), t.to(dev), wgt.to(dev))
    assert abs(loss - float((wgt * (y - t) ** 2).sum() / wgt.sum())) < 1e-5 * loss
    assert torch.allclose(gl.cpu(), 2 * wgt * (y - t) / wgt.sum(), rtol=1e-6, atol=1e-8)


def test_adam_matches_closed_form():
    from stfwi_amd import _abi
    from stfwi_amd.learned import _p, _s
    dev = torch.device("cuda", 0)
    g = torch.Generator().manual_seed(2)
    p0 = torch.randn(500, generator=g)
    grads = [torch.randn(500, generator=g) for _ in range(3)]
    p, m, v = p0.clone().to(dev), torch.zeros(500, device=dev), torch.zeros(500, device=dev)
    pr, mr, vr = p0.double().clone(), torch.zeros(500, dtype=torch.float64), torch.zeros(500, dtype=torch.float64)
    for t, gr in enumerate(grads, 1):
        _abi.check(_abi.lib().stfwi_adam(_p(p), _p(gr.to(dev)), _p(m), _p(v), 500, 1e-2, 0.9, 0.999, 1e-8, t, _s()))
        mr = 0.9 * mr + 0.1 * gr.double()
        vr = 0.999 * vr + 0.001 * gr.double() ** 2
        pr = pr - 1e-2 * (mr / (1 - 0.9 ** t)) / (torch.sqrt(vr / (1 - 0.999 ** t)) + 1e-8)
    assert torch.allclose(p.cpu().double(), pr, rtol=1e-5, atol=1e-6)


class _BfConv(torch.autograd.Function):
    """CPU reference of the kernels' convolution: BF16-rounded operands,
    float64 arithmetic, in the forward and in both gradients."""

    @staticmethod
    def forward(ctx, x, w, b, k):
        ctx.save_for_backward(x, w)
        ctx.k = k
        return F.conv2d(_bf(x), _bf(w), b, padding=k // 2)

    @staticmethod
    def backward(ctx, gy):
        x, w = ctx.saved_tensors
        k = ctx.k
        gx = torch.nn.grad.conv2d_input(x.shape, _bf(w), _bf(gy), padding=k // 2)
        gw = torch.nn.grad.conv2d_weight(_bf(x), w.shape, _bf(gy), padding=k // 2)
        return gx, gw, gy.sum(dim=(0, 2, 3)), None


def _cpu_unet(params, x, channels):
    c = lambda t, n, relu=True: (torch.relu if relu else (lambda z: z))(
        _BfConv.apply(t, params[n][0], params[n][1], params[n][0].shape[-1]))
    L = len(channels)
    skips, h = [], x
    for l in range(L):
        e = c(c(h, f"enc{l}a"), f"enc{l}b")
        skips.append(e)
        if l < L - 1:
            h = F.avg_pool2d(e, 2)
    d = skips[-1]
    for l in range(L - 2, -1, -1):
        cat = torch.cat([skips[l], F.interpolate(d, scale_factor=2, mode="nearest")], dim=1)
        d = c(c(cat, f"dec{l}a"), f"dec{l}b")
    return c(d, "out", relu=False)


@pytest.mark.parametrize("channels", [(16, 32, 64, 128), (8, 16)])
def test_unet_forward_backward_match_cpu_reference(channels):
    """The 4-level 16/32/64/128 U-Net with 3 inputs (and a 2-level one)
    against a CPU autograd U-Net of the same BF16 rounding: outputs, loss
    and every parameter gradient."""
    from stfwi_amd.learned import UNet, mse
    net = UNet(3, 1, channels, seed=3)
    g = torch.Generator().manual_seed(4)
    x = torch.randn((2, 3, 24, 40), generator=g)
    t = torch.randn((2, 1, 24, 40), generator=g)
    y = net.forward(x.cuda())
    loss, gy = mse(y, t.cuda())
    grads = net.backward(gy)
    assert set(grads) == set(net.params) and len(net.params) == 4 * len(channels) - 1
    params = {n: [ps[0].cpu().double().requires_grad_(True), ps[1].cpu().double().requires_grad_(True)]
              for n, ps in net.params.items()}
    yr = _cpu_unet(params, x.double(), channels)
    lr = ((yr - t.double()) ** 2).mean()
    lr.backward()
    assert _rel(y.cpu(), yr.detach()) < 1e-3
    assert abs(loss / float(lr.detach()) - 1) < 1e-3
    for n in params:
        for i in range(2):
            assert _rel(grads[n][i].cpu(), params[n][i].grad) < 1e-2, (n, i)


def test_training_reduces_loss_and_preconditioner_round_trip():
    """The preconditioner learns a fixed smoothing-and-weighting map of its
    inputs from pairs; the loss drops tenfold."""
    from stfwi_amd.learned import LearnedPreconditioner
    nz, nx = 30, 50
    lp = LearnedPreconditioner(nz, nx, seed=1)
    assert (lp.H, lp.W) == (32, 56)          # padded to multiples of 8
    rng = np.random.default_rng(0)
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    vs = 1500.0 + 20.0 * zz
    for i in range(8):
        gr = rng.standard_normal((nz, nx))
        il = 1.0 / (1.0 + 0.1 * zz + 0.01 * xx * (i + 1))
        step = -(gr * il)
        lp.add_pair(gr, il, vs, step)
    hist = lp.train(epochs=150, lr=3e-3, batch=4)
    assert hist[-1] < 0.1 * hist[0], (hist[0], hist[-1])
    d = lp.direction(rng.standard_normal((nz, nx)), il, vs)
    assert d.shape == (nz, nx) and d.dtype == torch.float64 and torch.isfinite(d).all()
    lp2 = LearnedPreconditioner(nz, nx, seed=9)
    lp2.net.load_state_dict(lp.net.state_dict())
    gtest = rng.standard_normal((nz, nx))
    assert torch.equal(lp.direction(gtest, il, vs), lp2.direction(gtest, il, vs))


def test_learned_vs_plain_evaluation_runs_end_to_end():
    """The learned-vs-plain evaluation path on the 152x500, nt=3000 survey of
    one held-out seed, cut to 2 + 1 + 1 iterations: the plain run, then the
    learned run stopped at the plain final RMSE, with the trained network
    when it is in the tree (stfwi_amd/weights), else an untrained one."""
    import os
    from stfwi_amd import learned_eval as E
    from stfwi_amd.learned import LearnedPreconditioner
    w = os.path.join(os.path.dirname(E.__file__), "weights", "final.pt")
    lp = LearnedPreconditioner.load(w) if os.path.exists(w) else LearnedPreconditioner(152, 500, seed=0)
    r = E.evaluate_seed(2000, {"final": lp}, plain_iters=(2, 1, 1))
    assert r["plain_iters"] == 4 and r["plain_s"] > 0 and np.isfinite(r["plain_rmse_final"])
    L = r["learned"]["final"]
    assert L["iters"] <= 4 and L["learned_steps"] <= L["iters"] <= L["learned_steps"] + L["fallbacks"]
    assert np.isfinite(L["rmse_final"])
    if L["reached"]:
        assert L["rmse_final"] <= r["plain_rmse_final"] + 1e-9 and 0 <= L["time_ratio"]


def test_trained_network_reaches_plain_rmse_with_fewer_evaluations():
    """What the shipped network gains over the method without it, on one
    layered model outside its training seeds (seed 3001, 25 shots, 152x500,
    nt = 3000): the plain run (bands 3/5/8 Hz x 20 L-BFGS iterations) ends
    at some Vs RMSE; the run with the network's update tried at the start of
    each band (final:stage:2) reaches that RMSE with fewer misfit-and-
    gradient evaluations.  Both runs are deterministic.  This is one model:
    over the held-out seeds 2000-2015 the gain is not reliable, which
    `bench.py --learned-vs-plain` measures."""
    import os
    from stfwi_amd import learned_eval as E
    wdir = os.path.join(os.path.dirname(E.__file__), "weights")
    assert os.path.exists(os.path.join(wdir, "final.pt")), "the trained network ships in stfwi_amd/weights"
    lp = E.variant(wdir, "final:stage:2")
    r = E.evaluate_seed(3001, {"v": lp})
    L = r["learned"]["v"]
    print(f"plain: RMSE {r['rmse_initial']:.2f} -> {r['plain_rmse_final']:.2f} m/s, {r['plain_evals']} evaluations, "
          f"{r['plain_s']:.1f} s; learned: reached {L['reached']} after {L['evals']} evaluations, "
          f"{L['run_s']:.1f} s, {L['learned_steps']} learned steps")
    assert r["plain_improves"] and L["reached"]
    assert L["learned_steps"] >= 1
    assert L["evals"] < r["plain_evals"]
