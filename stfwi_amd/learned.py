"""Learned gradient preconditioner for the plain L-BFGS method.

A small U-Net maps the raw Vs gradient and the illumination map of an FWI
iteration to a search direction.  It is trained on (gradient, illumination)
-> accepted model update pairs that the plain method produces, and the
driver falls back to the plain L-BFGS step whenever the learned direction
does not decrease the misfit.  The reference has no learned part; this module
has no reference counterpart.

Every convolution runs on the matrix cores through hand-written HIP kernels
(``stfwi_conv2d_bf16`` and its data/weight gradients: BF16 operands, FP32
accumulation); pooling, upsampling, the ReLU gradient, the MSE loss and Adam
are HIP kernels too.  Parameters and activations are FP32 device tensors;
torch only allocates and copies them.

Architecture (c1 = 16, c2 = 32 channels by default):
    x -> conv3 c1 -> conv3 c1 = e1 -> pool2 -> conv3 c2 -> conv3 c2 = e2
      -> up2, concat [e1, up(e2)] -> conv3 c1 -> conv3 c1 -> conv1 -> y
with ReLU after every 3x3 convolution.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np
import torch

from . import _abi

_VP = C.c_void_p


def _p(t):
    return _VP(t.data_ptr()) if t is not None else None


def _s():
    return _VP(torch.cuda.current_stream().cuda_stream)


def _lib():
    return _abi.lib()


# ---------------------------------------------------------------- layers --

def conv(x, w, b=None, relu=False):
    B, cin, h, wd = x.shape
    cout, _, k, _ = w.shape
    out = torch.empty((B, cout, h, wd), dtype=torch.float32, device=x.device)
    _abi.check(_lib().stfwi_conv2d_bf16(_p(x), B, cin, h, wd, _p(w), _p(b), cout, k, int(relu), _p(out), _s()))
    return out


def conv_dgrad(gy, w):
    B, cout, h, wd = gy.shape
    _, cin, k, _ = w.shape
    gx = torch.empty((B, cin, h, wd), dtype=torch.float32, device=gy.device)
    _abi.check(_lib().stfwi_conv2d_bf16_dgrad(_p(gy), B, cout, h, wd, _p(w), cin, k, _p(gx), _s()))
    return gx


_work = {}


def conv_wgrad(x, gy, w_shape, want_bias=True):
    B, cin, h, wd = x.shape
    cout, _, k, _ = w_shape
    n = int(_lib().stfwi_conv2d_bf16_wgrad_work(cin, cout, k))
    key = x.device
    if key not in _work or _work[key].numel() < n:
        _work[key] = torch.empty(n, dtype=torch.float32, device=x.device)
    dw = torch.empty(w_shape, dtype=torch.float32, device=x.device)
    db = torch.empty(cout, dtype=torch.float32, device=x.device) if want_bias else None
    _abi.check(_lib().stfwi_conv2d_bf16_wgrad(_p(x), _p(gy), B, cin, cout, h, wd, k, _p(dw), _p(db), 0,
                                              _p(_work[key]), _work[key].numel(), _s()))
    return dw, db


def relu_bwd_(g, y):
    _abi.check(_lib().stfwi_relu_bwd(_p(g), _p(y), g.numel(), _s()))
    return g


def pool2(x):
    B, c, h, w = x.shape
    out = torch.empty((B, c, h // 2, w // 2), dtype=torch.float32, device=x.device)
    _abi.check(_lib().stfwi_pool2(_p(x), B * c, h, w, _p(out), 0, _s()))
    return out


def pool2_bwd(gy, shape):
    B, c, h, w = shape
    gx = torch.empty(shape, dtype=torch.float32, device=gy.device)
    _abi.check(_lib().stfwi_pool2(_p(gy), B * c, h, w, _p(gx), 1, _s()))
    return gx


def up2_into(x, big, c0):
    """Nearest x2 upsampling of x (B, c, h, w) into channels [c0, c0 + c) of
    big (B, C, 2h, 2w)."""
    B, c, h, w = x.shape
    P = 4 * h * w
    dst = big[:, c0:]
    _abi.check(_lib().stfwi_up2(_p(x), B, c, h, w, _p(dst), P, c * h * w, big.shape[1] * P, 0, _s()))


def up2_bwd_from(gbig, c0, c, h, w):
    B = gbig.shape[0]
    P = 4 * h * w
    gx = torch.empty((B, c, h, w), dtype=torch.float32, device=gbig.device)
    _abi.check(_lib().stfwi_up2(_p(gbig[:, c0:]), B, c, h, w, _p(gx), P, c * h * w, gbig.shape[1] * P, 1, _s()))
    return gx


def mse(y, t, wgt=None):
    """(loss, dloss/dy) of sum w (y - t)^2 / sum w."""
    n = y.numel()
    wsum = float(wgt.sum()) if wgt is not None else float(n)
    g = torch.empty_like(y)
    part = torch.empty(256, dtype=torch.float32, device=y.device)
    loss = torch.empty(1, dtype=torch.float32, device=y.device)
    _abi.check(_lib().stfwi_mse(_p(y), _p(t), _p(wgt), n, 1.0 / wsum, _p(g), _p(part), _p(loss), _s()))
    return float(loss.item()) / wsum, g


# ------------------------------------------------------------------ U-Net --

class UNet:
    SPECS = (("e1a", "in", "c1", 3), ("e1b", "c1", "c1", 3), ("e2a", "c1", "c2", 3), ("e2b", "c2", "c2", 3),
             ("d1a", "c12", "c1", 3), ("d1b", "c1", "c1", 3), ("out", "c1", "out", 1))

    def __init__(self, in_ch=2, out_ch=1, c1=16, c2=32, seed=0, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        dims = {"in": in_ch, "out": out_ch, "c1": c1, "c2": c2, "c12": c1 + c2}
        self.c1, self.c2 = c1, c2
        g = torch.Generator().manual_seed(seed)
        self.params = {}
        for name, ci, co, k in self.SPECS:
            cin, cout = dims[ci], dims[co]
            std = math.sqrt(2.0 / (cin * k * k)) if name != "out" else math.sqrt(1.0 / (cin * k * k))
            w = (torch.randn((cout, cin, k, k), generator=g) * std).to(self.device)
            self.params[name] = [w, torch.zeros(cout, device=self.device)]
        self.adam = {n: [[torch.zeros_like(t) for t in ps], [torch.zeros_like(t) for t in ps]]
                     for n, ps in self.params.items()}
        self.step_count = 0
        self._cache = None

    # forward -------------------------------------------------------------
    def forward(self, x, keep=True):
        """x: (B, in_ch, H, W) FP32, H and W even."""
        P = self.params
        B, _, H, W = x.shape
        if H % 2 or W % 2:
            raise ValueError("the U-Net needs even H and W (pad the input)")
        e1a = conv(x, *P["e1a"], relu=True)
        e1 = conv(e1a, *P["e1b"], relu=True)
        p1 = pool2(e1)
        e2a = conv(p1, *P["e2a"], relu=True)
        e2 = conv(e2a, *P["e2b"], relu=True)
        cat = torch.empty((B, self.c1 + self.c2, H, W), dtype=torch.float32, device=x.device)
        cat[:, :self.c1].copy_(e1)
        up2_into(e2, cat, self.c1)
        d1a = conv(cat, *P["d1a"], relu=True)
        d1 = conv(d1a, *P["d1b"], relu=True)
        y = conv(d1, *P["out"], relu=False)
        if keep:
            self._cache = (x, e1a, e1, p1, e2a, e2, cat, d1a, d1)
        return y

    # backward ------------------------------------------------------------
    def backward(self, gy):
        """Parameter gradients {name: [dw, db]} of the last forward."""
        x, e1a, e1, p1, e2a, e2, cat, d1a, d1 = self._cache
        P = self.params
        G = {}
        G["out"] = list(conv_wgrad(d1, gy, P["out"][0].shape))
        g = relu_bwd_(conv_dgrad(gy, P["out"][0]), d1)
        G["d1b"] = list(conv_wgrad(d1a, g, P["d1b"][0].shape))
        g = relu_bwd_(conv_dgrad(g, P["d1b"][0]), d1a)
        G["d1a"] = list(conv_wgrad(cat, g, P["d1a"][0].shape))
        gcat = conv_dgrad(g, P["d1a"][0])
        h, w = e2.shape[2], e2.shape[3]
        g2 = relu_bwd_(up2_bwd_from(gcat, self.c1, self.c2, h, w), e2)
        G["e2b"] = list(conv_wgrad(e2a, g2, P["e2b"][0].shape))
        g2 = relu_bwd_(conv_dgrad(g2, P["e2b"][0]), e2a)
        G["e2a"] = list(conv_wgrad(p1, g2, P["e2a"][0].shape))
        g1 = pool2_bwd(conv_dgrad(g2, P["e2a"][0]), e1.shape)
        g1 += gcat[:, :self.c1]
        relu_bwd_(g1, e1)
        G["e1b"] = list(conv_wgrad(e1a, g1, P["e1b"][0].shape))
        g1 = relu_bwd_(conv_dgrad(g1, P["e1b"][0]), e1a)
        G["e1a"] = list(conv_wgrad(x, g1, P["e1a"][0].shape))
        return G

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for name, ps in self.params.items():
            m, v = self.adam[name]
            for i, t in enumerate(ps):
                _abi.check(_lib().stfwi_adam(_p(t), _p(grads[name][i]), _p(m[i]), _p(v[i]), t.numel(), lr, b1, b2,
                                             eps, self.step_count, _s()))

    def state_dict(self):
        return {f"{n}.{i}": t.detach().cpu() for n, ps in self.params.items() for i, t in enumerate(ps)}

    def load_state_dict(self, sd):
        for n, ps in self.params.items():
            for i in range(len(ps)):
                ps[i].copy_(sd[f"{n}.{i}"].to(self.device))


# ----------------------------------------------------- the preconditioner --

class LearnedPreconditioner:
    """(Vs gradient, illumination) -> Vs search direction on an (nz, nx)
    physical grid.  Inputs are normalised by their max |.|; the output is a
    unit-free direction the line search scales.  Training targets are the
    plain method's accepted Vs updates, normalised the same way."""

    def __init__(self, nz, nx, c1=16, c2=32, seed=0, device=None):
        self.nz, self.nx = int(nz), int(nx)
        self.H, self.W = self.nz + self.nz % 2, self.nx + self.nx % 2
        self.net = UNet(2, 1, c1, c2, seed, device)
        self.device = self.net.device
        self.pairs: list = []

    def _features(self, grad, illum):
        x = torch.zeros((1, 2, self.H, self.W), dtype=torch.float32, device=self.device)
        g = torch.as_tensor(grad, device=self.device).float().reshape(self.nz, self.nx)
        il = torch.ones_like(g) if illum is None else torch.as_tensor(illum, device=self.device).float().reshape(
            self.nz, self.nx)
        x[0, 0, :self.nz, :self.nx] = g / g.abs().max().clamp_min(1e-30)
        x[0, 1, :self.nz, :self.nx] = il / il.abs().max().clamp_min(1e-30)
        return x

    def direction(self, grad, illum=None):
        """Predicted Vs update direction (nz, nx), float64."""
        y = self.net.forward(self._features(grad, illum), keep=False)
        return y[0, 0, :self.nz, :self.nx].double()

    def add_pair(self, grad, illum, step):
        s = torch.as_tensor(step, device=self.device).float().reshape(self.nz, self.nx)
        t = torch.zeros((1, 1, self.H, self.W), dtype=torch.float32, device=self.device)
        t[0, 0, :self.nz, :self.nx] = s / s.abs().max().clamp_min(1e-30)
        self.pairs.append((self._features(grad, illum), t))

    def train(self, epochs=200, lr=1e-3, batch=4, seed=0):
        """Fit the stored pairs (MSE over the physical cells); returns the
        loss history (one value per epoch, mean over batches)."""
        if not self.pairs:
            return []
        rng = np.random.default_rng(seed)
        wgt = torch.zeros((1, 1, self.H, self.W), dtype=torch.float32, device=self.device)
        wgt[0, 0, :self.nz, :self.nx] = 1.0
        hist = []
        for _ in range(epochs):
            order = rng.permutation(len(self.pairs))
            tot, nb = 0.0, 0
            for i in range(0, len(order), batch):
                idx = order[i:i + batch]
                x = torch.cat([self.pairs[j][0] for j in idx])
                t = torch.cat([self.pairs[j][1] for j in idx])
                y = self.net.forward(x)
                loss, gy = mse(y, t, wgt.expand_as(t).contiguous())
                self.net.adam_step(self.net.backward(gy), lr=lr)
                tot += loss
                nb += 1
            hist.append(tot / nb)
        return hist
