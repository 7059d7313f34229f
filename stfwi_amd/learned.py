"""Learned gradient preconditioner for the plain L-BFGS method.

A U-Net maps three maps of an FWI iteration -- the raw Vs gradient, the
illumination and the current Vs model -- to a search direction.  It is
trained on pairs that the plain method produces (the inputs at an iterate
and the model update the plain method went on to accept), and the driver
falls back to the plain L-BFGS step whenever the learned direction does not
decrease the misfit.  The reference has no learned part; this module has no
reference counterpart.

Every convolution runs on the matrix cores through hand-written HIP kernels
(``stfwi_conv2d_bf16`` and its data/weight gradients: BF16 operands, FP32
accumulation); pooling, upsampling, the ReLU gradient, the MSE loss and Adam
are HIP kernels too.  Parameters and activations are FP32 device tensors;
torch only allocates and copies them.

Architecture, for channels (c0, .., c{L-1}) = (16, 32, 64, 128) by default:
    encoder level l: conv3 -> c_l, conv3 c_l -> c_l = e_l, then 2x2 average
        pooling into level l + 1 (not after the last level);
    decoder level l = L-2 .. 0: nearest 2x upsampling of the level below,
        concatenated after e_l, conv3 -> c_l, conv3 c_l -> c_l;
    a 1x1 convolution to the output;
with ReLU after every 3x3 convolution.  H and W must be multiples of
2^(L-1); the preconditioner pads the grid to that.
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
    def __init__(self, in_ch=3, out_ch=1, channels=(16, 32, 64, 128), seed=0, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.channels = tuple(int(c) for c in channels)
        L = len(self.channels)
        specs, prev = [], in_ch
        for l, c in enumerate(self.channels):
            specs += [(f"enc{l}a", prev, c, 3), (f"enc{l}b", c, c, 3)]
            prev = c
        for l in range(L - 2, -1, -1):
            c = self.channels[l]
            specs += [(f"dec{l}a", c + self.channels[l + 1], c, 3), (f"dec{l}b", c, c, 3)]
        specs.append(("out", self.channels[0], out_ch, 1))
        self.specs = specs
        g = torch.Generator().manual_seed(seed)
        self.params = {}
        for name, cin, cout, k in specs:
            std = math.sqrt((1.0 if name == "out" else 2.0) / (cin * k * k))
            w = (torch.randn((cout, cin, k, k), generator=g) * std).to(self.device)
            self.params[name] = [w, torch.zeros(cout, device=self.device)]
        self.adam = {n: [[torch.zeros_like(t) for t in ps], [torch.zeros_like(t) for t in ps]]
                     for n, ps in self.params.items()}
        self.step_count = 0
        self._cache = None

    @property
    def multiple(self):
        return 1 << (len(self.channels) - 1)

    def n_params(self):
        return sum(t.numel() for ps in self.params.values() for t in ps)

    # forward -------------------------------------------------------------
    def forward(self, x, keep=True):
        """x: (B, in_ch, H, W) FP32, H and W multiples of 2^(L-1)."""
        P, C, L = self.params, self.channels, len(self.channels)
        B, _, H, W = x.shape
        if H % self.multiple or W % self.multiple:
            raise ValueError(f"the U-Net needs H and W divisible by {self.multiple} (pad the input)")
        enc = []                      # per level: (input, first conv, second conv = skip)
        h = x
        for l in range(L):
            a = conv(h, *P[f"enc{l}a"], relu=True)
            e = conv(a, *P[f"enc{l}b"], relu=True)
            enc.append((h, a, e))
            if l < L - 1:
                h = pool2(e)
        dec = {}                      # per level: (concat, first conv, second conv)
        d = enc[-1][2]
        for l in range(L - 2, -1, -1):
            skip = enc[l][2]
            cat = torch.empty((B, C[l] + C[l + 1], skip.shape[2], skip.shape[3]), dtype=torch.float32,
                              device=x.device)
            cat[:, :C[l]].copy_(skip)
            up2_into(d, cat, C[l])
            a = conv(cat, *P[f"dec{l}a"], relu=True)
            d = conv(a, *P[f"dec{l}b"], relu=True)
            dec[l] = (cat, a, d)
        y = conv(d, *P["out"], relu=False)
        if keep:
            self._cache = (enc, dec, d)
        return y

    # backward ------------------------------------------------------------
    def backward(self, gy):
        """Parameter gradients {name: [dw, db]} of the last forward."""
        enc, dec, d_top = self._cache
        P, C, L = self.params, self.channels, len(self.channels)
        G = {}
        G["out"] = list(conv_wgrad(d_top, gy, P["out"][0].shape))
        g = relu_bwd_(conv_dgrad(gy, P["out"][0]), d_top)       # w.r.t. the last decoder output
        skip_grad = {}
        for l in range(0, L - 1):
            cat, a, d = dec[l]
            G[f"dec{l}b"] = list(conv_wgrad(a, g, P[f"dec{l}b"][0].shape))
            g = relu_bwd_(conv_dgrad(g, P[f"dec{l}b"][0]), a)
            G[f"dec{l}a"] = list(conv_wgrad(cat, g, P[f"dec{l}a"][0].shape))
            gcat = conv_dgrad(g, P[f"dec{l}a"][0])
            skip_grad[l] = gcat[:, :C[l]]
            below = dec[l + 1][2] if l + 1 < L - 1 else enc[L - 1][2]
            g = relu_bwd_(up2_bwd_from(gcat, C[l], C[l + 1], below.shape[2], below.shape[3]), below)
        # g: gradient w.r.t. the deepest encoder output, ReLU applied
        for l in range(L - 1, -1, -1):
            h, a, e = enc[l]
            if l < L - 1:
                g = g + skip_grad[l]
                relu_bwd_(g, e)
            G[f"enc{l}b"] = list(conv_wgrad(a, g, P[f"enc{l}b"][0].shape))
            g = relu_bwd_(conv_dgrad(g, P[f"enc{l}b"][0]), a)
            G[f"enc{l}a"] = list(conv_wgrad(h, g, P[f"enc{l}a"][0].shape))
            if l > 0:
                g = pool2_bwd(conv_dgrad(g, P[f"enc{l}a"][0]), enc[l - 1][2].shape)
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

def _unit(t):
    return t / t.abs().max().clamp_min(1e-30)


class LearnedPreconditioner:
    """(Vs gradient, illumination, Vs) -> Vs update on an (nz, nx) physical
    grid.  The gradient and the illumination enter divided by their largest
    |value|, the model divided by `vs_ref` m/s.  The network predicts the
    update in units of `step_ref` m/s, so its output carries a step length
    as well as a direction (``absolute``: the driver tries it as the step
    itself).  Training targets are model updates divided by `step_ref`."""

    absolute = True

    def __init__(self, nz, nx, channels=(16, 32, 64, 128), seed=0, device=None, vs_ref=3000.0, step_ref=100.0,
                 policy="every", maxls=10):
        self.nz, self.nx = int(nz), int(nx)
        self.net = UNet(3, 1, channels, seed, device)
        m = self.net.multiple
        self.H, self.W = -(-self.nz // m) * m, -(-self.nx // m) * m
        self.device = self.net.device
        self.vs_ref = float(vs_ref)
        self.step_ref = float(step_ref)
        # how the FWI driver uses it: when a proposal is made ('every'
        # iteration, each 'stage' start, or the 'run' start) and how many
        # misfit evaluations its backtracking may spend before the plain
        # step is taken
        self.policy = policy
        self.maxls = int(maxls)
        # factor on the predicted update before it is tried
        self.scale = 1.0
        self.pairs: list = []

    def features(self, grad, illum, vs):
        x = torch.zeros((1, 3, self.H, self.W), dtype=torch.float32, device=self.device)
        g = torch.as_tensor(grad, device=self.device).float().reshape(self.nz, self.nx)
        il = torch.ones_like(g) if illum is None else torch.as_tensor(illum, device=self.device).float().reshape(
            self.nz, self.nx)
        v = torch.as_tensor(vs, device=self.device).float().reshape(self.nz, self.nx)
        x[0, 0, :self.nz, :self.nx] = _unit(g)
        x[0, 1, :self.nz, :self.nx] = _unit(il)
        x[0, 2, :self.nz, :self.nx] = v / self.vs_ref
        return x

    def target(self, step):
        t = torch.zeros((1, 1, self.H, self.W), dtype=torch.float32, device=self.device)
        t[0, 0, :self.nz, :self.nx] = torch.as_tensor(step, device=self.device).float().reshape(
            self.nz, self.nx) / self.step_ref
        return t

    def direction(self, grad, illum=None, vs=None):
        """Predicted Vs update (nz, nx) in m/s, float64."""
        if vs is None:
            vs = torch.full((self.nz, self.nx), self.vs_ref, device=self.device)
        y = self.net.forward(self.features(grad, illum, vs), keep=False)
        return y[0, 0, :self.nz, :self.nx].double() * self.step_ref

    def add_pair(self, grad, illum, vs, step):
        self.pairs.append((self.features(grad, illum, vs), self.target(step)))

    def train(self, epochs=200, lr=1e-3, batch=4, seed=0, pairs=None, flip=False):
        """Fit the pairs (MSE over the physical cells); returns the loss
        history (one value per epoch, mean over batches).  flip=True mirrors
        the physical columns of every other batch (a survey with sources
        and receivers spread evenly along the surface is nearly symmetric
        under x -> nx - 1 - x)."""
        pairs = self.pairs if pairs is None else pairs
        if not pairs:
            return []
        rng = np.random.default_rng(seed)
        wgt = torch.zeros((1, 1, self.H, self.W), dtype=torch.float32, device=self.device)
        wgt[0, 0, :self.nz, :self.nx] = 1.0
        hist = []
        nx = self.nx
        for _ in range(epochs):
            order = rng.permutation(len(pairs))
            tot, nb = 0.0, 0
            for i in range(0, len(order), batch):
                idx = order[i:i + batch]
                x = torch.cat([pairs[j][0] for j in idx])
                t = torch.cat([pairs[j][1] for j in idx])
                if flip and rng.random() < 0.5:
                    x[..., :nx] = x[..., :nx].flip(-1)
                    t[..., :nx] = t[..., :nx].flip(-1)
                y = self.net.forward(x)
                loss, gy = mse(y, t, wgt.expand_as(t).contiguous())
                self.net.adam_step(self.net.backward(gy), lr=lr)
                tot += loss
                nb += 1
            hist.append(tot / nb)
        return hist

    def save(self, path, optimizer=True):
        """Weights and settings, with the Adam moments unless optimizer is
        False (with them, training can continue from the file)."""
        net = self.net
        d = {"state": net.state_dict(), "channels": net.channels, "nz": self.nz, "nx": self.nx,
             "vs_ref": self.vs_ref, "step_ref": self.step_ref}
        if optimizer:
            d["adam"] = {f"{n}.{j}.{i}": t.detach().cpu() for n, mv in net.adam.items() for j in range(2)
                         for i, t in enumerate(mv[j])}
            d["adam_step"] = net.step_count
        torch.save(d, path)

    @classmethod
    def load(cls, path, device=None):
        d = torch.load(path, map_location="cpu", weights_only=True)
        lp = cls(d["nz"], d["nx"], tuple(d["channels"]), device=device, vs_ref=d["vs_ref"], step_ref=d["step_ref"])
        lp.net.load_state_dict(d["state"])
        if "adam" in d:
            for n, mv in lp.net.adam.items():
                for j in range(2):
                    for i in range(len(mv[j])):
                        mv[j][i].copy_(d["adam"][f"{n}.{j}.{i}"].to(lp.device))
            lp.net.step_count = int(d["adam_step"])
        return lp


class PairCollector:
    """Training pairs from plain-method runs (run_multiscale_fwi(collect=)).

    Every accepted step of a Vs-only stage gives the inputs at its starting
    iterate.  Each pair receives three targets: the accepted step itself
    ('step'), the update from its iterate to the stage's final model
    ('ahead', the rest of the plain stage in one step) and the update to the
    run's final model ('final', the rest of the plain run)."""

    MODES = ("step", "ahead", "final")

    def __init__(self, lp: LearnedPreconditioner):
        self.lp = lp
        self.stage_open: list = []
        self.run_open: list = []
        self.pairs = {m: [] for m in self.MODES}

    def add_step(self, grad, illum, vs, step):
        self.stage_open.append((self.lp.features(grad, illum, vs), self.lp.target(step),
                                torch.as_tensor(vs, device=self.lp.device).detach().clone()))

    def end_stage(self, vs_final):
        vf = torch.as_tensor(vs_final, device=self.lp.device)
        for x, t_step, vs in self.stage_open:
            self.pairs["step"].append((x, t_step))
            self.pairs["ahead"].append((x, self.lp.target(vf - vs)))
            self.run_open.append((x, vs))
        self.stage_open = []

    def end_run(self, vs_final):
        vf = torch.as_tensor(vs_final, device=self.lp.device)
        for x, vs in self.run_open:
            self.pairs["final"].append((x, self.lp.target(vf - vs)))
        self.run_open = []

    def __len__(self):
        return len(self.pairs["step"])
