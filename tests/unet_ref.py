"""FP32/FP64 references for the U-Net kernels (test infrastructure).

* ``direct_conv3d``: brute-force 'same' 3-D cross-correlation in float64 over
  NDHWC arrays, weights [co][taps][ci] with tap = dz*9 + dy*3 + dx (offsets
  -1..1 along z, y, x), the layout the library's kernels use;
* ``pool_ref``/``pool_bwd_ref``/``up_ref``/``up_bwd_ref``/``wgrad_ref``: the
  other kernels of the network, in numpy;
* ``torch_unet``: the surrogate's network restated with torch.nn.functional in
  float32 (conv3d, max_pool3d, nearest interpolate, cat), to check the
  hand-written forward and backward end to end.
"""
from __future__ import annotations

import numpy as np


def direct_conv3d(x: np.ndarray, w: np.ndarray, b=None, ksize: int = 3, relu: bool = False) -> np.ndarray:
    """x (n, d, h, w, ci), w (co, taps, ci) -> (n, d, h, w, co), float64."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    n, D, H, W, ci = x.shape
    co = w.shape[0]
    r = ksize // 2
    xp = np.zeros((n, D + 2 * r, H + 2 * r, W + 2 * r, ci))
    xp[:, r:r + D, r:r + H, r:r + W] = x
    y = np.zeros((n, D, H, W, co))
    for t in range(ksize ** 3):
        dz, dy, dx = t // (ksize * ksize), (t // ksize) % ksize, t % ksize
        sl = xp[:, dz:dz + D, dy:dy + H, dx:dx + W]          # x[v + offset]
        y += np.einsum("ndhwc,oc->ndhwo", sl, w[:, t, :])
    if b is not None:
        y += np.asarray(b, np.float64)
    if relu:
        y = np.maximum(y, 0.0)
    return y


def pool_ref(x):
    """2x2x2 max of NDHWC x and the window offset of the first maximum (z, y, x order)."""
    n, d, h, w, c = x.shape
    win = x.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).transpose(0, 1, 3, 5, 2, 4, 6, 7).reshape(
        n, d // 2, h // 2, w // 2, 8, c)
    return win.max(4), win.argmax(4)


def pool_bwd_ref(x, dy):
    n, d, h, w, c = x.shape
    _, am = pool_ref(x)
    win = (np.arange(8)[None, None, None, None, :, None] == am[:, :, :, :, None, :]) * dy[:, :, :, :, None, :]
    return win.reshape(n, d // 2, h // 2, w // 2, 2, 2, 2, c).transpose(0, 1, 4, 2, 5, 3, 6, 7).reshape(x.shape)


def up_ref(x):
    return x.repeat(2, 1).repeat(2, 2).repeat(2, 3)


def up_bwd_ref(dy):
    n, D, H, W, c = dy.shape
    return dy.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).sum((2, 4, 6))


def wgrad_ref(x, dz, ksize):
    n, d, h, w, ci = x.shape
    r = ksize // 2
    xp = np.zeros((n, d + 2 * r, h + 2 * r, w + 2 * r, ci))
    xp[:, r:r + d, r:r + h, r:r + w] = x
    out = np.zeros((dz.shape[-1], ksize ** 3, ci))
    for t in range(ksize ** 3):
        a, b, c = t // (ksize * ksize), (t // ksize) % ksize, t % ksize
        out[:, t] = np.einsum("ndhwo,ndhwc->oc", dz, xp[:, a:a + d, b:b + h, c:c + w])
    return out, dz.sum((0, 1, 2, 3))


def to_torch_weight(w, ksize: int):
    """[co][taps][ci] -> torch's [co][ci][kd][kh][kw]."""
    co, taps, ci = w.shape
    return w.reshape(co, ksize, ksize, ksize, ci).permute(0, 4, 1, 2, 3).contiguous()


def _bf16_round_fn():
    import torch

    class RoundBF16(torch.autograd.Function):
        """Identity up to bf16 rounding, forward and backward: marks where the
        kernels store a tensor, or the gradient of a tensor, in bf16."""

        @staticmethod
        def forward(ctx, t):
            return t.to(torch.bfloat16).to(t.dtype)

        @staticmethod
        def backward(ctx, g):
            return g.to(torch.bfloat16).to(g.dtype)

    return RoundBF16.apply


def torch_unet(x, params: dict, layers, bf16_points: bool = False):
    """Forward of the surrogate network in torch, in x's dtype (NCDHW inside).
    x: (n, d, h, w, IN_CH) float32; params: name -> (W [co][taps][ci], b [co]).
    Returns the head output (n, d, h, w, HEAD_CH).  With bf16_points the
    activations and their gradients are rounded to bf16 where the kernels
    store them (conv outputs, pooled and upsampled tensors, skip inputs)."""
    import torch
    import torch.nn.functional as F

    L = {l.name: l for l in layers}
    rnd = _bf16_round_fn() if bf16_points else (lambda t: t)

    def conv(name, t):
        l = L[name]
        W, b = params[name]
        y = F.conv3d(t, to_torch_weight(W, l.ksize), b, padding=l.ksize // 2)
        return rnd(F.relu(y) if l.relu else y)

    pool = lambda z: rnd(F.max_pool3d(z, 2))
    up = lambda z: rnd(F.interpolate(z, scale_factor=2, mode="nearest"))
    t = x.permute(0, 4, 1, 2, 3)
    e1 = conv("enc1b", conv("enc1a", t))
    e2 = conv("enc2b", conv("enc2a", pool(e1)))
    e3 = conv("enc3b", conv("enc3a", pool(e2)))
    b4 = conv("botb", conv("bota", pool(e3)))
    d3 = conv("dec3b", conv("dec3a", torch.cat([up(b4), rnd(e3)], 1)))
    d2 = conv("dec2b", conv("dec2a", torch.cat([up(d3), rnd(e2)], 1)))
    d1 = conv("dec1b", conv("dec1a", torch.cat([up(d2), rnd(e1)], 1)))
    out = conv("head", d1)
    return out.permute(0, 2, 3, 4, 1)
