"""3-D U-Net flow surrogate, trained and run with the library's own kernels.

The reference starts every solve from rest (f = w_q, engine.js:260-261) and has
no surrogate.  This network maps the voxel description of a box to its
averaged steady fields (ux, uy, uz, rho); its prediction seeds the lattice at
equilibrium (abx_sim_init_eq).

Architecture: 4 levels with 16/32/64/128 channels, two 3x3x3 convolutions +
ReLU per level, 2x2x2 max pooling down, nearest 2x upsampling up with the
encoder output concatenated as a skip, a 1x1x1 head.  Weights and activations
are bf16, accumulation FP32 (MFMA), master weights FP32 with Adam.

Every compute step is a kernel of libairbox_hip.so (abx_conv3d_*,
abx_maxpool2*, abx_upsample2*, abx_mse_loss, abx_adam; each ReLU's backward
is applied in the store of the gradient kernel before it);
torch only allocates device memory and provides the stream.  Forward and
backward are written out by hand.  A skip concatenation (decoder input =
[upsampled | skip]) is two contiguous tensors per level, "upK" and "eK", that
the decoder's first convolution, its input gradient and its weight gradient
read or write as one input (abx_conv3d_*_parts): no copy is made, and the
pooling and upsampling kernels touch whole rows instead of channel slices.

In training, the head, the loss and the head's backward are one kernel
(abx_head_loss_backward), between the decoder's forward and its backward.

The weight gradients (and their partial-sum reductions) run on a second
stream, each after an event that marks the main stream's work so far (its dZ
and input are then written), while the main stream carries on with the input
gradients; the main stream waits for them before Adam.  Only the weight
gradients use the workspace, so they stay in order on their own stream.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import abx

IN_CH = 16          # input channels as stored (the features, zero padded)
OUT_CH = 4          # ux, uy, uz, rho
HEAD_CH = 16        # head output channels as computed (OUT_CH used)
WIDTHS = (16, 32, 64, 128)


def _torch():
    import torch
    try:
        torch.empty(1, device="cuda")     # initialises torch's HIP context (or fails without a GPU)
    except Exception as e:
        raise abx.AbxError(f"the U-Net surrogate needs a GPU visible to torch ({e})") from None
    return torch


@dataclass
class Layer:
    name: str
    ci: int
    co: int
    ksize: int
    relu: bool
    w_off: int = 0      # offset of W [co][taps][ci] in the flat parameter vector
    b_off: int = 0

    @property
    def taps(self) -> int:
        return 27 if self.ksize == 3 else 1

    @property
    def nw(self) -> int:
        return self.co * self.taps * self.ci


def layers() -> List[Layer]:
    w1, w2, w3, w4 = WIDTHS
    L = [Layer("enc1a", IN_CH, w1, 3, True), Layer("enc1b", w1, w1, 3, True),
         Layer("enc2a", w1, w2, 3, True), Layer("enc2b", w2, w2, 3, True),
         Layer("enc3a", w2, w3, 3, True), Layer("enc3b", w3, w3, 3, True),
         Layer("bota", w3, w4, 3, True), Layer("botb", w4, w4, 3, True),
         Layer("dec3a", w4 + w3, w3, 3, True), Layer("dec3b", w3, w3, 3, True),
         Layer("dec2a", w3 + w2, w2, 3, True), Layer("dec2b", w2, w2, 3, True),
         Layer("dec1a", w2 + w1, w1, 3, True), Layer("dec1b", w1, w1, 3, True),
         Layer("head", w1, HEAD_CH, 1, False)]
    off = 0
    for l in L:
        l.w_off = off
        off += l.nw
        l.b_off = off
        off += l.co
    return L


def n_params() -> int:
    L = layers()
    return L[-1].b_off + L[-1].co


class T:
    """A channel slice [c0, c0+c) of an NDHWC bf16 tensor."""

    def __init__(self, t, c0: int = 0, c: Optional[int] = None):
        self.t = t
        self.c0 = c0
        self.c = t.shape[-1] - c0 if c is None else c
        self.s = t.shape[-1]

    @property
    def p(self) -> int:
        return self.t.data_ptr() + 2 * self.c0

    def sl(self, c0: int, c: int) -> "T":
        return T(self.t, self.c0 + c0, c)


def _ck(rc: int) -> None:
    if rc != 0:
        raise abx.AbxError(abx.load().abx_last_error().decode("utf-8"))


class UNet:
    """The network, its parameters (FP32 master + Adam moments + bf16 packs) and
    activation buffers for one batch shape (n, d, h, w)."""

    def __init__(self, seed: int = 0, device: str = "cuda"):
        torch = _torch()
        self.torch = torch
        self.dev = torch.device(device)
        self.L = layers()
        self.by = {l.name: l for l in self.L}
        P = n_params()
        rng = np.random.default_rng(seed)
        w = np.zeros(P, np.float32)
        for l in self.L:
            std = math.sqrt(2.0 / (l.taps * l.ci)) if l.relu else math.sqrt(1.0 / (l.taps * l.ci))
            W = (rng.standard_normal((l.co, l.taps, l.ci)) * std).astype(np.float32)
            if l.name == "head":
                W[OUT_CH:] = 0.0   # unused head channels stay zero (their gradient is zero)
            w[l.w_off:l.w_off + l.nw] = W.ravel()
        self.w = torch.from_numpy(w).to(self.dev)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.step_no = 0
        self.packs: Dict[str, Tuple] = {}
        for l in self.L:
            fw = torch.empty(l.nw, dtype=torch.bfloat16, device=self.dev)
            bw = torch.empty(l.nw, dtype=torch.bfloat16, device=self.dev)
            self.packs[l.name] = (fw, bw)
        self.repack()
        self.shape = None
        self.ws = None
        # weight gradients on a second stream (default priority), joined before Adam
        self.overlap = True
        # train_step: head, loss and head backward in one kernel
        self.fused_head = True
        self._side = None
        self._evs: List = []
        self._nev = 0

    # -- parameters ---------------------------------------------------------
    @property
    def stream(self) -> int:
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def _event(self):
        """A reusable event (recording it again after its wait was enqueued
        does not change that wait)."""
        if self._nev == len(self._evs):
            self._evs.append(self.torch.cuda.Event())
        ev = self._evs[self._nev]
        self._nev += 1
        return ev

    def _wstream(self) -> int:
        """The stream of the next weight gradient: the side stream, made to
        wait for everything enqueued on the main stream so far, or the main
        stream itself without overlap."""
        if not self.overlap:
            return self.stream
        torch = self.torch
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.dev)
        ev = self._event()
        ev.record(torch.cuda.current_stream(self.dev))
        self._side.wait_event(ev)
        return self._side.cuda_stream

    def _join(self) -> None:
        """The main stream waits for the weight gradients enqueued so far."""
        if self._side is not None:
            ev = self._event()
            ev.record(self._side)
            self.torch.cuda.current_stream(self.dev).wait_event(ev)

    def wptr(self, l: Layer, t=None) -> int:
        t = self.w if t is None else t
        return t.data_ptr() + 4 * l.w_off

    def bptr(self, l: Layer, t=None) -> int:
        t = self.w if t is None else t
        return t.data_ptr() + 4 * l.b_off

    def repack(self) -> None:
        """Both bf16 layouts of every layer's weights, in one launch."""
        if getattr(self, "_pack_table", None) is None:
            descs = (abx.PackDesc * len(self.L))()
            for i, l in enumerate(self.L):
                fw, bw = self.packs[l.name]
                descs[i] = abx.PackDesc(self.wptr(l), fw.data_ptr(), bw.data_ptr(), l.co, l.ksize ** 3, l.ci)
            self._pack_table = descs
        _ck(abx.load().abx_conv3d_pack_all(self.stream, len(self.L), self._pack_table))

    def state(self) -> dict:
        return {"w": self.w.cpu().numpy(), "m": self.m.cpu().numpy(), "v": self.v.cpu().numpy(),
                "step": np.int64(self.step_no)}

    def load_state(self, st: dict) -> None:
        torch = self.torch
        self.w.copy_(torch.from_numpy(np.asarray(st["w"], np.float32)))
        if "m" in st:
            self.m.copy_(torch.from_numpy(np.asarray(st["m"], np.float32)))
            self.v.copy_(torch.from_numpy(np.asarray(st["v"], np.float32)))
            self.step_no = int(st["step"])
        self.repack()

    # -- buffers --------------------------------------------------------------
    def _alloc(self, n: int, d: int, h: int, w: int) -> None:
        if self.shape == (n, d, h, w):
            return
        if d % 8 or h % 8 or w % 8:
            raise abx.AbxError(f"U-Net grid {d}x{h}x{w} must be a multiple of 8 per axis")
        torch = self.torch
        w1, w2, w3, w4 = WIDTHS
        dims = [(d >> k, h >> k, w >> k) for k in range(4)]
        E = lambda lvl, c: torch.zeros((n,) + dims[lvl] + (c,), dtype=torch.bfloat16, device=self.dev)
        A = {}
        A["a1"], A["up1"], A["e1"], A["p1"] = E(0, w1), E(0, w2), E(0, w1), E(1, w1)
        A["a2"], A["up2"], A["e2"], A["p2"] = E(1, w2), E(1, w3), E(1, w2), E(2, w2)
        A["a3"], A["up3"], A["e3"], A["p3"] = E(2, w3), E(2, w4), E(2, w3), E(3, w3)
        A["a4"], A["b4"] = E(3, w4), E(3, w4)
        A["d3a"], A["d3b"] = E(2, w3), E(2, w3)
        A["d2a"], A["d2b"] = E(1, w2), E(1, w2)
        A["d1a"], A["d1b"] = E(0, w1), E(0, w1)
        A["out"] = E(0, HEAD_CH)
        self.A = A
        # gradients
        G = {k: torch.zeros_like(v) for k, v in A.items()}
        self.G = G
        self.dims = dims
        self.shape = (n, d, h, w)
        lib = abx.load()
        need = 0
        for l in self.L:
            lvl = self._level(l)
            dd, hh, ww = dims[lvl]
            need = max(need, lib.abx_conv3d_wgrad_workspace(n, dd, hh, ww, l.ci, l.co, l.ksize))
        self.ws = torch.empty(need // 4 + 1, dtype=torch.float32, device=self.dev)
        self.ws_bytes = need

    @staticmethod
    def _level(l: Layer) -> int:
        return {"enc1a": 0, "enc1b": 0, "enc2a": 1, "enc2b": 1, "enc3a": 2, "enc3b": 2, "bota": 3, "botb": 3,
                "dec3a": 2, "dec3b": 2, "dec2a": 1, "dec2b": 1, "dec1a": 0, "dec1b": 0, "head": 0}[l.name]

    # -- kernels ----------------------------------------------------------------
    def _conv(self, name: str, x: T, y: T, accumulate: bool = False) -> None:
        l = self.by[name]
        n = self.shape[0]
        d, h, w = self.dims[self._level(l)]
        assert x.c == l.ci and y.c == l.co, (name, x.c, y.c)
        _ck(abx.load().abx_conv3d_forward(self.stream, x.p, x.s, n, d, h, w, l.ci, self.packs[name][0].data_ptr(),
                                          self.bptr(l), l.co, l.ksize, int(l.relu), int(accumulate), y.p, y.s))

    def _conv_parts(self, name: str, x1: T, x2: T, y: T) -> None:
        """Layer name over the concatenation [x1 | x2]."""
        l = self.by[name]
        n = self.shape[0]
        d, h, w = self.dims[self._level(l)]
        assert x1.c + x2.c == l.ci and y.c == l.co, (name, x1.c, x2.c, y.c)
        _ck(abx.load().abx_conv3d_forward_parts(self.stream, x1.p, x1.s, x1.c, x2.p, x2.s, n, d, h, w, l.ci,
                                                self.packs[name][0].data_ptr(), self.bptr(l), l.co, int(l.relu), 0,
                                                y.p, y.s, l.co, None, 0))

    def _dgrad_parts(self, name: str, dz: T, dx1: T, dx2: T) -> None:
        """Input gradient of layer name into the concatenation [dx1 | dx2]."""
        l = self.by[name]
        n = self.shape[0]
        d, h, w = self.dims[self._level(l)]
        assert dz.c == l.co and dx1.c + dx2.c == l.ci, (name, dz.c, dx1.c, dx2.c)
        _ck(abx.load().abx_conv3d_forward_parts(self.stream, dz.p, dz.s, l.co, None, 0, n, d, h, w, l.co,
                                                self.packs[name][1].data_ptr(), None, l.ci, 0, 0, dx1.p, dx1.s,
                                                dx1.c, dx2.p, dx2.s))

    def _wgrad_parts(self, name: str, x1: T, x2: T, dz: T) -> None:
        l = self.by[name]
        n = self.shape[0]
        d, h, w = self.dims[self._level(l)]
        _ck(abx.load().abx_conv3d_wgrad_parts(self._wstream(), x1.p, x1.s, x1.c, x2.p, x2.s, dz.p, dz.s, n, d, h, w,
                                              l.ci, l.co, self.ws.data_ptr(), self.ws_bytes, self.wptr(l, self.g),
                                              self.bptr(l, self.g), 0))

    def _dgrad(self, name: str, dz: T, dx: T, gate: Optional[T] = None) -> None:
        """dx = input gradient of layer name; with gate (the ReLU output that
        is this layer's input) the ReLU's backward is applied in the store."""
        l = self.by[name]
        n = self.shape[0]
        d, h, w = self.dims[self._level(l)]
        assert dz.c == l.co and dx.c == l.ci, (name, dz.c, dx.c)
        lib = abx.load()
        if gate is None:
            _ck(lib.abx_conv3d_forward(self.stream, dz.p, dz.s, n, d, h, w, l.co, self.packs[name][1].data_ptr(),
                                       None, l.ci, l.ksize, 0, 0, dx.p, dx.s))
        else:
            assert gate.c == l.ci, (name, gate.c)
            _ck(lib.abx_conv3d_forward_gated(self.stream, dz.p, dz.s, n, d, h, w, l.co,
                                             self.packs[name][1].data_ptr(), None, l.ci, l.ksize, 0, 0, dx.p, dx.s,
                                             gate.p, gate.s))

    def _wgrad(self, name: str, x: T, dz: T) -> None:
        l = self.by[name]
        n = self.shape[0]
        d, h, w = self.dims[self._level(l)]
        _ck(abx.load().abx_conv3d_wgrad(self._wstream(), x.p, x.s, dz.p, dz.s, n, d, h, w, l.ci, l.co, l.ksize,
                                        self.ws.data_ptr(), self.ws_bytes, self.wptr(l, self.g),
                                        self.bptr(l, self.g), 0))

    def _pool(self, x: T, y: T, lvl: int) -> None:
        n = self.shape[0]
        d, h, w = self.dims[lvl]
        _ck(abx.load().abx_maxpool2(self.stream, x.p, x.s, n, d, h, w, x.c, y.p, y.s))

    def _pool_bwd(self, x: T, dy: T, dx: T, lvl: int) -> None:
        """dx += pooling gradient, then the backward of the ReLU whose output
        is x (the pooled encoder output), in one pass."""
        n = self.shape[0]
        d, h, w = self.dims[lvl]
        _ck(abx.load().abx_maxpool2_backward_gated(self.stream, x.p, x.s, dy.p, dy.s, n, d, h, w, x.c, dx.p, dx.s, 1))

    def _up(self, x: T, y: T, lvl: int) -> None:
        n = self.shape[0]
        d, h, w = self.dims[lvl]     # coarse level
        _ck(abx.load().abx_upsample2(self.stream, x.p, x.s, n, d, h, w, x.c, y.p, y.s))

    def _up_bwd(self, dy: T, dx: T, lvl: int, gate: T) -> None:
        """dx = upsampling gradient, masked by the backward of the ReLU whose
        output (gate) was upsampled."""
        n = self.shape[0]
        d, h, w = self.dims[lvl]
        _ck(abx.load().abx_upsample2_backward_gated(self.stream, dy.p, dy.s, n, d, h, w, dx.c, dx.p, dx.s, gate.p,
                                                    gate.s))

    # -- network ----------------------------------------------------------------
    def forward(self, x, head: bool = True) -> "object":
        """x: (n, d, h, w, IN_CH) bf16 on the device.  Returns the head output
        (n, d, h, w, HEAD_CH) bf16 (channels >= OUT_CH are zero); with
        head=False the head is left to head_loss_backward (out is not
        written)."""
        n, d, h, w, c = x.shape
        assert c == IN_CH and x.dtype == self.torch.bfloat16 and x.is_contiguous()
        self._alloc(n, d, h, w)
        A = self.A
        w1, w2, w3, w4 = WIDTHS
        X = T(x)
        e1, e2, e3 = T(A["e1"]), T(A["e2"]), T(A["e3"])
        self._conv("enc1a", X, T(A["a1"]))
        self._conv("enc1b", T(A["a1"]), e1)
        self._pool(e1, T(A["p1"]), 0)
        self._conv("enc2a", T(A["p1"]), T(A["a2"]))
        self._conv("enc2b", T(A["a2"]), e2)
        self._pool(e2, T(A["p2"]), 1)
        self._conv("enc3a", T(A["p2"]), T(A["a3"]))
        self._conv("enc3b", T(A["a3"]), e3)
        self._pool(e3, T(A["p3"]), 2)
        self._conv("bota", T(A["p3"]), T(A["a4"]))
        self._conv("botb", T(A["a4"]), T(A["b4"]))
        self._up(T(A["b4"]), T(A["up3"]), 3)
        self._conv_parts("dec3a", T(A["up3"]), e3, T(A["d3a"]))
        self._conv("dec3b", T(A["d3a"]), T(A["d3b"]))
        self._up(T(A["d3b"]), T(A["up2"]), 2)
        self._conv_parts("dec2a", T(A["up2"]), e2, T(A["d2a"]))
        self._conv("dec2b", T(A["d2a"]), T(A["d2b"]))
        self._up(T(A["d2b"]), T(A["up1"]), 1)
        self._conv_parts("dec1a", T(A["up1"]), e1, T(A["d1a"]))
        self._conv("dec1b", T(A["d1a"]), T(A["d1b"]))
        if head:
            self._conv("head", T(A["d1b"]), T(A["out"]))
        self.x_in = x
        return A["out"]

    def backward(self, dout) -> None:
        """Gradients of every parameter into self.g (overwritten) from the
        gradient of the head output (n, d, h, w, HEAD_CH) bf16, or with dout
        None from the head's gradients that head_loss_backward wrote."""
        A, G = self.A, self.G
        w1, w2, w3, w4 = WIDTHS
        a = lambda k: T(A[k])
        g = lambda k: T(G[k])
        self._nev = 0
        if dout is not None:
            dO = T(dout)
            self._wgrad("head", a("d1b"), dO)
            self._dgrad("head", dO, g("d1b"), gate=a("d1b"))
        # level 1 decoder
        self._wgrad("dec1b", a("d1a"), g("d1b"))
        self._dgrad("dec1b", g("d1b"), g("d1a"), gate=a("d1a"))
        self._wgrad_parts("dec1a", a("up1"), a("e1"), g("d1a"))
        self._dgrad_parts("dec1a", g("d1a"), g("up1"), g("e1"))
        self._up_bwd(g("up1"), g("d2b"), 1, gate=a("d2b"))
        # level 2 decoder
        self._wgrad("dec2b", a("d2a"), g("d2b"))
        self._dgrad("dec2b", g("d2b"), g("d2a"), gate=a("d2a"))
        self._wgrad_parts("dec2a", a("up2"), a("e2"), g("d2a"))
        self._dgrad_parts("dec2a", g("d2a"), g("up2"), g("e2"))
        self._up_bwd(g("up2"), g("d3b"), 2, gate=a("d3b"))
        # level 3 decoder
        self._wgrad("dec3b", a("d3a"), g("d3b"))
        self._dgrad("dec3b", g("d3b"), g("d3a"), gate=a("d3a"))
        self._wgrad_parts("dec3a", a("up3"), a("e3"), g("d3a"))
        self._dgrad_parts("dec3a", g("d3a"), g("up3"), g("e3"))
        self._up_bwd(g("up3"), g("b4"), 3, gate=a("b4"))
        # bottleneck
        self._wgrad("botb", a("a4"), g("b4"))
        self._dgrad("botb", g("b4"), g("a4"), gate=a("a4"))
        self._wgrad("bota", a("p3"), g("a4"))
        self._dgrad("bota", g("a4"), g("p3"))
        # level 3 encoder: skip gradient (from dec3a) + pooling gradient
        e3, de3 = a("e3"), g("e3")
        self._pool_bwd(e3, g("p3"), de3, 2)
        self._wgrad("enc3b", a("a3"), de3)
        self._dgrad("enc3b", de3, g("a3"), gate=a("a3"))
        self._wgrad("enc3a", a("p2"), g("a3"))
        self._dgrad("enc3a", g("a3"), g("p2"))
        # level 2 encoder
        e2, de2 = a("e2"), g("e2")
        self._pool_bwd(e2, g("p2"), de2, 1)
        self._wgrad("enc2b", a("a2"), de2)
        self._dgrad("enc2b", de2, g("a2"), gate=a("a2"))
        self._wgrad("enc2a", a("p1"), g("a2"))
        self._dgrad("enc2a", g("a2"), g("p1"))
        # level 1 encoder
        e1, de1 = a("e1"), g("e1")
        self._pool_bwd(e1, g("p1"), de1, 0)
        self._wgrad("enc1b", a("a1"), de1)
        self._dgrad("enc1b", de1, g("a1"), gate=a("a1"))
        self._wgrad("enc1a", T(self.x_in), g("a1"))
        self._join()

    def adam(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.step_no += 1
        _ck(abx.load().abx_adam(self.stream, self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                                self.v.data_ptr(), self.w.numel(), lr, b1, b2, eps, self.step_no))
        self.repack()

    def loss_and_grad(self, target, mask, chan_w, count: Optional[int] = None) -> Tuple["object", "object"]:
        """Masked weighted MSE of the head output against target (n, d, h, w,
        OUT_CH) bf16 over voxels where mask (n, d, h, w) uint8 is set.  Returns
        (loss as a device scalar, gradient of the head output).  count: the
        number of set mask voxels when the caller knows it; otherwise it is
        counted on the device and read back, which waits for the GPU."""
        torch = self.torch
        out = self.A["out"]
        nvox = out.numel() // HEAD_CH
        lib = abx.load()
        nparts = lib.abx_mse_partials(nvox)
        if not hasattr(self, "_lp") or self._lp.numel() != nparts:
            self._lp = torch.empty(nparts, dtype=torch.float32, device=self.dev)
            self._dout = torch.empty_like(out)
        cnt = float(mask.sum().item()) if count is None else float(count)
        scale = 1.0 / max(cnt * OUT_CH, 1.0)
        _ck(lib.abx_mse_loss(self.stream, out.data_ptr(), HEAD_CH, target.data_ptr(), target.shape[-1],
                             mask.data_ptr(), nvox, OUT_CH, chan_w.data_ptr(), scale, self._lp.data_ptr(),
                             self._dout.data_ptr(), HEAD_CH))
        return self._lp.sum(), self._dout

    def head_loss_backward(self, target, mask, chan_w, count: Optional[int] = None) -> "object":
        """After forward(x, head=False): the head, the loss of loss_and_grad
        and the head's input and weight gradients in one kernel
        (abx_head_loss_backward).  Returns the loss as a device scalar."""
        torch = self.torch
        l = self.by["head"]
        d1b = T(self.A["d1b"])
        nvox = self.A["d1b"].numel() // d1b.s
        lib = abx.load()
        nparts = lib.abx_head_loss_partials(nvox)
        if not hasattr(self, "_hlp") or self._hlp.numel() != nparts:
            self._hlp = torch.empty(nparts, dtype=torch.float32, device=self.dev)
            self._hws_bytes = lib.abx_head_loss_workspace(nvox)
            self._hws = torch.empty(self._hws_bytes // 4, dtype=torch.float32, device=self.dev)
        cnt = float(mask.sum().item()) if count is None else float(count)
        scale = 1.0 / max(cnt * OUT_CH, 1.0)
        gx = T(self.G["d1b"])
        _ck(lib.abx_head_loss_backward(self.stream, d1b.p, d1b.s, l.ci, self.packs["head"][0].data_ptr(),
                                       self.bptr(l), OUT_CH, target.data_ptr(), target.shape[-1], mask.data_ptr(),
                                       nvox, chan_w.data_ptr(), scale, self._hlp.data_ptr(), gx.p, gx.s,
                                       self._hws.data_ptr(), self._hws_bytes, self.wptr(l, self.g),
                                       self.bptr(l, self.g)))
        return self._hlp.sum()

    def train_step(self, x, target, mask, chan_w, lr: float = 1e-3, count: Optional[int] = None) -> "object":
        if self.fused_head:
            self.forward(x, head=False)
            loss = self.head_loss_backward(target, mask, chan_w, count)
            self.backward(None)
        else:
            self.forward(x)
            loss, dout = self.loss_and_grad(target, mask, chan_w, count)
            self.backward(dout)
        self.adam(lr)
        return loss
