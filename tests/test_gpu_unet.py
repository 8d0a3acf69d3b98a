"""U-Net surrogate kernels on the GPU against FP32/FP64 references.

Operands are rounded to bf16 before the reference sees them, so the only
differences left are FP32 accumulation order and the bf16 rounding of the
kernels' bf16 outputs.  Bars (written per test): bf16 outputs within
2^-7 relative plus 2^-9 of the output's range; FP32 weight gradients within
1e-3 relative of the range.
"""
from __future__ import annotations

import numpy as np
import pytest

from airboxcalc_amd import abx
from tests.gpu_util import say
from tests import unet_ref as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    abx.device(0)   # gfx950 check; raises without the GPU
    return torch.device("cuda:0")


def bf(a) -> "torch.Tensor":
    """float tensor -> bf16 tensor on the GPU."""
    return torch.as_tensor(a, dtype=torch.float32).to(_dev()).to(torch.bfloat16)


def f32(t) -> np.ndarray:
    return t.float().cpu().numpy()


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def ck(rc):
    if rc != 0:
        raise abx.AbxError(abx.load().abx_last_error().decode())


def pack(W, ksize, flip=False):
    co, taps, ci = W.shape
    w32 = torch.as_tensor(W, dtype=torch.float32).to(_dev()).contiguous()
    out = torch.empty(co * taps * ci, dtype=torch.bfloat16, device=_dev())
    ck(abx.load().abx_conv3d_pack(stream(), w32.data_ptr(), co, ksize, ci, int(flip), out.data_ptr()))
    return out, w32


def close_bf16(got, ref, what):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    rng = max(np.abs(ref).max(), 1e-30)
    tol = 2.0 ** -7 * np.abs(ref) + 2.0 ** -9 * rng
    bad = np.abs(got - ref) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} outside bf16 bar, max err {np.abs(got - ref).max():.3e} (range {rng:.3e})"


CASES = [
    # n, d, h, w, ci, co, ksize  (d, h, w not multiples of the 2x4x16 tile)
    (2, 6, 5, 20, 16, 32, 3),
    (1, 4, 8, 16, 48, 16, 3),
    (1, 3, 3, 12, 192, 64, 3),
    (1, 2, 2, 8, 64, 128, 3),
    (2, 5, 4, 17, 16, 16, 1),
    (1, 3, 7, 33, 96, 32, 3),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv_forward_matches_torch(case):
    n, d, h, w, ci, co, k = case
    say("conv forward", case)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((n, d, h, w, ci)).astype(np.float32)
    W = (rng.standard_normal((co, k ** 3, ci)) / np.sqrt(k ** 3 * ci)).astype(np.float32)
    b = rng.standard_normal(co).astype(np.float32)
    xt = bf(X)
    wp, _ = pack(W, k)
    bt = torch.as_tensor(b).to(_dev())
    y = torch.empty((n, d, h, w, co), dtype=torch.bfloat16, device=_dev())
    ck(abx.load().abx_conv3d_forward(stream(), xt.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), bt.data_ptr(), co,
                                     k, 1, 0, y.data_ptr(), co))
    torch.cuda.synchronize()
    Xr = xt.float()
    Wr = wp.float().reshape(co, k ** 3, ci)
    import torch.nn.functional as F
    ref = F.relu(F.conv3d(Xr.permute(0, 4, 1, 2, 3), R.to_torch_weight(Wr, k), bt, padding=k // 2))
    close_bf16(f32(y), ref.permute(0, 2, 3, 4, 1).cpu().numpy(), f"forward {case}")


def test_conv_forward_brute_force():
    """Tiny case against a float64 direct convolution (no torch in the reference)."""
    n, d, h, w, ci, co = 1, 3, 3, 5, 16, 16
    say("conv forward brute force")
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n, d, h, w, ci)).astype(np.float32)
    W = rng.standard_normal((co, 27, ci)).astype(np.float32) * 0.1
    b = rng.standard_normal(co).astype(np.float32)
    xt = bf(X)
    wp, _ = pack(W, 3)
    bt = torch.as_tensor(b).to(_dev())
    y = torch.empty((n, d, h, w, co), dtype=torch.bfloat16, device=_dev())
    ck(abx.load().abx_conv3d_forward(stream(), xt.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), bt.data_ptr(), co,
                                     3, 0, 0, y.data_ptr(), co))
    torch.cuda.synchronize()
    ref = R.direct_conv3d(f32(xt), f32(wp).reshape(co, 27, ci), b, 3, relu=False)
    close_bf16(f32(y), ref, "brute force")


@pytest.mark.parametrize("co,ci", [(16, 16), (48, 16), (64, 16), (16, 48), (32, 32)])
def test_conv_forward16_matches_general_kernel(co, ci):
    """The 16-channel forward kernel (k_conv_fwd16) pairs the taps
    differently from the general one (so that one input fragment serves
    several output rows), so the FP32 sums are taken in another order: both
    outputs lie within the bf16 bar of the float64 direct convolution, and
    nearly all elements agree to one bf16 ulp (outputs that cancel to near
    zero, or that the ReLU cuts on one side only, can differ by more), on a
    grid that is not a multiple of either tile."""
    n, d, h, w = 2, 5, 9, 37
    say("conv forward16 vs general", co)
    rng = np.random.default_rng(4)
    xt = bf(rng.standard_normal((n, d, h, w, ci)))
    W = (rng.standard_normal((co, 27, ci)) / 20).astype(np.float32)
    wp, _ = pack(W, 3)
    bt = torch.as_tensor(rng.standard_normal(co).astype(np.float32)).to(_dev())
    outs = []
    try:
        for k in (3, 2):   # bit 0: k_conv_fwd16; bit 1: transposed weight gradient (the default)
            ck(abx.load().abx_conv3d_set_kernel(k))
            y = torch.empty((n, d, h, w, co), dtype=torch.bfloat16, device=_dev())
            ck(abx.load().abx_conv3d_forward(stream(), xt.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), bt.data_ptr(),
                                             co, 3, 1, 0, y.data_ptr(), co))
            outs.append(y)
        torch.cuda.synchronize()
    finally:
        ck(abx.load().abx_conv3d_set_kernel(3))
    y16, yg = f32(outs[0]).astype(np.float64), f32(outs[1]).astype(np.float64)
    big = np.maximum(np.abs(y16), np.abs(yg))
    ulp = np.where(big > 0, 2.0 ** (np.floor(np.log2(np.where(big > 0, big, 1.0))) - 7), 0.0)
    assert (np.abs(y16 - yg) <= ulp).mean() > 0.99
    ref = R.direct_conv3d(f32(xt), f32(wp).reshape(co, 27, ci), bt.cpu().numpy(), 3, relu=True)
    close_bf16(y16, ref, "k_conv_fwd16 vs float64")
    close_bf16(yg, ref, "k_conv_fwd vs float64")


@pytest.mark.parametrize("case", [(2, 5, 9, 37, 16, 16, 1, 0), (1, 6, 5, 20, 48, 16, 0, 1), (1, 9, 7, 33, 16, 48, 1, 1),
                                  (1, 3, 6, 18, 32, 32, 0, 0), (1, 17, 13, 40, 16, 16, 0, 1)],
                         ids=lambda c: "x".join(map(str, c)))
def test_conv_forward16_brick_paths_equal(case):
    """k_conv_fwd16 loads its input brick and weights straight into LDS
    (global_load_lds) and holds the weights in registers (the default), or
    reads them from LDS for every MFMA (flag bit 6; three workgroups per CU,
    two with bit 4), or stages brick and weights through registers (flag bit
    3). The LDS contents and the MFMA order are the same,
    so the outputs are equal bit for bit: with and without accumulation and
    ReLU, over several 16-channel input chunks, on grids that are not
    multiples of the 16 x 4 x 8 tile (the zero halo comes from a zero source
    on the direct path)."""
    n, d, h, w, ci, co, relu, acc = case
    say("conv forward16 brick paths", case)
    rng = np.random.default_rng(12)
    xt = bf(rng.standard_normal((n, d, h, w, ci)))
    W = (rng.standard_normal((co, 27, ci)) / 20).astype(np.float32)
    wp, _ = pack(W, 3)
    bt = torch.as_tensor(rng.standard_normal(co).astype(np.float32)).to(_dev())
    y0 = bf(rng.standard_normal((n, d, h, w, co)))
    lib = abx.load()
    outs = []
    try:
        for flags in (3, 11, 67, 83):
            ck(lib.abx_conv3d_set_kernel(flags))
            y = y0.clone()
            ck(lib.abx_conv3d_forward(stream(), xt.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), bt.data_ptr(), co,
                                      3, relu, acc, y.data_ptr(), co))
            torch.cuda.synchronize()
            outs.append(y.view(torch.int16).cpu())
    finally:
        ck(lib.abx_conv3d_set_kernel(3))
    assert torch.equal(outs[0], outs[1])
    assert torch.equal(outs[0], outs[2])
    assert torch.equal(outs[0], outs[3])
    ref = R.direct_conv3d(f32(xt), f32(wp).reshape(co, 27, ci), bt.cpu().numpy(), 3, relu=False)
    if acc:
        ref = ref + f32(y0)
    if relu:
        ref = np.maximum(ref, 0)
    close_bf16(f32(outs[0].view(torch.bfloat16)), ref, "k_conv_fwd16 (LDS-direct brick) vs float64")


@pytest.mark.parametrize("case", [(2, 5, 9, 37, 16, 16, 3), (1, 6, 5, 20, 32, 16, 3), (1, 4, 8, 16, 16, 32, 3),
                                  (1, 3, 7, 33, 32, 32, 3), (2, 5, 4, 17, 16, 32, 1), (1, 3, 6, 18, 32, 64, 3),
                                  (1, 3, 5, 16, 16, 64, 3), (1, 2, 3, 16, 32, 64, 1)],
                         ids=lambda c: "x".join(map(str, c)))
def test_wgrad_transposed_reads_equal_gathered(case):
    """The tap-group weight gradient on transposing LDS reads
    (k_conv_wgrad_tr; the default for 1x1x1, selected for 3x3x3 with flag
    bit 2) builds the same fragments as the 16-bit gathering kernel and
    issues the same MFMAs in the same order: equal bit for bit, weights and
    bias, on grids that are not multiples of the tile, for 16, 32 and 64
    output channels per workgroup."""
    n, d, h, w, ci, co, k = case
    say("wgrad transposed reads vs gathered", case)
    rng = np.random.default_rng(6)
    xt = bf(rng.standard_normal((n, d, h, w, ci)))
    dzt = bf(rng.standard_normal((n, d, h, w, co)))
    lib = abx.load()
    ws_bytes = lib.abx_conv3d_wgrad_workspace(n, d, h, w, ci, co, k)
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=_dev())
    outs = []
    try:
        for flags in (7, 5):
            ck(lib.abx_conv3d_set_kernel(flags))
            dw = torch.empty(co * k ** 3 * ci, dtype=torch.float32, device=_dev())
            db = torch.empty(co, dtype=torch.float32, device=_dev())
            ck(lib.abx_conv3d_wgrad(stream(), xt.data_ptr(), ci, dzt.data_ptr(), co, n, d, h, w, ci, co, k,
                                    ws.data_ptr(), ws_bytes, dw.data_ptr(), db.data_ptr(), 0))
            torch.cuda.synchronize()
            outs.append((dw.cpu(), db.cpu()))
    finally:
        ck(lib.abx_conv3d_set_kernel(3))
    assert torch.equal(outs[0][0].view(torch.int32), outs[1][0].view(torch.int32))
    assert torch.equal(outs[0][1].view(torch.int32), outs[1][1].view(torch.int32))


@pytest.mark.parametrize("case", [(2, 5, 9, 37, 16, 16, 3, 3, 0), (1, 4, 6, 20, 32, 48, 3, 0, 1),
                                  (1, 3, 5, 18, 16, 64, 3, 3, 1), (2, 5, 4, 17, 16, 16, 1, 3, 0),
                                  (1, 9, 7, 21, 48, 32, 3, 3, 0), (1, 6, 5, 19, 32, 16, 3, 67, 0)],
                         ids=lambda c: "x".join(map(str, c)))
def test_conv_forward_gated_equals_forward_then_relu_backward(case):
    """abx_conv3d_forward_gated (the ReLU backward folded into the store)
    equals abx_conv3d_forward followed by abx_relu_backward bit for bit, on
    the 16-channel-tile kernel and the general one, with and without
    accumulation, the gate a channel slice of a wider tensor; the 16-channel
    kernel stages the gate values in LDS with the brick of the last input
    chunk (one, two or three chunks here), or with flag bit 6 reads them from
    global memory."""
    n, d, h, w, ci, co, k, flags, acc = case
    say("conv forward gated", case)
    rng = np.random.default_rng(11)
    xt = bf(rng.standard_normal((n, d, h, w, ci)))
    W = (rng.standard_normal((co, k ** 3, ci)) / 20).astype(np.float32)
    wp, _ = pack(W, k)
    gate_wide = bf(rng.standard_normal((n, d, h, w, co + 16)))     # half of the values not positive
    gate = gate_wide[..., 8:8 + co]
    y0 = bf(rng.standard_normal((n, d, h, w, co)))
    lib = abx.load()
    try:
        ck(lib.abx_conv3d_set_kernel(flags))
        ya, yb = y0.clone(), y0.clone()
        ck(lib.abx_conv3d_forward(stream(), xt.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), None, co, k, 0, acc,
                                  ya.data_ptr(), co))
        ck(lib.abx_relu_backward(stream(), ya.data_ptr(), co, gate.data_ptr(), co + 16, n * d * h * w, co))
        ck(lib.abx_conv3d_forward_gated(stream(), xt.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), None, co, k, 0,
                                        acc, yb.data_ptr(), co, gate.data_ptr(), co + 16))
        torch.cuda.synchronize()
    finally:
        ck(lib.abx_conv3d_set_kernel(3))
    assert torch.equal(ya.view(torch.int16), yb.view(torch.int16))
    assert (f32(yb) == 0).mean() > 0.4


def test_pool_and_upsample_gated_backward_equal_backward_then_relu():
    """abx_maxpool2_backward_gated (x is the ReLU output) and
    abx_upsample2_backward_gated equal the plain gradient kernel followed by
    abx_relu_backward, bit for bit, on channel slices of wider tensors."""
    say("pool / upsample gated backward")
    rng = np.random.default_rng(13)
    n, d, h, w, c = 2, 4, 6, 8, 16
    lib = abx.load()
    xw = bf(rng.standard_normal((n, d, h, w, c + 8)))
    x = xw[..., 8:]
    dy = bf(rng.standard_normal((n, d // 2, h // 2, w // 2, c)))
    dx0 = bf(rng.standard_normal((n, d, h, w, c)))
    da, db = dx0.clone(), dx0.clone()
    ck(lib.abx_maxpool2_backward(stream(), x.data_ptr(), c + 8, dy.data_ptr(), c, n, d, h, w, c, da.data_ptr(), c, 1))
    ck(lib.abx_relu_backward(stream(), da.data_ptr(), c, x.data_ptr(), c + 8, n * d * h * w, c))
    ck(lib.abx_maxpool2_backward_gated(stream(), x.data_ptr(), c + 8, dy.data_ptr(), c, n, d, h, w, c, db.data_ptr(),
                                       c, 1))
    gw = bf(rng.standard_normal((n, d, h, w, c + 8)))
    gate = gw[..., :c]
    fine = bf(rng.standard_normal((n, 2 * d, 2 * h, 2 * w, c)))
    ua = torch.empty((n, d, h, w, c), dtype=torch.bfloat16, device=_dev())
    ub = torch.empty_like(ua)
    ck(lib.abx_upsample2_backward(stream(), fine.data_ptr(), c, n, d, h, w, c, ua.data_ptr(), c))
    ck(lib.abx_relu_backward(stream(), ua.data_ptr(), c, gate.data_ptr(), c + 8, n * d * h * w, c))
    ck(lib.abx_upsample2_backward_gated(stream(), fine.data_ptr(), c, n, d, h, w, c, ub.data_ptr(), c,
                                        gate.data_ptr(), c + 8))
    torch.cuda.synchronize()
    assert torch.equal(da.view(torch.int16), db.view(torch.int16))
    assert torch.equal(ua.view(torch.int16), ub.view(torch.int16))
    assert (f32(db) == 0).mean() > 0.3 and (f32(ub) == 0).mean() > 0.3


def test_pack_all_equals_per_layer_pack():
    """abx_conv3d_pack_all (every layer, both layouts, one launch) writes what
    abx_conv3d_pack writes layer by layer, bit for bit."""
    say("pack all layers in one launch")
    rng = np.random.default_rng(14)
    lib = abx.load()
    shapes = [(16, 3, 16), (32, 3, 48), (64, 1, 16), (16, 3, 32)]
    ws, ref, got = [], [], []
    descs = (abx.PackDesc * len(shapes))()
    for i, (co, k, ci) in enumerate(shapes):
        w = torch.as_tensor((rng.standard_normal(co * k ** 3 * ci) / 10).astype(np.float32)).to(_dev())
        ws.append(w)
        pair_ref = [torch.empty(co * k ** 3 * ci, dtype=torch.bfloat16, device=_dev()) for _ in range(2)]
        for flip in (0, 1):
            ck(lib.abx_conv3d_pack(stream(), w.data_ptr(), co, k, ci, flip, pair_ref[flip].data_ptr()))
        pair = [torch.empty_like(pair_ref[0]) for _ in range(2)]
        descs[i] = abx.PackDesc(w.data_ptr(), pair[0].data_ptr(), pair[1].data_ptr(), co, k ** 3, ci)
        ref.append(pair_ref)
        got.append(pair)
    ck(lib.abx_conv3d_pack_all(stream(), len(shapes), descs))
    torch.cuda.synchronize()
    for r, g in zip(ref, got):
        for a, b in zip(r, g):
            assert torch.equal(a.view(torch.int16), b.view(torch.int16))


def test_mse_vector_kernel_equals_generic():
    """The MSE loss on 16-byte rows (k_mse_v) and the generic kernel (taken
    when the gradient buffer is not 16-byte aligned) give the same loss
    partials and gradients bit for bit; unused channels get zero gradient."""
    say("mse vector vs generic")
    rng = np.random.default_rng(12)
    nvox, C, ps = 5000, 4, 16
    pred = bf(rng.standard_normal((nvox, ps)))
    tgt = bf(rng.standard_normal((nvox, C)))
    mask = torch.as_tensor((rng.random(nvox) < 0.7).astype(np.uint8)).to(_dev())
    cw = torch.as_tensor(np.array([1.0, 0.5, 2.0, 1.5], np.float32)).to(_dev())
    lib = abx.load()
    nparts = lib.abx_mse_partials(nvox)
    outs = []
    for off in (0, 1):                                # off 1: dpred 2 bytes past a 16-byte boundary
        lp = torch.empty(nparts, dtype=torch.float32, device=_dev())
        buf = torch.full((nvox * ps + 8,), 7.0, dtype=torch.bfloat16, device=_dev())
        dp = buf[off:off + nvox * ps]
        ck(lib.abx_mse_loss(stream(), pred.data_ptr(), ps, tgt.data_ptr(), C, mask.data_ptr(), nvox, C, cw.data_ptr(),
                            1.0 / (nvox * C), lp.data_ptr(), dp.data_ptr(), ps))
        torch.cuda.synchronize()
        outs.append((lp.cpu(), dp.cpu().view(torch.int16).reshape(nvox, ps)))
    assert torch.equal(outs[0][0].view(torch.int32), outs[1][0].view(torch.int32))
    assert torch.equal(outs[0][1], outs[1][1])
    assert (outs[0][1][:, C:] == 0).all()


def test_conv_forward_slices_and_accumulate():
    """Input and output are channel slices of wider tensors; accumulate adds."""
    n, d, h, w, ci, co = 1, 4, 4, 16, 32, 16
    say("conv forward slices")
    rng = np.random.default_rng(3)
    big_in = bf(rng.standard_normal((n, d, h, w, ci + 24)))
    big_out = bf(rng.standard_normal((n, d, h, w, co + 16)))
    before = f32(big_out)
    W = (rng.standard_normal((co, 27, ci)) / 20).astype(np.float32)
    wp, _ = pack(W, 3)
    xoff, yoff = 16, 8
    ck(abx.load().abx_conv3d_forward(stream(), big_in.data_ptr() + 2 * xoff, ci + 24, n, d, h, w, ci, wp.data_ptr(),
                                     None, co, 3, 0, 1, big_out.data_ptr() + 2 * yoff, co + 16))
    torch.cuda.synchronize()
    after = f32(big_out)
    X = f32(big_in)[..., xoff:xoff + ci]
    ref = R.direct_conv3d(X, f32(wp).reshape(co, 27, ci)) + before[..., yoff:yoff + co]
    close_bf16(after[..., yoff:yoff + co], ref, "accumulated slice")
    # channels outside the slice are untouched
    assert np.array_equal(after[..., :yoff], before[..., :yoff])
    assert np.array_equal(after[..., yoff + co:], before[..., yoff + co:])


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv_backward_matches_autograd(case):
    n, d, h, w, ci, co, k = case
    say("conv backward", case)
    rng = np.random.default_rng(4)
    xt = bf(rng.standard_normal((n, d, h, w, ci)))
    W = (rng.standard_normal((co, k ** 3, ci)) / np.sqrt(k ** 3 * ci)).astype(np.float32)
    dzt = bf(rng.standard_normal((n, d, h, w, co)))
    wp, _ = pack(W, k)
    wb, _ = pack(W, k, flip=True)
    lib = abx.load()
    # input gradient: forward over dZ with the flipped transpose
    dx = torch.empty((n, d, h, w, ci), dtype=torch.bfloat16, device=_dev())
    ck(lib.abx_conv3d_forward(stream(), dzt.data_ptr(), co, n, d, h, w, co, wb.data_ptr(), None, ci, k, 0, 0,
                              dx.data_ptr(), ci))
    # weight gradient
    ws_bytes = lib.abx_conv3d_wgrad_workspace(n, d, h, w, ci, co, k)
    assert ws_bytes > 0
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=_dev())
    dw = torch.empty(co * k ** 3 * ci, dtype=torch.float32, device=_dev())
    db = torch.empty(co, dtype=torch.float32, device=_dev())
    ck(lib.abx_conv3d_wgrad(stream(), xt.data_ptr(), ci, dzt.data_ptr(), co, n, d, h, w, ci, co, k, ws.data_ptr(),
                            ws_bytes, dw.data_ptr(), db.data_ptr(), 0))
    torch.cuda.synchronize()
    import torch.nn.functional as F
    Xr = xt.float().permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    Wr = R.to_torch_weight(wp.float().reshape(co, k ** 3, ci), k).requires_grad_(True)
    Br = torch.zeros(co, device=_dev(), requires_grad=True)
    out = F.conv3d(Xr, Wr, Br, padding=k // 2)
    out.backward(dzt.float().permute(0, 4, 1, 2, 3).contiguous())
    close_bf16(f32(dx), Xr.grad.permute(0, 2, 3, 4, 1).cpu().numpy(), f"input grad {case}")
    ref_dw = Wr.grad.permute(0, 2, 3, 4, 1).reshape(co, k ** 3, ci).cpu().numpy()
    got_dw = dw.cpu().numpy().reshape(co, k ** 3, ci)
    rng_w = np.abs(ref_dw).max()
    err = np.abs(got_dw - ref_dw).max()
    assert err <= 1e-3 * rng_w, f"weight grad {case}: max err {err:.3e} vs range {rng_w:.3e}"
    ref_db = Br.grad.cpu().numpy()
    assert np.abs(db.cpu().numpy() - ref_db).max() <= 1e-3 * max(np.abs(ref_db).max(), 1e-30)
    # accumulate adds to the previous contents (old + sum of partials: FP32 order differs from 2 * sum)
    ck(lib.abx_conv3d_wgrad(stream(), xt.data_ptr(), ci, dzt.data_ptr(), co, n, d, h, w, ci, co, k, ws.data_ptr(),
                            ws_bytes, dw.data_ptr(), None, 1))
    torch.cuda.synchronize()
    np.testing.assert_allclose(dw.cpu().numpy().reshape(co, k ** 3, ci), 2 * got_dw, rtol=0, atol=1e-5 * rng_w)
    # deterministic: the same call gives the same bits
    ck(lib.abx_conv3d_wgrad(stream(), xt.data_ptr(), ci, dzt.data_ptr(), co, n, d, h, w, ci, co, k, ws.data_ptr(),
                            ws_bytes, dw.data_ptr(), None, 0))
    torch.cuda.synchronize()
    assert np.array_equal(dw.cpu().numpy().reshape(co, k ** 3, ci), got_dw)


@pytest.mark.parametrize("case", [(2, 5, 9, 37, 16, 16), (1, 6, 5, 20, 32, 16), (1, 5, 7, 18, 48, 16),
                                  (2, 4, 8, 16, 16, 32), (1, 3, 7, 33, 32, 32), (1, 3, 7, 33, 96, 32),
                                  (1, 3, 6, 18, 16, 64), (1, 3, 5, 16, 32, 64), (1, 3, 3, 12, 192, 64),
                                  (1, 2, 2, 8, 64, 128), (8, 12, 6, 8, 128, 128), (8, 61, 30, 50, 16, 16)],
                         ids=lambda c: "x".join(map(str, c)))
def test_wgrad27_every_channel_blocking(case):
    """k_wgrad27 (every tap in one workgroup, the 3x3x3 default) for every
    (output, input) channel blocking it is compiled for, on grids that are
    not multiples of its 16x4x4 tile (the last case has enough tiles for the
    16x4x8 tiles of large grids): weights and bias within 1e-3 of the
    range of torch's FP32 autograd on the same bf16 operands, within the
    same bar of the tap-group kernel and of k_wgrad27 staging its tiles
    through registers (flag bit 5; the default loads them straight into LDS
    where the blocking allows more workgroups per CU, so the partial sums
    may be grouped in more rows), and the same bits on a second call."""
    n, d, h, w, ci, co = case
    say("wgrad27", case)
    rng = np.random.default_rng(8)
    xt = bf(rng.standard_normal((n, d, h, w, ci)))
    dzt = bf(rng.standard_normal((n, d, h, w, co)))
    lib = abx.load()
    ws_bytes = lib.abx_conv3d_wgrad_workspace(n, d, h, w, ci, co, 3)
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=_dev())

    def run(flags):
        ck(lib.abx_conv3d_set_kernel(flags))
        dw = torch.empty(co * 27 * ci, dtype=torch.float32, device=_dev())
        db = torch.empty(co, dtype=torch.float32, device=_dev())
        ck(lib.abx_conv3d_wgrad(stream(), xt.data_ptr(), ci, dzt.data_ptr(), co, n, d, h, w, ci, co, 3,
                                ws.data_ptr(), ws_bytes, dw.data_ptr(), db.data_ptr(), 0))
        torch.cuda.synchronize()
        return dw.cpu().numpy().reshape(co, 27, ci), db.cpu().numpy()

    try:
        dw, db = run(3)
        dw2, db2 = run(3)
        dw_old, db_old = run(7)
        dw_st, db_st = run(35)
    finally:
        ck(lib.abx_conv3d_set_kernel(3))
    assert np.array_equal(dw, dw2) and np.array_equal(db, db2)
    import torch.nn.functional as F
    Xr = xt.float().permute(0, 4, 1, 2, 3).contiguous()
    Wr = torch.zeros((co, ci, 3, 3, 3), device=_dev(), requires_grad=True)
    Br = torch.zeros(co, device=_dev(), requires_grad=True)
    F.conv3d(Xr, Wr, Br, padding=1).backward(dzt.float().permute(0, 4, 1, 2, 3).contiguous())
    ref_dw = Wr.grad.permute(0, 2, 3, 4, 1).reshape(co, 27, ci).cpu().numpy()
    ref_db = Br.grad.cpu().numpy()
    rng_w, rng_b = np.abs(ref_dw).max(), max(np.abs(ref_db).max(), 1e-30)
    assert np.abs(dw - ref_dw).max() <= 1e-3 * rng_w, f"weights {case}: {np.abs(dw - ref_dw).max():.3e} of {rng_w:.3e}"
    assert np.abs(db - ref_db).max() <= 1e-3 * rng_b, f"bias {case}"
    assert np.abs(dw - dw_old).max() <= 1e-3 * rng_w and np.abs(db - db_old).max() <= 1e-3 * rng_b
    assert np.abs(dw - dw_st).max() <= 1e-3 * rng_w and np.abs(db - db_st).max() <= 1e-3 * rng_b


@pytest.mark.parametrize("case", [(2, 5, 9, 37, 32, 16, 16), (1, 6, 5, 20, 64, 32, 32), (1, 3, 5, 18, 128, 64, 64),
                                  (2, 4, 6, 16, 16, 32, 16)],
                         ids=lambda c: "x".join(map(str, c)))
def test_conv_parts_equal_the_concatenated_tensor(case):
    """A concatenation held as two tensors (abx_conv3d_forward_parts,
    abx_conv3d_wgrad_parts) against the same channels in one tensor: the
    forward and the input gradient written to two parts equal the one-tensor
    results bit for bit; the weight and bias gradients agree within 1e-3 of
    the range (their partial sums are grouped by input block, which differs
    when 48 or 32 channels no longer divide both parts)."""
    n, d, h, w, c1, c2, co = case
    ci = c1 + c2
    say("conv parts", case)
    rng = np.random.default_rng(9)
    lib = abx.load()
    xa = bf(rng.standard_normal((n, d, h, w, c1)))
    xb = bf(rng.standard_normal((n, d, h, w, c2)))
    xc = torch.cat([xa, xb], -1).contiguous()
    W = (rng.standard_normal((co, 27, ci)) / np.sqrt(27 * ci)).astype(np.float32)
    wp, _ = pack(W, 3)
    wb, _ = pack(W, 3, flip=True)
    bias = torch.as_tensor(rng.standard_normal(co).astype(np.float32)).to(_dev())
    # forward: [xa | xb] -> y
    y1 = torch.empty((n, d, h, w, co), dtype=torch.bfloat16, device=_dev())
    y2 = torch.empty_like(y1)
    ck(lib.abx_conv3d_forward(stream(), xc.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), bias.data_ptr(), co, 3, 1,
                              0, y1.data_ptr(), co))
    ck(lib.abx_conv3d_forward_parts(stream(), xa.data_ptr(), c1, c1, xb.data_ptr(), c2, n, d, h, w, ci, wp.data_ptr(),
                                    bias.data_ptr(), co, 1, 0, y2.data_ptr(), co, co, None, 0))
    # input gradient: dz -> [dxa | dxb]
    dz = bf(rng.standard_normal((n, d, h, w, co)))
    dxc = torch.empty((n, d, h, w, ci), dtype=torch.bfloat16, device=_dev())
    dxa = torch.empty((n, d, h, w, c1), dtype=torch.bfloat16, device=_dev())
    dxb = torch.empty((n, d, h, w, c2), dtype=torch.bfloat16, device=_dev())
    ck(lib.abx_conv3d_forward(stream(), dz.data_ptr(), co, n, d, h, w, co, wb.data_ptr(), None, ci, 3, 0, 0,
                              dxc.data_ptr(), ci))
    ck(lib.abx_conv3d_forward_parts(stream(), dz.data_ptr(), co, co, None, 0, n, d, h, w, co, wb.data_ptr(), None, ci,
                                    0, 0, dxa.data_ptr(), c1, c1, dxb.data_ptr(), c2))
    # weight gradient
    ws_bytes = lib.abx_conv3d_wgrad_workspace(n, d, h, w, ci, co, 3)
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=_dev())
    outs = []
    for parts in (False, True):
        dw = torch.empty(co * 27 * ci, dtype=torch.float32, device=_dev())
        db = torch.empty(co, dtype=torch.float32, device=_dev())
        if parts:
            ck(lib.abx_conv3d_wgrad_parts(stream(), xa.data_ptr(), c1, c1, xb.data_ptr(), c2, dz.data_ptr(), co, n, d,
                                          h, w, ci, co, ws.data_ptr(), ws_bytes, dw.data_ptr(), db.data_ptr(), 0))
        else:
            ck(lib.abx_conv3d_wgrad(stream(), xc.data_ptr(), ci, dz.data_ptr(), co, n, d, h, w, ci, co, 3,
                                    ws.data_ptr(), ws_bytes, dw.data_ptr(), db.data_ptr(), 0))
        torch.cuda.synchronize()
        outs.append((dw.cpu().numpy(), db.cpu().numpy()))
    assert torch.equal(y1.view(torch.int16), y2.view(torch.int16))
    assert torch.equal(dxc[..., :c1].contiguous().view(torch.int16), dxa.view(torch.int16))
    assert torch.equal(dxc[..., c1:].contiguous().view(torch.int16), dxb.view(torch.int16))
    (w1_, b1_), (w2_, b2_) = outs
    assert np.abs(w2_ - w1_).max() <= 1e-3 * np.abs(w1_).max()
    assert np.abs(b2_ - b1_).max() <= 1e-3 * np.abs(b1_).max()
    # parts are refused where they are not supported: a gate, a 1x1x1 convolution, an unaligned split
    assert lib.abx_conv3d_wgrad_parts(stream(), xa.data_ptr(), c1, 8, xb.data_ptr(), c2, dz.data_ptr(), co, n, d, h,
                                      w, ci, co, ws.data_ptr(), ws_bytes, dw.data_ptr(), None, 0) != 0
    assert lib.abx_conv3d_forward_parts(stream(), xa.data_ptr(), c1, c1, None, 0, n, d, h, w, ci, wp.data_ptr(),
                                        None, co, 0, 0, y2.data_ptr(), co, co, None, 0) != 0


def test_wgrad_workspace_too_small_is_an_error():
    say("wgrad workspace check")
    lib = abx.load()
    n, d, h, w, ci, co = 1, 2, 4, 16, 16, 16
    need = lib.abx_conv3d_wgrad_workspace(n, d, h, w, ci, co, 3)
    x = torch.zeros((n, d, h, w, ci), dtype=torch.bfloat16, device=_dev())
    ws = torch.empty(need // 4, dtype=torch.float32, device=_dev())
    dw = torch.empty(co * 27 * ci, dtype=torch.float32, device=_dev())
    rc = lib.abx_conv3d_wgrad(stream(), x.data_ptr(), ci, x.data_ptr(), co, n, d, h, w, ci, co, 3, ws.data_ptr(),
                              need - 4, dw.data_ptr(), None, 0)
    assert rc != 0 and b"workspace" in lib.abx_last_error()


def _distinct_windows(n, d, h, w, c, rng):
    """bf16-exact values, distinct inside every 2x2x2 window."""
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    pos = ((z & 1) * 4 + (y & 1) * 2 + (x & 1))[None, ..., None]
    base = rng.integers(0, 24, size=(n, d, h, w, c))
    sign = np.where(rng.random((n, d, h, w, c)) < 0.5, -1, 1)
    return (sign * (base * 8 + pos)).astype(np.float32)


@pytest.mark.parametrize("c", [24, 16, 32, 64])
def test_pool_and_upsample_match_torch(c):
    """Pooling and upsampling against torch; 16, 32 and 64 channels take the
    x-pair-lane kernels, 24 the per-voxel ones."""
    import torch.nn.functional as F
    n, d, h, w = 2, 4, 6, 10
    say("pool / upsample", c)
    rng = np.random.default_rng(5)
    lib = abx.load()
    X = bf(_distinct_windows(n, d, h, w, c, rng))
    Y = torch.empty((n, d // 2, h // 2, w // 2, c), dtype=torch.bfloat16, device=_dev())
    ck(lib.abx_maxpool2(stream(), X.data_ptr(), c, n, d, h, w, c, Y.data_ptr(), c))
    dY = bf(rng.standard_normal((n, d // 2, h // 2, w // 2, c)))
    dX = bf(rng.standard_normal((n, d, h, w, c)))      # accumulate onto this
    dX0 = f32(dX)
    ck(lib.abx_maxpool2_backward(stream(), X.data_ptr(), c, dY.data_ptr(), c, n, d, h, w, c, dX.data_ptr(), c, 1))
    U = torch.empty((n, 2 * d, 2 * h, 2 * w, c), dtype=torch.bfloat16, device=_dev())
    ck(lib.abx_upsample2(stream(), X.data_ptr(), c, n, d, h, w, c, U.data_ptr(), c))
    dU = bf(rng.standard_normal((n, 2 * d, 2 * h, 2 * w, c)))
    dXu = torch.empty_like(X)
    ck(lib.abx_upsample2_backward(stream(), dU.data_ptr(), c, n, d, h, w, c, dXu.data_ptr(), c))
    torch.cuda.synchronize()
    Xr = X.float().permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    P = F.max_pool3d(Xr, 2)
    assert np.array_equal(f32(Y), P.detach().permute(0, 2, 3, 4, 1).cpu().numpy())
    P.backward(dY.float().permute(0, 4, 1, 2, 3).contiguous())
    close_bf16(f32(dX), dX0 + Xr.grad.permute(0, 2, 3, 4, 1).cpu().numpy(), "maxpool backward")
    Xr2 = X.float().permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    Ur = F.interpolate(Xr2, scale_factor=2, mode="nearest")
    assert np.array_equal(f32(U), Ur.detach().permute(0, 2, 3, 4, 1).cpu().numpy())
    Ur.backward(dU.float().permute(0, 4, 1, 2, 3).contiguous())
    close_bf16(f32(dXu), Xr2.grad.permute(0, 2, 3, 4, 1).cpu().numpy(), "upsample backward")


@pytest.mark.parametrize("c", [16, 24, 32])
def test_maxpool_backward_ties_go_to_the_first_maximum(c):
    """With ties in every window (values 0, 1, 2), the gradient goes to the
    first maximum in the fine order (z, y, x), whichever kernel runs; the
    rest of the window gets zero."""
    n, d, h, w = 2, 4, 6, 8
    say("maxpool backward ties", c)
    rng = np.random.default_rng(17)
    xv = rng.integers(0, 3, (n, d, h, w, c)).astype(np.float32)
    gv = rng.standard_normal((n, d // 2, h // 2, w // 2, c)).astype(np.float32)
    X, dY = bf(xv), bf(gv)
    gv = f32(dY)
    dX = torch.empty_like(X)
    ck(abx.load().abx_maxpool2_backward(stream(), X.data_ptr(), c, dY.data_ptr(), c, n, d, h, w, c, dX.data_ptr(), c,
                                        0))
    torch.cuda.synchronize()
    ref = np.zeros((n, d, h, w, c), np.float32)
    win = xv.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).transpose(0, 1, 3, 5, 2, 4, 6, 7)
    win = win.reshape(n, d // 2, h // 2, w // 2, 8, c)
    first = win.argmax(axis=4)          # numpy returns the first maximum
    for o in range(8):
        a, b, e = o >> 2, (o >> 1) & 1, o & 1
        ref[:, a::2, b::2, e::2, :] = np.where(first == o, gv, 0.0)
    assert np.array_equal(f32(dX), ref)


def _bfr(a):
    """Round float64 values to bf16 (nearest even) and back."""
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).double().numpy()


def test_unet_forward_backward_stepwise():
    """Every stored activation and gradient buffer of the hand-written forward
    and backward is recomputed in float64 from the network's own stored inputs
    (so rounding differences cannot cascade) and must agree to the bf16 bar:
    this checks each kernel call and all the wiring (slices, skips, masks,
    which buffer feeds which)."""
    from airboxcalc_amd import unet as U
    say("unet stepwise forward/backward")
    net = U.UNet(seed=3)
    rng = np.random.default_rng(6)
    n, d, h, w = 2, 8, 16, 24
    x = bf(rng.standard_normal((n, d, h, w, U.IN_CH)))
    net.forward(x)
    dout = bf(np.concatenate([rng.standard_normal((n, d, h, w, U.OUT_CH)),
                              np.zeros((n, d, h, w, U.HEAD_CH - U.OUT_CH))], -1))
    net.backward(dout)
    torch.cuda.synchronize()
    A = {k: f32(v).astype(np.float64) for k, v in net.A.items()}
    G = {k: f32(v).astype(np.float64) for k, v in net.G.items()}
    X, DO = f32(x).astype(np.float64), f32(dout).astype(np.float64)
    Wm = net.w.cpu().numpy().astype(np.float64)
    Gw = net.g.cpu().numpy().astype(np.float64)
    L = net.by
    Wb = {nm: f32(net.packs[nm][0]).astype(np.float64).reshape(l.co, l.taps, l.ci) for nm, l in L.items()}
    B = {nm: Wm[l.b_off:l.b_off + l.co] for nm, l in L.items()}
    w1, w2, w3, w4 = U.WIDTHS

    def conv(nm, xin):
        l = L[nm]
        return _bfr(R.direct_conv3d(xin, Wb[nm], B[nm], l.ksize, l.relu))

    def dgrad(nm, dz):
        l = L[nm]
        flipped = np.transpose(Wb[nm][:, ::-1, :], (2, 1, 0))
        return _bfr(R.direct_conv3d(dz, flipped, None, l.ksize))

    def check(tag, got, ref, ulps=1):
        rng_ = max(np.abs(ref).max(), 1e-30)
        tol = ulps * 2.0 ** -7 * np.abs(ref) + 1e-5 * rng_     # one bf16 ulp is at most 2^-7 |x|
        bad = np.abs(got - ref) > tol
        assert not bad.any(), f"{tag}: {bad.sum()} of {bad.size} off, max err {np.abs(got - ref).max():.3e}"

    cat = lambda D, k: np.concatenate([D["up" + k], D["e" + k]], -1)   # [upsampled | skip]
    c1, c2, c3 = cat(A, "1"), cat(A, "2"), cat(A, "3")
    e1, e2, e3 = c1[..., w2:], c2[..., w3:], c3[..., w4:]
    # forward
    check("a1", A["a1"], conv("enc1a", X))
    check("e1", e1, conv("enc1b", A["a1"]))
    check("p1", A["p1"], R.pool_ref(e1)[0], 0)
    check("a2", A["a2"], conv("enc2a", A["p1"]))
    check("e2", e2, conv("enc2b", A["a2"]))
    check("p2", A["p2"], R.pool_ref(e2)[0], 0)
    check("a3", A["a3"], conv("enc3a", A["p2"]))
    check("e3", e3, conv("enc3b", A["a3"]))
    check("p3", A["p3"], R.pool_ref(e3)[0], 0)
    check("a4", A["a4"], conv("bota", A["p3"]))
    check("b4", A["b4"], conv("botb", A["a4"]))
    check("up3", c3[..., :w4], R.up_ref(A["b4"]), 0)
    check("d3a", A["d3a"], conv("dec3a", c3))
    check("d3b", A["d3b"], conv("dec3b", A["d3a"]))
    check("up2", c2[..., :w3], R.up_ref(A["d3b"]), 0)
    check("d2a", A["d2a"], conv("dec2a", c2))
    check("d2b", A["d2b"], conv("dec2b", A["d2a"]))
    check("up1", c1[..., :w2], R.up_ref(A["d2b"]), 0)
    check("d1a", A["d1a"], conv("dec1a", c1))
    check("d1b", A["d1b"], conv("dec1b", A["d1a"]))
    check("out", A["out"], conv("head", A["d1b"]))
    # backward: activation gradients (ReLU-masked where the network masks them)
    m = lambda a: (a > 0).astype(np.float64)
    gc1, gc2, gc3 = cat(G, "1"), cat(G, "2"), cat(G, "3")
    check("g d1b", G["d1b"], dgrad("head", DO) * m(A["d1b"]))
    check("g d1a", G["d1a"], dgrad("dec1b", G["d1b"]) * m(A["d1a"]))
    ref_c1 = dgrad("dec1a", G["d1a"])
    check("g up1", gc1[..., :w2], ref_c1[..., :w2])
    check("g d2b", G["d2b"], _bfr(R.up_bwd_ref(gc1[..., :w2])) * m(A["d2b"]))
    check("g d2a", G["d2a"], dgrad("dec2b", G["d2b"]) * m(A["d2a"]))
    ref_c2 = dgrad("dec2a", G["d2a"])
    check("g up2", gc2[..., :w3], ref_c2[..., :w3])
    check("g d3b", G["d3b"], _bfr(R.up_bwd_ref(gc2[..., :w3])) * m(A["d3b"]))
    check("g d3a", G["d3a"], dgrad("dec3b", G["d3b"]) * m(A["d3a"]))
    ref_c3 = dgrad("dec3a", G["d3a"])
    check("g up3", gc3[..., :w4], ref_c3[..., :w4])
    check("g b4", G["b4"], _bfr(R.up_bwd_ref(gc3[..., :w4])) * m(A["b4"]))
    check("g a4", G["a4"], dgrad("botb", G["b4"]) * m(A["a4"]))
    check("g p3", G["p3"], dgrad("bota", G["a4"]))
    # skip gradient + pooling gradient, then the ReLU mask
    check("g e3", gc3[..., w4:], _bfr(ref_c3[..., w4:] + R.pool_bwd_ref(e3, G["p3"])) * m(e3), 2)
    check("g a3", G["a3"], dgrad("enc3b", gc3[..., w4:]) * m(A["a3"]))
    check("g p2", G["p2"], dgrad("enc3a", G["a3"]))
    check("g e2", gc2[..., w3:], _bfr(ref_c2[..., w3:] + R.pool_bwd_ref(e2, G["p2"])) * m(e2), 2)
    check("g a2", G["a2"], dgrad("enc2b", gc2[..., w3:]) * m(A["a2"]))
    check("g p1", G["p1"], dgrad("enc2a", G["a2"]))
    check("g e1", gc1[..., w2:], _bfr(ref_c1[..., w2:] + R.pool_bwd_ref(e1, G["p1"])) * m(e1), 2)
    check("g a1", G["a1"], dgrad("enc1b", gc1[..., w2:]) * m(A["a1"]))
    # weight and bias gradients from the same stored tensors (FP32 sums)
    pairs = {"head": (A["d1b"], DO), "dec1b": (A["d1a"], G["d1b"]), "dec1a": (c1, G["d1a"]),
             "dec2b": (A["d2a"], G["d2b"]), "dec2a": (c2, G["d2a"]), "dec3b": (A["d3a"], G["d3b"]),
             "dec3a": (c3, G["d3a"]), "botb": (A["a4"], G["b4"]), "bota": (A["p3"], G["a4"]),
             "enc3b": (A["a3"], gc3[..., w4:]), "enc3a": (A["p2"], G["a3"]), "enc2b": (A["a2"], gc2[..., w3:]),
             "enc2a": (A["p1"], G["a2"]), "enc1b": (A["a1"], gc1[..., w2:]), "enc1a": (X, G["a1"])}
    for nm, (xin, dz) in pairs.items():
        l = L[nm]
        rw, rb = R.wgrad_ref(xin, dz, l.ksize)
        gw = Gw[l.w_off:l.w_off + l.nw].reshape(rw.shape)
        gb = Gw[l.b_off:l.b_off + l.co]
        assert np.abs(gw - rw).max() <= 1e-5 * max(np.abs(rw).max(), 1e-30), nm
        assert np.abs(gb - rb).max() <= 1e-5 * max(np.abs(rb).max(), 1e-30), nm


def test_unet_close_to_plain_fp32_network():
    """End to end against the FP32 network in torch: bf16 storage costs a few
    parts in a thousand of the output and, through ReLU masks that flip near
    zero, a few percent of the weight gradients."""
    from airboxcalc_amd import unet as U
    say("unet vs FP32 network")
    net = U.UNet(seed=3)
    rng = np.random.default_rng(6)
    n, d, h, w = 2, 8, 16, 24
    x = bf(rng.standard_normal((n, d, h, w, U.IN_CH)))
    out = net.forward(x)
    dout = bf(np.concatenate([rng.standard_normal((n, d, h, w, U.OUT_CH)),
                              np.zeros((n, d, h, w, U.HEAD_CH - U.OUT_CH))], -1))
    net.backward(dout)
    torch.cuda.synchronize()
    Wn = net.w.cpu().numpy()
    P = {}
    for l in net.L:
        ww = torch.as_tensor(Wn[l.w_off:l.w_off + l.nw].reshape(l.co, l.taps, l.ci)).to(torch.bfloat16).double()
        P[l.name] = (ww.requires_grad_(True), torch.as_tensor(Wn[l.b_off:l.b_off + l.co]).double().requires_grad_(True))
    ref = R.torch_unet(x.double().cpu(), P, net.L)
    o, r = f32(out), ref.detach().numpy()
    rel = float(np.linalg.norm(o - r) / np.linalg.norm(r))
    assert rel < 1e-2, rel
    ref.backward(dout.double().cpu())
    G = net.g.cpu().numpy()
    for l in net.L:
        gw = G[l.w_off:l.w_off + l.nw]
        rw = P[l.name][0].grad.numpy().ravel()
        if l.name == "head":
            gw, rw = gw.reshape(l.co, -1)[:U.OUT_CH].ravel(), rw.reshape(l.co, -1)[:U.OUT_CH].ravel()
        cos = float(gw @ rw / (np.linalg.norm(gw) * np.linalg.norm(rw)))
        assert cos > 0.98, (l.name, cos)


def test_unet_fused_head_loss_equals_separate_kernels():
    """train_step's head, loss and head backward in one kernel
    (abx_head_loss_backward) against the head convolution, abx_mse_loss and
    the head's gated input and weight gradients as separate kernels, on a
    mask that leaves 30 % of the volume out and unequal channel weights: the
    loss within 1e-4 (relative), the decoder output's gradient within the
    bf16 bar, the head's weight and bias gradients within 1e-3 of their range
    and all the parameters' gradients within 1e-2 (relative L2: the head's
    FP32 sums run in another order, so an output may round to the next bf16
    value); and the same bits on a second fused pass."""
    from airboxcalc_amd import unet as U
    say("unet fused head and loss")
    rng = np.random.default_rng(13)
    n, d, h, w = 2, 16, 16, 24
    X = np.zeros((n, d, h, w, U.IN_CH), np.float32)
    X[..., :5] = rng.standard_normal((n, d, h, w, 5))
    x, y = bf(X), bf(0.5 * rng.standard_normal((n, d, h, w, U.OUT_CH)))
    mask = torch.as_tensor((rng.random((n, d, h, w)) < 0.7).astype(np.uint8)).to(_dev())
    cw = torch.tensor([1.0, 0.5, 2.0, 1.5], dtype=torch.float32, device=_dev())
    cnt = int(mask.sum().item())
    net = U.UNet(seed=4)
    net.forward(x)
    loss_s, dout = net.loss_and_grad(y, mask, cw, count=cnt)
    net.backward(dout)
    torch.cuda.synchronize()
    ls, g_s, gx_s = float(loss_s.item()), net.g.clone(), net.G["d1b"].clone()
    runs = []
    for _ in range(2):
        net.forward(x, head=False)
        loss_f = net.head_loss_backward(y, mask, cw, count=cnt)
        net.backward(None)
        torch.cuda.synchronize()
        runs.append((float(loss_f.item()), net.g.clone(), net.G["d1b"].clone()))
    lf, g_f, gx_f = runs[0]
    assert runs[1][0] == lf
    assert torch.equal(runs[1][1].view(torch.int32), g_f.view(torch.int32))
    assert torch.equal(runs[1][2].view(torch.int16), gx_f.view(torch.int16))
    assert abs(lf - ls) <= 1e-4 * abs(ls), (lf, ls)
    close_bf16(f32(gx_f), f32(gx_s), "gradient of the decoder output")
    hl = net.by["head"]
    gs_np, gf_np = g_s.cpu().numpy().astype(np.float64), g_f.cpu().numpy().astype(np.float64)
    for off, cnt_ in ((hl.w_off, hl.nw), (hl.b_off, hl.co)):
        a, b = gf_np[off:off + cnt_], gs_np[off:off + cnt_]
        assert np.abs(a - b).max() <= 1e-3 * max(np.abs(b).max(), 1e-30), "head gradient"
    assert np.linalg.norm(gf_np - gs_np) <= 1e-2 * np.linalg.norm(gs_np)


def test_unet_side_stream_weight_gradients_equal_one_stream():
    """The weight gradients run on a second stream (each after an event on
    the main stream, joined before Adam) by default: three training steps
    give the same parameters bit for bit as with every kernel on the main
    stream."""
    from airboxcalc_amd import unet as U
    say("unet side-stream weight gradients")
    rng = np.random.default_rng(9)
    n, d, h, w = 2, 16, 16, 24
    X = np.zeros((n, d, h, w, U.IN_CH), np.float32)
    X[..., :5] = rng.standard_normal((n, d, h, w, 5))
    x, y = bf(X), bf(0.5 * rng.standard_normal((n, d, h, w, U.OUT_CH)))
    mask = torch.ones((n, d, h, w), dtype=torch.uint8, device=_dev())
    cw = torch.ones(U.OUT_CH, dtype=torch.float32, device=_dev())
    ws = []
    for overlap in (True, False):
        net = U.UNet(seed=3)
        net.overlap = overlap
        for _ in range(3):
            net.train_step(x, y, mask, cw, lr=1e-3, count=n * d * h * w)
        torch.cuda.synchronize()
        ws.append(net.w.cpu())
    assert torch.equal(ws[0].view(torch.int32), ws[1].view(torch.int32))


def test_unet_training_reduces_loss():
    from airboxcalc_amd import unet as U
    say("unet training smoke")
    net = U.UNet(seed=0)
    rng = np.random.default_rng(7)
    n, d, h, w = 4, 8, 16, 16
    X = np.zeros((n, d, h, w, U.IN_CH), np.float32)
    X[..., :5] = rng.standard_normal((n, d, h, w, 5))
    # a target the network can represent: a smoothed linear map of the inputs
    Y = np.stack([X[..., 0] + 0.5 * X[..., 1], X[..., 2] - X[..., 3], 0.3 * X[..., 4], X[..., 0] * 0.2], -1)
    x, y = bf(X), bf(Y)
    mask = torch.ones((n, d, h, w), dtype=torch.uint8, device=_dev())
    cw = torch.ones(U.OUT_CH, dtype=torch.float32, device=_dev())
    losses = [float(net.train_step(x, y, mask, cw, lr=1e-3).item()) for _ in range(60)]
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0], losses[::10]
