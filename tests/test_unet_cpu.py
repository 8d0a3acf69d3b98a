"""CPU checks of the surrogate's references and host logic (no GPU calls)."""
from __future__ import annotations

import numpy as np
import pytest

from tests import unet_ref as R

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("ksize", [1, 3])
def test_direct_conv_matches_torch_cpu(ksize):
    """Pins the brute-force reference: tap order (dz, dy, dx) and padding."""
    import torch.nn.functional as F
    rng = np.random.default_rng(0)
    n, d, h, w, ci, co = 2, 3, 4, 5, 6, 7
    x = rng.standard_normal((n, d, h, w, ci))
    W = rng.standard_normal((co, ksize ** 3, ci))
    b = rng.standard_normal(co)
    got = R.direct_conv3d(x, W, b, ksize, relu=True)
    ref = F.relu(F.conv3d(torch.from_numpy(x).permute(0, 4, 1, 2, 3), R.to_torch_weight(torch.from_numpy(W), ksize),
                          torch.from_numpy(b), padding=ksize // 2)).permute(0, 2, 3, 4, 1).numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_flipped_transpose_gives_the_input_gradient():
    """The input gradient is a forward convolution over dZ with W'[ci][t][co] =
    W[co][26-t][ci] (the packing the library does for for_input_grad)."""
    rng = np.random.default_rng(1)
    n, d, h, w, ci, co = 1, 3, 4, 5, 4, 3
    x = torch.from_numpy(rng.standard_normal((n, d, h, w, ci))).requires_grad_(True)
    W = rng.standard_normal((co, 27, ci))
    dz = rng.standard_normal((n, d, h, w, co))
    import torch.nn.functional as F
    y = F.conv3d(x.permute(0, 4, 1, 2, 3), R.to_torch_weight(torch.from_numpy(W), 3), padding=1)
    y.backward(torch.from_numpy(dz).permute(0, 4, 1, 2, 3))
    Wf = np.transpose(W[:, ::-1, :], (2, 1, 0))       # [ci][t][co], tap 26-t
    got = R.direct_conv3d(dz, Wf, ksize=3)
    np.testing.assert_allclose(got, x.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_unet_layer_table():
    from airboxcalc_amd import unet as U
    L = U.layers()
    assert [l.name for l in L][:2] == ["enc1a", "enc1b"] and L[-1].name == "head"
    # parameters are packed back to back, weights then bias, per layer
    off = 0
    for l in L:
        assert l.w_off == off and l.b_off == off + l.nw
        off = l.b_off + l.co
        assert l.ci % 16 == 0 and l.co % 16 == 0
    assert U.n_params() == off
    by = {l.name: l for l in L}
    # decoder inputs are [upsampled | skip]
    assert by["dec3a"].ci == by["botb"].co + by["enc3b"].co
    assert by["dec2a"].ci == by["dec3b"].co + by["enc2b"].co
    assert by["dec1a"].ci == by["dec2b"].co + by["enc1b"].co
    assert [by[f"enc{i}b"].co for i in (1, 2, 3)] + [by["botb"].co] == list(U.WIDTHS)


def test_pool_upsample_wgrad_references_match_torch_cpu():
    """Pins the numpy references of the other network kernels against torch."""
    import torch.nn.functional as F
    rng = np.random.default_rng(2)
    n, d, h, w, c = 2, 4, 6, 8, 3
    # distinct values inside every window, so the maximum is unique
    x = rng.permutation(n * d * h * w * c).reshape(n, d, h, w, c).astype(np.float64)
    xt = torch.from_numpy(x).permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    p = F.max_pool3d(xt, 2)
    dy = rng.standard_normal(p.shape)
    p.backward(torch.from_numpy(dy))
    np.testing.assert_array_equal(R.pool_ref(x)[0], p.detach().permute(0, 2, 3, 4, 1).numpy())
    np.testing.assert_allclose(R.pool_bwd_ref(x, np.transpose(dy, (0, 2, 3, 4, 1))),
                               xt.grad.permute(0, 2, 3, 4, 1).numpy(), rtol=0, atol=0)
    ut = torch.from_numpy(x).permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    u = F.interpolate(ut, scale_factor=2, mode="nearest")
    du = rng.standard_normal(u.shape)
    u.backward(torch.from_numpy(du))
    np.testing.assert_array_equal(R.up_ref(x), u.detach().permute(0, 2, 3, 4, 1).numpy())
    np.testing.assert_allclose(R.up_bwd_ref(np.transpose(du, (0, 2, 3, 4, 1))), ut.grad.permute(0, 2, 3, 4, 1).numpy(),
                               rtol=1e-12, atol=1e-12)
    for k in (1, 3):
        ci, co = 4, 5
        W = rng.standard_normal((co, k ** 3, ci))
        xi = torch.from_numpy(rng.standard_normal((n, d, h, w, ci))).permute(0, 4, 1, 2, 3).contiguous()
        Wt = R.to_torch_weight(torch.from_numpy(W), k).requires_grad_(True)
        bt = torch.zeros(co, dtype=torch.float64, requires_grad=True)
        y = F.conv3d(xi, Wt, bt, padding=k // 2)
        dz = rng.standard_normal(y.shape)
        y.backward(torch.from_numpy(dz))
        dw, db = R.wgrad_ref(xi.permute(0, 2, 3, 4, 1).numpy(), np.transpose(dz, (0, 2, 3, 4, 1)), k)
        np.testing.assert_allclose(dw, Wt.grad.permute(0, 2, 3, 4, 1).reshape(co, k ** 3, ci).numpy(), rtol=1e-10,
                                   atol=1e-10)
        np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-10, atol=1e-10)
