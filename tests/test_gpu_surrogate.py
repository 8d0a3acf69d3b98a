"""Surrogate warm start on the GPU: the device-side resample + equilibrium
(abx_sim_init_eq_resample) against the host path, and a warm-started solve."""
from __future__ import annotations

import numpy as np
import pytest

from airboxcalc_amd import abx
from airboxcalc_amd import geometry as G
from airboxcalc_amd import surrogate as S
from tests.gpu_util import say

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _net():
    from airboxcalc_amd import unet as U
    net = U.UNet(seed=0)
    norm, meta = S.load(S.WEIGHTS, net)
    return net, norm


def test_device_resample_matches_host_path():
    say("warm start: device resample vs host")
    dev = abx.device(0)
    net, norm = _net()
    box = S.family([1000002])[0]
    tris = G.airbox_triangles(box.spec)
    gf = abx.Geometry(dev, tris, G.prepare_params(box.spec, 3.0), max_cells=0)
    gc = abx.Geometry(dev, tris, G.prepare_params(box.spec, S.FAMILY_DX), max_cells=0)
    try:
        pred = S.predict(net, S.features(gc.grid(), box.K, box.aniso), norm)
        ru = S.resample(pred, gc, gf)
        a = abx.Sim(dev, gf.grid(), box.K, box.aniso, S.U0, S.TAU)
        a.init_eq(ru)
        b = abx.Sim(dev, gf.grid(), box.K, box.aniso, S.U0, S.TAU)
        S.warm_start_init(net, norm, dev, box, gf, b, tris)
        fa, fb = a.read_f(), b.read_f()
        assert np.isfinite(fb).all()
        np.testing.assert_allclose(fb, fa, rtol=0, atol=2e-6)
        np.testing.assert_allclose(b.read_rho_u(), ru, rtol=0, atol=2e-5)
        a.destroy()
        b.destroy()
    finally:
        gf.free()
        gc.free()


def test_warm_started_solve_reaches_steady():
    from airboxcalc_amd import steady
    say("warm start: held-out box to steady at the family cell size")
    dev = abx.device(0)
    net, norm = _net()
    box = S.family([1000005])[0]
    tris = G.airbox_triangles(box.spec)
    g = abx.Geometry(dev, tris, G.prepare_params(box.spec, S.FAMILY_DX), max_cells=0)
    try:
        p = dict(K=box.K, aniso=box.aniso, u=S.U0, tau=S.TAU)
        cold = steady.solve_to_steady(dev, g, p, window=1000, max_steps=20000)
        warm = steady.solve_to_steady(dev, g, p, window=1000, max_steps=20000,
                                      init_fn=lambda sim: S.warm_start_init(net, norm, dev, box, g, sim, tris))
        assert cold.converged and warm.converged
        assert warm.steps <= cold.steps
        assert abs(warm.metrics["gammaCore"] - cold.metrics["gammaCore"]) < 0.01
    finally:
        g.free()
