"""abx_evaluate_metrics_noalloc against abx_evaluate, through the C ABI.

The allocation-free entry point keeps its working buffers in one fixed device
array per GPU.  Its metrics and filter map must equal abx_evaluate's bit for
bit (the same kernels on the same data), on two calls in a row where the
second has a smaller grid (the array still holds the first call's contents),
and grids too large for the array must be refused with an error."""
import ctypes as C

import numpy as np
import pytest

from airboxcalc_amd import abx, engine
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, say

pytestmark = pytest.mark.gpu


def _noalloc(sim, eg, divisor):
    M = abx.Metrics()
    fmap = np.full(eg.nx * eg.nz, -7.0, np.float32)
    rc = abx.load().abx_evaluate_metrics_noalloc(sim.h, C.byref(eg), 0, divisor, C.byref(M),
                                                 fmap.ctypes.data_as(C.c_void_p))
    abx._check(rc)
    return M, fmap


def _solved(name, precision):
    fx = GU.load(name)
    g = gpu_geometry(fx)
    r = dict(fx["desc"]["run"])
    res = engine.simulate_hip(dev(), g, r, precision=precision, keep_sim=True)
    return g, res.sim, max(res.nacc, 1)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_noalloc_metrics_equal_evaluate_on_two_calls(precision):
    lib = abx.load()
    cases = []
    for name in ("small_yaw", "small"):  # the second grid is the smaller one
        g, sim, div = _solved(name, precision)
        cases.append((name, g, sim, div))
    eg0, eg1 = cases[0][1].eval_view(), cases[1][1].eval_view()
    n0 = eg0.nx * eg0.ny * eg0.nz
    n1 = eg1.nx * eg1.ny * eg1.nz
    assert n1 < n0 and eg1.nx * eg1.nz < eg0.nx * eg0.nz
    for name, g, sim, div in cases:
        say("noalloc metrics", name, precision)
        eg = g.eval_view()
        want = sim.evaluate(eg, divisor=div, field3d=False)
        M, fmap = _noalloc(sim, eg, div)
        for k in abx.METRIC_KEYS:
            assert getattr(M, k) == want["metrics"][k] or (np.isnan(getattr(M, k)) and np.isnan(want["metrics"][k])), k
        for k in ("ic", "jc", "jlow", "jb", "kB"):
            assert getattr(M, k) == want[k], k
        assert M.rhoRef == want["rhoRef"] and M.pin == want["pin"]
        assert np.array_equal(fmap, want["fmap"], equal_nan=True)
        assert lib.abx_evaluate_metrics_noalloc_bytes(eg.nx, eg.ny, eg.nz) > 0
    for _, _, sim, _ in cases:
        sim.destroy()


def test_noalloc_refuses_a_grid_beyond_its_scratch():
    """A lattice whose layer needs more than the 8 MiB array: 1500 x 4 x 1500
    cells, all solid except a thin inlet channel, so the refusal comes before
    any launch."""
    lib = abx.load()
    nx, ny, nz = 1500, 4, 1500
    need = lib.abx_evaluate_metrics_noalloc_bytes(nx, ny, nz)
    assert need > 8 << 20
    fluid = np.zeros((nx, ny, nz), np.uint8)
    fluid[700:704, 1:3, 0:6] = 1
    opening = np.zeros((nx, ny), np.uint8)
    opening[700:704, 1:3] = 1
    outlet = np.zeros_like(fluid)
    dc = np.arange(ny, dtype=np.float32)
    grid = abx.GridArrays(nx, ny, nz, fluid.ravel(), opening.ravel(), outlet.ravel(), dc, 0, 2, 1.5)
    sim = abx.Sim(dev(), grid, precision="f32")
    xc = np.arange(nx, dtype=np.float32)
    zc = np.arange(nz, dtype=np.float32)
    eg = abx.EvalGeom(nx=nx, ny=ny, nz=nz, top=0, jm=2, h=1.0, pleat=1.5, mn_z=0.0,
                      opening=opening.ctypes.data_as(C.c_void_p), xc=xc.ctypes.data_as(C.c_void_p),
                      dc=dc.ctypes.data_as(C.c_void_p), zc=zc.ctypes.data_as(C.c_void_p), oAx=0, dAx=1, iAx=2,
                      sd=-1, si=-1, seal=0.0, cy=1.0, sy=0.0)
    say("noalloc refusal")
    with pytest.raises(abx.AbxError) as e:
        _noalloc(sim, eg, 1.0)
    assert "more than the 8388608 bytes available" in str(e.value)
    sim.destroy()
