"""The benchmark workloads themselves, checked against the reference.

- The 4 mm airbox under the reference's full protocol (2000 steps averaging
  the last 500): the FP32 kernel's metrics within the stated tolerances of
  engine.js's own (gamma, gamma-core, lowCore +-0.005; zeta +-2 %).  The
  float-storage/double-arithmetic kernel equals engine.js exactly there
  (test_gpu_lbm.test_simulate_metrics_equal_engine_js).
- The 1 mm airbox of the throughput benchmark (seed 1, 10.5 M fluid cells in
  the row-span lattice with the default block order): 12 steps against the
  oracle's cpuStep, bit for bit in the reference's arithmetic, within the
  reference's gate in FP32.
- The 64 seeded batch variants (200-360 mm, dx 2-4 mm): batched members
  equal solo runs bit for bit, and members in the reference's arithmetic
  equal the oracle after 12 steps.
"""
import os

import numpy as np
import pytest

from airboxcalc_amd import abx, batch as BT, engine
from airboxcalc_amd import geometry as G
from oracle import oracle as O
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, say

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, os.cpu_count() or 1))


def _run_oracle(S, n, acc_from, f=None):
    """n cpuSteps on THREADS threads from f (default f0); the fields of the
    steps after acc_from accumulated.  Returns (f, acc)."""
    f = S.f0() if f is None else f.copy()
    f2 = np.empty_like(f)
    acc = np.zeros(4 * S.NF, np.float32)
    for s in range(1, n + 1):
        O.cpu_step(S, f, f2, acc if s > acc_from else None, threads=THREADS)
        f, f2 = f2, f
    return f, acc


@pytest.mark.parametrize("name", ["c1", "small", "small_yaw", "small_noout", "seed5_dx3"])
def test_f32_metrics_within_tolerances_of_engine_js(name):
    fx = GU.load(name)
    say("f32 metrics vs engine.js", name)
    g = gpu_geometry(fx)
    r = fx["desc"]["run"]
    res = engine.simulate_hip(dev(), g, r, precision="f32", keep_sim=True)
    ev = res.sim.evaluate(g.eval_view(), divisor=max(res.nacc, 1), field3d=False)["metrics"]
    res.sim.destroy()
    want = {k: GU.x64(v) for k, v in fx["sim"]["sim"]["metrics"].items()}
    d = {k: ev[k] - want[k] for k in ("gamma", "gammaCore", "lowCore", "zeta")}
    print("[gpu] f32 - engine.js:", {k: f"{v:.2e}" for k, v in d.items()}, flush=True)
    for k in ("gamma", "gammaCore", "lowCore"):
        assert abs(d[k]) <= 0.005, k
    assert abs(d["zeta"]) <= 0.02 * abs(want["zeta"])
    hist = np.array(res.hist)
    want_h = np.array([[GU.x64(e) for e in row] for row in fx["sim"]["sim"]["hist"]])
    assert np.array_equal(hist[:, 0], want_h[:, 0])
    # gamma of the probes once the flow has reached the filter (mass ratio
    # above 0.1); before, the mean filter velocity is near 0 and gamma is the
    # ratio of two vanishing sums
    on = want_h[:, 2] > 0.1
    if on.any():
        assert np.abs(hist[on, 1] - want_h[on, 1]).max() <= 0.005


def test_dx1_benchmark_lattice_matches_oracle():
    spec, _ = G.seeded_variant(1, "roofline")
    p = G.prepare_params(spec, 1.0, outlet=True, click=True)
    tris = G.airbox_triangles(spec)
    say("dx=1 mm airbox: prepare")
    g = abx.Geometry(dev(), tris, p, max_cells=0)
    og = O.prepare(tris, p, max_cells=0)
    assert (g.nx, g.ny, g.nz) == (og.nx, og.ny, og.nz) == (356, 202, 244)
    assert np.array_equal(g.array("fluid"), og.fluid)
    S = O.setup_sim(og, 3.0, 10.0, 0.06, 0.51)
    say("dx=1 mm airbox: oracle 12 steps on", THREADS, "threads")
    f_ref, acc_ref = _run_oracle(S, 12, 6)
    for precision in ("f32_f64", "f32"):
        say("dx=1 mm airbox:", precision)
        sim = abx.Sim(dev(), g.grid(), 3.0, 10.0, 0.06, 0.51, precision=precision)
        I = sim.info
        assert sim.NF == S.NF == 10493429
        # the row-span layout: a stored position per fluid cell plus the solid
        # cells inside the rows' spans and the final segment's padding
        assert S.NF <= I.plane < 1.02 * S.NF
        sim.steps(6)
        sim.steps(6, doacc=True)
        f = sim.read_f()
        if precision == "f32_f64":
            assert np.array_equal(f, f_ref)
            assert np.array_equal(sim.read_acc(1.0), acc_ref)
        else:
            d = np.abs(f.astype(np.float64) - f_ref)
            assert d[::7].max() < 1e-3       # the reference's gate (engine.js:772)
            assert d.max() < 1e-5
            np.testing.assert_allclose(sim.read_acc(1.0), acc_ref, rtol=0, atol=1e-5)
        sim.destroy()


def test_batch_variants_match_solo_runs_and_the_oracle():
    variants = BT.seeded_variants(range(64))
    say("64 batch variants: prepare")
    geoms = BT.prepare_all(dev(), variants)
    nf = np.array([gm.nf for gm in geoms])
    pick = {int(nf.argmin()), int(nf.argmax())}
    for dx in (2.0, 2.5, 3.0, 4.0):
        pick.add(next(i for i, v in enumerate(variants) if v.dx == dx))
    pick = sorted(pick)
    print("[gpu] members checked:", [(i, variants[i].dx, int(nf[i])) for i in pick], flush=True)
    plan = [(3, False), (4, False), (5, True), (2, True)]
    # every variant batched against the picked ones stepped alone
    sims = [abx.Sim(dev(), gm.grid(), K=v.K, aniso=v.aniso, u=BT.U0, tau=BT.TAU) for v, gm in zip(variants, geoms)]
    b = abx.Batch(dev(), sims)
    b.set_launch(4, False, True)
    say("64 batch variants: batched steps")
    for n, acc in plan:
        b.steps(n, acc)
    b.sync()
    for i in pick:
        say("solo member", i)
        s = abx.Sim(dev(), geoms[i].grid(), K=variants[i].K, aniso=variants[i].aniso, u=BT.U0, tau=BT.TAU)
        for n, acc in plan:
            s.steps(n, acc)
        assert np.array_equal(sims[i].read_f(), s.read_f()), i
        assert np.array_equal(sims[i].read_acc(), s.read_acc()), i
        s.destroy()
    b.destroy()
    for s in sims:
        s.destroy()
    # the picked members in the reference's arithmetic against the oracle
    members = [abx.Sim(dev(), geoms[i].grid(), K=variants[i].K, aniso=variants[i].aniso, u=BT.U0, tau=BT.TAU,
                       precision="f32_f64") for i in pick]
    b = abx.Batch(dev(), members)
    say("picked members f32_f64: batched steps")
    b.steps(6)
    b.steps(6, True)
    b.sync()
    for i, m in zip(pick, members):
        v = variants[i]
        og = O.prepare(G.airbox_triangles(v.spec), G.prepare_params(v.spec, v.dx), max_cells=0)
        S = O.setup_sim(og, v.K, v.aniso, BT.U0, BT.TAU)
        f_ref, acc_ref = _run_oracle(S, 12, 6)
        assert np.array_equal(m.read_f(), f_ref), i
        assert np.array_equal(m.read_acc(1.0), acc_ref), i
    b.destroy()
