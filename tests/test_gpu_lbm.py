"""Stream-collide on the GPU against the reference's CPU backend.

ABX_F32_F64 (float storage, double arithmetic in cpuStep's order) must give
engine.js's cpuStep populations bit for bit: the fixtures hold the SHA-256 of
engine.js's own Float32Arrays after 1, 2 and 12 steps.  ABX_F32 (the WGSL
kernel's float arithmetic) must pass the reference's own GPU/CPU gate: max
|f_GPU - f_CPU| < 1e-3 after 12 steps over every 7th element
(engine.js:597-611, 772).  ABX_F64 must equal the oracle's f64-storage step
bit for bit.
"""
import numpy as np
import pytest

from airboxcalc_amd import abx, engine
from oracle import oracle as O
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, grid_of_oracle, oracle_geometry, say, selftest_ns

pytestmark = pytest.mark.gpu

RUN = GU.names(lambda r: "setup" in r)


def _sim(fx, precision, B=1, ns=None):
    g = gpu_geometry(fx)
    r = fx["desc"]["run"]
    return g, abx.Sim(dev(), g.grid(), r["K"], r["aniso"], r["u"], r["tau"], precision=precision, B=B, ns=ns)


@pytest.mark.parametrize("name", RUN)
def test_setup_scalars(name):
    fx = GU.load(name)
    say("setup", name)
    _, s = _sim(fx, "f32")
    st = fx["setup"]
    I = s.info
    assert (I.NF, I.Ain, I.Af, I.nM, I.nT) == (st["NF"], st["Ain"], st["Af"], st["nM"], st["nT"])
    assert I.sig == GU.x64(st["sig"]) and I.cs2s == GU.x64(st["cs2s"])


@pytest.mark.parametrize("name", RUN)
def test_f32_f64_matches_engine_js_bit_for_bit(name):
    fx = GU.load(name)
    say("f32_f64 trajectories", name)
    og = oracle_geometry(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    ns = selftest_ns(S.flag)
    _, sim = _sim(fx, "f32_f64", B=2, ns=np.concatenate([np.zeros_like(ns), ns]))
    f0 = sim.read_f(0)
    assert np.array_equal(f0, S.f0())
    for s in range(1, 13):
        sim.steps(1, doacc=s > 6)
        if s in (1, 2, 12):
            for b, vn in enumerate(("plain", "ns")):
                assert GU.sha(sim.read_f(b)) == fx["sha"][f"f_{vn}_{s}"], (vn, s)
    for b, vn in enumerate(("plain", "ns")):
        assert GU.sha(sim.read_acc(1.0, b)) == fx["sha"][f"acc_{vn}"], vn


@pytest.mark.parametrize("name", RUN)
def test_f32_passes_reference_gate(name):
    fx = GU.load(name)
    say("f32 gate", name)
    og = oracle_geometry(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    ns = selftest_ns(S.flag)
    _, sim = _sim(fx, "f32", B=2, ns=np.concatenate([np.zeros_like(ns), ns]))
    sim.steps(12)
    md = 0.0
    for b, nsb in enumerate((None, ns)):
        ref = O.run_steps(S, 12, ns=nsb)
        got = sim.read_f(b)
        md = max(md, float(np.abs(got[::7].astype(np.float64) - ref[::7]).max()))
        # all elements, not only every 7th, and far inside the gate
        assert np.abs(got.astype(np.float64) - ref).max() < 1e-5
    assert md < 1e-3


@pytest.mark.parametrize("name", ["small", "small_yaw"])
def test_f64_matches_oracle_bit_for_bit(name):
    fx = GU.load(name)
    say("f64", name)
    og = oracle_geometry(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    _, sim = _sim(fx, "f64")
    acc_ref = np.zeros(4 * S.NF)
    f = S.f0(np.float64)
    f2 = np.empty_like(f)
    for s in range(1, 21):
        O.cpu_step(S, f, f2, acc_ref if s > 10 else None)
        f, f2 = f2, f
        sim.steps(1, doacc=s > 10)
    assert np.array_equal(sim.read_f(), f)
    assert np.array_equal(sim.read_acc(1.0, f64=True), acc_ref)


@pytest.mark.parametrize("precision", ["f32", "f32_f64", "f64"])
def test_probe_matches_oracle(precision):
    fx = GU.load("small")
    say("probe", precision)
    og = oracle_geometry(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    _, sim = _sim(fx, precision)
    sim.steps(200)
    f = sim.read_f().astype(np.float32)
    gam, mr, qo = sim.probe(200)
    g2, m2, q2 = O.probe(S, f, 200)
    # f32_f64 and f64 form u_y in double like probeCPU; f32 in float like the
    # WGSL probe (engine.js:455-461), so its u_y carries float rounding
    tol = {"f32": 1e-5, "f32_f64": 1e-12, "f64": 1e-6}[precision]
    assert gam == pytest.approx(g2, rel=tol) and mr == pytest.approx(m2, rel=tol)


@pytest.mark.parametrize("name", ["small", "small_noout", "small_yaw", "c1", "c1_seed3", "seed5_dx3"])
def test_simulate_metrics_equal_engine_js(name):
    """simulateGPU's loop on the f32_f64 kernel: the averaged acc is the CPU
    backend's, so evaluate gives engine.js's metrics exactly.  "c1" is the
    4 mm airbox with the reference's full protocol (2000 steps averaging the
    last 500, probes every 100)."""
    fx = GU.load(name)
    say("simulate", name)
    g = gpu_geometry(fx)
    r = fx["desc"]["run"]
    res = engine.simulate_hip(dev(), g, r, precision="f32_f64")
    og = oracle_geometry(fx)
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    ev = O.evaluate(og, S, r["u"], res.acc)
    sim = fx["sim"]["sim"]
    for k, e in sim["metrics"].items():
        assert ev["metrics"][k] == GU.x64(e), k
    want = np.array([[GU.x64(e) for e in row] for row in sim["hist"]])
    got = np.array(res.hist)
    assert np.array_equal(got[:, 0], want[:, 0])
    # the probe's sums are tree reductions on the GPU: rounding at 1e-16,
    # amplified where the mean filter velocity is still near 0 (seed5_dx3's
    # first probes, gamma ~ 19)
    np.testing.assert_allclose(got[:, 1:], want[:, 1:], rtol=1e-11, atol=1e-15)


def test_init_eq_cold_start_and_moments():
    fx = GU.load("small")
    say("init_eq")
    og = oracle_geometry(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    _, sim = _sim(fx, "f32_f64")
    NF = sim.NF
    ru = np.zeros(4 * NF, np.float32)
    ru[:NF] = 1
    sim.init_eq(ru)
    assert np.array_equal(sim.read_f(), S.f0())
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.95, 1.05, NF).astype(np.float32)
    u = rng.uniform(-0.05, 0.05, (3, NF)).astype(np.float32)
    sim.init_eq(np.concatenate([rho, u.ravel()]))
    f = sim.read_f().reshape(19, NF).astype(np.float64)
    cx = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
    cy = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
    cz = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
    np.testing.assert_allclose(f.sum(0), rho, rtol=2e-6)
    for c, comp in ((cx, 0), (cy, 1), (cz, 2)):
        np.testing.assert_allclose((c[:, None] * f).sum(0), rho * u[comp], atol=2e-7)


@pytest.mark.parametrize("precision", ["f32", "f32_f64"])
def test_write_read_round_trip_and_launch_variants_agree(precision):
    """Every launch variant of the stream-collide (AA pattern in one lattice,
    two lattices with pull streaming, every XCD block order, non-temporal
    population accesses) computes the same steps, bit for bit, on an
    arbitrary state with averaging on."""
    fx = GU.load("small_yaw")
    say("round trip / launch variants", precision)
    _, a = _sim(fx, precision)
    rng = np.random.default_rng(3)
    f = (np.repeat([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, a.NF) *
         rng.uniform(0.99, 1.01, 19 * a.NF)).astype(np.float32)
    a.write_f(f)
    assert np.array_equal(a.read_f(), f)
    a.steps(10, doacc=True)
    want_f, want_acc = a.read_f(), a.read_acc()
    for kern, xm, nt in ((4, 0, 3), (4, 1, 0), (4, 3, 3), (4, -1, 0), (4, -1, 3), (4, -1, 19 + 256 * 19),
                         (4, 8, 3 + 256 * 11), (4, -1, 256 * 8), (4, -1, 3 + 256 * 1), (4, -1, 3 + 256 * 2),
                         (4, -1, 3 + 256 * 16), (4, -1, 256 * 3), (4, 4, 19 + 256 * 32), (4, 16, 3 + 256 * 48),
                         (2, 0, 0), (2, 5, 0), (2, -1, 0)):
        _, b = _sim(fx, precision)
        b.set_launch(kern, xm, nt)
        b.write_f(f)
        b.steps(10, doacc=True)
        assert np.array_equal(b.read_f(), want_f), (kern, xm, nt)
        assert np.array_equal(b.read_acc(), want_acc), (kern, xm, nt)
        b.destroy()
    # occupancy caps (reserved LDS) change nothing but the launch
    for occ in ((0, 0), (8, 24), (32, 16), (4, 4)):
        _, b = _sim(fx, precision)
        b.set_occupancy(*occ)
        b.write_f(f)
        b.steps(10, doacc=True)
        assert np.array_equal(b.read_f(), want_f), occ
        assert np.array_equal(b.read_acc(), want_acc), occ
        b.destroy()
    # switching scheme mid-run (AA pattern <-> two lattices) carries the state over
    _, c = _sim(fx, precision)
    c.write_f(f)
    c.steps(3, doacc=True)
    c.set_launch(kernel=2)
    c.steps(4, doacc=True)
    c.set_launch(kernel=4, nt=19 + 256 * 11)
    c.steps(1, doacc=True)
    c.set_launch(kernel=2)
    c.steps(1, doacc=True)
    c.set_launch(kernel=4)
    c.steps(1, doacc=True)
    assert np.array_equal(c.read_f(), want_f)
    assert np.array_equal(c.read_acc(), want_acc)


def _ragged_grid(seed, nx, ny, nz):
    """A grid of random fluid inside a solid shell: rows with several fluid
    runs (solid cells inside a row's span), rows of one cell, rows without
    fluid, many rows per 64-position segment when nz is small; inlet cells at
    k = 0 under an opening, outlet cells, a porous layer."""
    rng = np.random.default_rng(seed)
    fl = np.zeros((nx, ny, nz), np.uint8)
    fl[1:-1, 1:-1, 1:-1] = rng.random((nx - 2, ny - 2, nz - 2)) < 0.7
    fl[3, :, :] = 0                                   # a plane without fluid
    fl[5, 4, :] = 0
    fl[5, 4, nz // 2] = 1                             # a row of one cell
    opening = np.zeros((nx, ny), np.uint8)
    opening[2:nx - 2, 2:ny - 2] = rng.random((nx - 4, ny - 4)) < 0.5
    fl[:, :, 0] = opening                             # inlet cells
    outlet = np.zeros_like(fl)
    outlet[:, ny - 2, 1:-1] = fl[:, ny - 2, 1:-1]
    dc = np.arange(ny, dtype=np.float32) * 2.0
    top, jm, pleat = 1, 4, 9.0
    return fl, opening, outlet, dc, top, jm, pleat


def _oracle_sim_from_arrays(fl, opening, outlet, dc, top, jm, pleat, K, aniso, u, tau):
    """The oracle's setupSim on a grid given as arrays (the arrays stay
    owned here)."""
    import ctypes as C
    nx, ny, nz = fl.shape
    g = O.GeomS()
    g.nx, g.ny, g.nz, g.N, g.nf, g.top, g.jm, g.pleat = nx, ny, nz, fl.size, int(fl.sum()), top, jm, pleat
    g.fluid = fl.ctypes.data_as(C.POINTER(C.c_uint8))
    g.opening = opening.ctypes.data_as(C.POINTER(C.c_uint8))
    g.outlet = outlet.ctypes.data_as(C.POINTER(C.c_uint8))
    g.dc = dc.ctypes.data_as(C.POINTER(C.c_float))
    S = O.SimS()
    assert O.lib().orc_setup_sim(C.byref(g), K, aniso, u, tau, C.byref(S)) == 0
    NF = S.NF
    return O.Sim(s=S, NF=NF, cid=None, cells=None, nb=None, flag=None, layerM=None, layerT=None, Ain=S.Ain,
                 Af=S.Af, Lp=S.Lp, sig=S.sig, cs2s=S.cs2s, u0=S.u0)


@pytest.mark.parametrize("seed,dims", [(1, (23, 19, 70)), (2, (17, 13, 9)), (3, (12, 40, 130))])
def test_row_span_layout_on_ragged_fluid(seed, dims):
    """The row-span lattice (rows stored from their first to their last fluid
    cell, one after another) on ragged random fluid: the reference's
    arithmetic equals the oracle's cpuStep bit for bit after even and odd
    steps with averaging, as do the two-lattice scheme and every XCD order,
    and the probe equals the oracle's."""
    fl, opening, outlet, dc, top, jm, pleat = _ragged_grid(seed, *dims)
    K, aniso, u, tau = 3.0, 10.0, 0.06, 0.51
    S = _oracle_sim_from_arrays(fl, opening, outlet, dc, top, jm, pleat, K, aniso, u, tau)
    assert S.Ain > 0 and S.Af > 0
    grid = abx.GridArrays(*dims, fl.ravel(), opening.ravel(), outlet.ravel(), dc, top, jm, pleat)
    rng = np.random.default_rng(seed)
    f = (S.f0().astype(np.float64) * rng.uniform(0.99, 1.01, 19 * S.NF)).astype(np.float32)
    f_ref = f.copy()
    f2 = np.empty_like(f_ref)
    acc_ref = np.zeros(4 * S.NF, np.float32)
    for st in range(1, 8):
        O.cpu_step(S, f_ref, f2, acc_ref if st > 3 else None)
        f_ref, f2 = f2, f_ref
    say("row-span layout", seed, dims)
    for kern, xm in ((4, -1), (4, 0), (4, 1), (4, 2), (2, -1), (2, 0)):
        sim = abx.Sim(dev(), grid, K, aniso, u, tau, precision="f32_f64")
        assert sim.NF == S.NF and sim.info.plane >= S.NF
        sim.set_launch(kern, xm)
        sim.write_f(f)
        sim.steps(3)
        sim.steps(4, doacc=True)
        assert np.array_equal(sim.read_f(), f_ref), (kern, xm)
        assert np.array_equal(sim.read_acc(1.0), acc_ref), (kern, xm)
        if S.s.nM and S.s.nT:
            assert sim.probe(7)[:2] == pytest.approx(O.probe(S, f_ref, 7)[:2], rel=1e-12, nan_ok=True)
        sim.destroy()


@pytest.mark.parametrize("name", ["small", "small_yaw", "c1_seed3"])
def test_gpu_evaluate_field3d_streamlines(name):
    """evaluate/field3d/streamlines on the GPU against engine.js's outputs on
    the same averaged field (f32_f64 kernel, so acc is engine.js's)."""
    fx = GU.load(name)
    say("evaluate", name)
    g = gpu_geometry(fx)
    r = fx["desc"]["run"]
    res = engine.simulate_hip(dev(), g, r, precision="f32_f64", keep_sim=True)
    sim = res.sim
    ev = sim.evaluate(g.eval_view(), divisor=max(res.nacc, 1))
    want = fx["sim"]["sim"]
    for k, e in want["metrics"].items():
        assert ev["metrics"][k] == pytest.approx(GU.x64(e), rel=1e-12, abs=1e-14), k
    arr = GU.arrays(name) if name != "c1_seed3" else None
    # per-cell maps: bit-exact where no sum enters, one float ulp where one does
    for pl in ("planInlet", "planLow", "lon"):
        for q, k in enumerate(("speed", "ua", "ub")):
            assert GU.sha(ev[pl][q]) == fx["sha"][f"sim_{pl}_{k}"], (pl, k)
    assert GU.sha(ev["vol"]["spd"]) == fx["sha"]["sim_vol_spd"]
    if arr is not None:
        for k in ("fmap",):
            np.testing.assert_array_max_ulp(ev["fmap"], arr["sim_fmap"], maxulp=1)
        np.testing.assert_array_max_ulp(np.nan_to_num(ev["vol"]["pst"]), np.nan_to_num(arr["sim_vol_pst"]), maxulp=2)
        np.testing.assert_array_max_ulp(np.nan_to_num(ev["vol"]["loss"]), np.nan_to_num(arr["sim_vol_loss"]),
                                        maxulp=2)
    L = sim.streamlines(g.eval_view(), divisor=max(res.nacc, 1))
    for k in ("pts", "spd", "off"):
        assert GU.sha(L[k]) == fx["sha"][f"sim_lines_{k}"], k
    sim.destroy()


def test_pattern_time_moves_populations_without_collision():
    """abx_sim_pattern_time runs the AA steps' loads and stores with the
    collision left out: it returns positive per-step times, and the fluid
    populations afterwards are the same values, moved (a streaming with
    bounce-back only permutes them)."""
    fx = GU.load("small_yaw")
    say("pattern time")
    _, sim = _sim(fx, "f32")
    rng = np.random.default_rng(21)
    f = rng.uniform(0.0, 1.0, 19 * sim.NF).astype(np.float32)
    sim.write_f(f)
    me, mo = sim.pattern_time(3)
    assert me > 0 and mo > 0
    g = sim.read_f()
    assert not np.array_equal(g, f)
    assert np.array_equal(np.sort(g), np.sort(f))
    sim.destroy()
