"""The C oracle (oracle/engine_oracle.c) against the reference's own outputs.

Every fixture under tests/golden/ was written by running src/engine.js under
Node (tests/golden/make_golden.py).  Integer, byte and float32-storage results
must match bit for bit (SHA-256 of the raw arrays); float64 scalars must be
the same IEEE value.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_util as GU

PREP_CASES = GU.names(lambda r: "prepare" in r)
RUN_CASES = GU.names(lambda r: "setup" in r)
SIM_CASES = GU.names(lambda r: "sim" in r)
ERR_CASES = GU.names(lambda r: "errors" in r)


def _geom(fx):
    d = fx["desc"]
    cy, sy = GU.x64(fx["frame"]["cy"]), GU.x64(fx["frame"]["sy"])
    return O.prepare(GU.tris_of(fx), d["prepare"], cy=cy, sy=sy)


@pytest.mark.parametrize("name", GU.names(lambda r: "tris" in r.get("sha", {})))
def test_stl_triangles(name):
    fx = GU.load(name)
    assert GU.sha(GU.tris_of(fx)) == fx["sha"]["tris"]


@pytest.mark.parametrize("name", GU.names(lambda r: "fine" in r.get("sha", {})))
def test_voxelize_bit_exact(name):
    fx = GU.load(name)
    v = fx["voxel"]
    d = fx["desc"]
    cy, sy = GU.x64(fx["frame"]["cy"]), GU.x64(fx["frame"]["sy"])
    g = O.prepare(GU.tris_of(fx), d["prepare"], cy=cy, sy=sy, max_cells=0) if "prepare" in fx else None
    if g is not None:
        assert GU.sha(g.T) == fx["sha"]["T"]
        T = g.T
    else:
        pytest.skip("prepare failed before the fine grid")
    origin = np.array([GU.x64(e) for e in v["origin"]])
    fine = O.voxelize(T, origin, v["nxf"], v["nyf"], v["nzf"], GU.x64(v["hf"]))
    assert GU.sha(fine) == fx["sha"]["fine"]
    assert GU.sha(g.fine) == fx["sha"]["fine"]


@pytest.mark.parametrize("name", PREP_CASES)
def test_prepare(name):
    fx = GU.load(name)
    g = _geom(fx)
    P = fx["prepare"]
    for k in ("nx", "ny", "nz", "N", "nf", "top", "jm"):
        assert getattr(g, k) == P[k], k
    assert [GU.x64(e) for e in P["mn"]] == list(g.mn)
    assert [GU.x64(e) for e in P["mx"]] == list(g.mx)
    for k in ("solid", "fluid", "opening", "outlet", "xc", "dc", "zc", "openPts", "filterPts", "outletPts"):
        assert GU.sha(getattr(g, k)) == fx["sha"][k], k
    S = P["summary"]
    for k in ("cells", "nHoles", "pleatOk", "grid"):
        assert g.summary[k] == S[k], k
    for k in ("inlet_cm2", "filter_cm2", "outlet_cm2", "outlet_cm2_real"):
        assert g.summary[k] == GU.x64(S[k]), k
    assert g.summary["volume_l"] == pytest.approx(GU.x64(S["volume_l"]), rel=1e-15)


@pytest.mark.parametrize("name", RUN_CASES)
def test_setup_and_12_step_trajectories(name):
    fx = GU.load(name)
    g = _geom(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(g, r["K"], r["aniso"], r["u"], r["tau"])
    st = fx["setup"]
    assert (S.NF, S.Ain, S.Af, len(S.layerM), len(S.layerT)) == (st["NF"], st["Ain"], st["Af"], st["nM"], st["nT"])
    assert S.sig == GU.x64(st["sig"]) and S.cs2s == GU.x64(st["cs2s"])
    for k in ("cid", "nb", "flag", "layerM", "layerT"):
        assert GU.sha(getattr(S, k)) == fx["sha"][k], k
    ns_st = np.zeros(S.NF, np.float32)
    idx = np.arange(0, S.NF, 17)
    ns_st[idx[S.flag[idx] == 0]] = 1
    # threads > 1: the multi-core CPU baseline's split of the same step
    for vn, ns, threads in (("plain", None, 1), ("ns", ns_st, 1), ("plain", None, 5), ("ns", ns_st, 3)):
        f = S.f0()
        f2 = np.empty_like(f)
        acc = np.zeros(4 * S.NF, np.float32)
        for s in range(1, 13):
            O.cpu_step(S, f, f2, acc if s > 6 else None, ns, threads=threads)
            f, f2 = f2, f
            if s in (1, 2, 12):
                assert GU.sha(f) == fx["sha"][f"f_{vn}_{s}"], (vn, s, threads)
        assert GU.sha(acc) == fx["sha"][f"acc_{vn}"], (vn, threads)
        gam, mr, _ = O.probe(S, f, 12)
        t = fx["traj"][vn]
        assert np.array_equal(np.array([gam, mr]), np.array([GU.x64(t["gamma"]), GU.x64(t["massRatio"])]),
                              equal_nan=True)


@pytest.mark.parametrize("name", SIM_CASES)
def test_simulate_evaluate(name):
    fx = GU.load(name)
    g = _geom(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(g, r["K"], r["aniso"], r["u"], r["tau"])
    for rn, sim in fx["sim"].items():
        ns = None
        if rn == "simns":
            ns = np.zeros(S.NF, np.float32)
            idx = np.arange(0, S.NF, 17)
            ns[idx[S.flag[idx] == 0]] = 1
        acc, hist = O.simulate(S, r["steps"], r["avg"], r["every"], ns)
        want_hist = np.array([[GU.x64(e) for e in row] for row in sim["hist"]])
        assert np.array_equal(hist, want_hist, equal_nan=True)
        ev = O.evaluate(g, S, r["u"], acc, ns)
        for k, e in sim["metrics"].items():
            assert ev["metrics"][k] == GU.x64(e) or (np.isnan(ev["metrics"][k]) and np.isnan(GU.x64(e))), k
        assert GU.sha(ev["fmap"]) == fx["sha"][f"{rn}_fmap"]
        for pl in ("planInlet", "planLow", "lon"):
            for q, k in enumerate(("speed", "ua", "ub")):
                assert GU.sha(ev[pl][q]) == fx["sha"][f"{rn}_{pl}_{k}"], (pl, k)
        vol = O.field3d(g, S, r["u"], acc, ev["fluid"], ev["rhoRef"])
        assert vol["kB"] == sim["vol"]["kB"]
        for k in ("spd", "pst", "loss"):
            assert GU.sha(vol[k]) == fx["sha"][f"{rn}_vol_{k}"], k
        L = O.streamlines(g, S, r["u"], acc, ev["fluid"])
        for k in ("pts", "spd", "off"):
            assert GU.sha(L[k]) == fx["sha"][f"{rn}_lines_{k}"], k


@pytest.mark.parametrize("name", ERR_CASES)
def test_prepare_errors(name):
    fx = GU.load(name)
    msg = fx["errors"]["prepare"]
    with pytest.raises(O.OracleError) as ei:
        _geom(fx)
    assert str(ei.value) == msg
