"""The Node.js drop-in (airboxcalc_amd/js/engine_hip.js over the N-API addon)
against the reference's own outputs (engine.js golden fixtures).

With the f32_f64 kernel the whole chain prepare -> simulateHIP -> result is
compared with what engine.js's prepare/simulate produced: masks and
coordinates bit for bit, per-cell maps and streamlines bit for bit, metrics
to 1e-12 relative (GPU tree reductions), hist likewise."""
import json
import os
import subprocess

import numpy as np
import pytest

from airboxcalc_amd import geometry as G
from tests import golden_util as GU

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "tests", "js", "run_hip_case.js")


def _run(tmp_path, name, precision="f32_f64", selftest=False, run=True):
    fx = GU.load(name)
    d = fx["desc"]
    stl = str(tmp_path / "box.stl")
    (G.write_stl_ascii if d.get("ascii") else G.write_stl_binary)(stl, G.airbox_triangles(G.spec_from_dict(d["spec"])))
    case = {"stl": stl, "prepare": d["prepare"], "run": d.get("run") if run else None, "precision": precision,
            "selftest": selftest}
    cj = str(tmp_path / "case.json")
    with open(cj, "w") as fh:
        json.dump(case, fh)
    print("[gpu] node case", name, precision, flush=True)
    p = subprocess.run(["node", RUNNER, cj], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return fx, json.loads(p.stdout)


@pytest.mark.parametrize("name", ["small", "small_yaw", "c1_seed3"])
def test_node_simulate_matches_engine_js(tmp_path, name):
    fx, out = _run(tmp_path, name, selftest=True)
    assert out["selftest"] < 1e-3
    for k in ("tris", "solid", "fluid", "opening", "outlet", "xc", "dc", "zc", "openPts", "filterPts", "outletPts"):
        assert out["sha"][k] == fx["sha"][k], k
    P = fx["prepare"]
    for k in ("nx", "ny", "nz", "N", "nf", "top", "jm"):
        assert out["g"][k] == P[k], k
    for k in ("volume_l", "inlet_cm2", "filter_cm2", "outlet_cm2", "outlet_cm2_real"):
        assert out["hex"][k] == P["summary"][k]["hex"], k
    sim = fx["sim"]["sim"]
    for k, e in sim["metrics"].items():
        assert out["metrics"][k] == pytest.approx(GU.x64(e), rel=1e-12, abs=1e-14), k
    want = np.array([[GU.x64(e) for e in row] for row in sim["hist"]])
    got = np.array(out["hist"], dtype=float)
    assert np.array_equal(got[:, 0], want[:, 0])
    np.testing.assert_allclose(got[:, 1:], want[:, 1:], rtol=1e-12, atol=1e-15)
    assert [p["step"] for p in out["progress"]] == list(want[:, 0].astype(int))
    for pl in ("planInlet", "planLow", "lon"):
        for k in ("speed", "ua", "ub"):
            assert out["sha"][f"{pl}_{k}"] == fx["sha"][f"sim_{pl}_{k}"], (pl, k)
    assert out["sha"]["vol_spd"] == fx["sha"]["sim_vol_spd"]
    for k in ("pts", "spd", "off"):
        assert out["sha"][f"lines_{k}"] == fx["sha"][f"sim_lines_{k}"], k
    assert out["device_result"].startswith("GPU · ")


def test_node_f32_run_and_prepare_errors(tmp_path):
    fx, out = _run(tmp_path, "small", precision="f32")
    sim = fx["sim"]["sim"]
    assert out["metrics"]["gammaCore"] == pytest.approx(GU.x64(sim["metrics"]["gammaCore"]), abs=5e-4)
    for name in ("err_grid", "err_noopen", "err_lid", "err_outlet"):
        fx, out = _run(tmp_path, name, run=False)
        assert out["error"] == fx["errors"]["prepare"], name


def test_worker_protocol_run_matches_direct_call(tmp_path):
    """The HIP worker (airboxcalc_amd/js/worker_hip.js) runs a case through the
    reference's message protocol; its result message equals runHIP called
    directly, bit for bit, and carries the reference's device/status texts."""
    fx = GU.load("small_yaw")
    d = fx["desc"]
    stl = str(tmp_path / "box.stl")
    G.write_stl_binary(stl, G.airbox_triangles(G.spec_from_dict(d["spec"])))
    cj = str(tmp_path / "case.json")
    with open(cj, "w") as fh:
        json.dump({"stl": stl, "prepare": d["prepare"], "run": d["run"]}, fh)
    print("[gpu] node worker case small_yaw", flush=True)
    p = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "run_worker_case.js"), cj], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = json.loads(p.stdout)
    assert out["gpuinfo"]["ok"] and out["gpuinfo"]["name"]
    assert out["loaded"]["transferred"] and out["prepared"]["cells"] == fx["prepare"]["nf"]
    assert not out["errors"], out["errors"]
    r, dr = out["result"], out["direct"]
    assert r["device"].startswith("GPU · ")
    assert out["status"] == ["Grafikkarte wird geprüft …", "Rechnet auf der Grafikkarte (" + out["gpuinfo"]["name"] + ")"]
    assert out["progress"] == len(r["hist"])
    assert r["metrics"] == dr["metrics"] and r["hist"] == dr["hist"]
    assert r["sha"] == dr["sha"]


def test_worker_warm_start_and_criterion_match_direct_call(tmp_path):
    """The worker's run message may ask for the surrogate warm start and the
    steady criterion (warmStart, criterion in the run parameters): its result
    equals runHIP called directly with the same options, bit for bit,
    including the criterion's windows."""
    fx = GU.load("small_yaw")
    d = fx["desc"]
    stl = str(tmp_path / "box.stl")
    G.write_stl_binary(stl, G.airbox_triangles(G.spec_from_dict(d["spec"])))
    run = dict(d["run"], warmStart="family", criterion={"window": 200, "maxSteps": 2000})
    cj = str(tmp_path / "case.json")
    with open(cj, "w") as fh:
        json.dump({"stl": stl, "prepare": d["prepare"], "run": run}, fh)
    print("[gpu] node worker warm start + criterion", flush=True)
    p = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "run_worker_case.js"), cj], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = json.loads(p.stdout)
    assert not out["errors"], out["errors"]
    r, dr = out["result"], out["direct"]
    assert r["steady"] is not None and r["steady"] == dr["steady"]
    assert r["steady"]["steps"] <= 2000 and len(r["steady"]["windows"]) == len(r["hist"])
    assert r["metrics"] == dr["metrics"] and r["hist"] == dr["hist"]
    assert r["sha"] == dr["sha"]

