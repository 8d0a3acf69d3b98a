"""Failure paths of the reference, on the GPU.

- Non-finite flow: progressInfo throws "Rechnung instabil bei Schritt N. ..."
  when the outflow sum over the layer above the seal is not finite
  (engine.js:335).  The GPU probe, the simulate loop and the Node drop-in must
  fail with that exact text at the step the reference fails at.
- The self-test gate: runSim throws "Abweichung GPU/CPU x.xe-y" when the GPU
  and CPU populations differ by the gate or more after 12 steps
  (engine.js:771-772); the HIP worker answers the run with that error.
The unstable runs use u = 0.3 and tau = 0.5 on the small box, which the
reference's cpuStep (the oracle) blows up at step 40 with probes every 10
steps.  NaN and Inf are ordinary float values on the GPU: no index depends on
them.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from airboxcalc_amd import abx, engine
from airboxcalc_amd import geometry as G
from oracle import oracle as O
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, oracle_geometry, say

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MSG = "Rechnung instabil bei Schritt {}. Zellgröße vergrößern oder Viskosität (tau) erhöhen."
UNSTABLE = dict(K=3.0, aniso=10.0, u=0.3, tau=0.5, steps=200, avg=50, every=10)


def _oracle_failure(og, run):
    S = O.setup_sim(og, run["K"], run["aniso"], run["u"], run["tau"])
    with pytest.raises(O.OracleError) as e:
        O.simulate(S, run["steps"], run["avg"], run["every"])
    return str(e.value)


def test_oracle_fails_like_the_reference():
    """The oracle's simulate raises the reference's message on this run (the
    oracle restates progressInfo, engine.js:330-339)."""
    fx = GU.load("small")
    msg = _oracle_failure(oracle_geometry(fx), UNSTABLE)
    assert msg == MSG.format(40)


@pytest.mark.parametrize("precision", ["f32_f64", "f32"])
def test_simulate_fails_at_the_reference_step(precision):
    fx = GU.load("small")
    say("unstable simulate", precision)
    want = _oracle_failure(oracle_geometry(fx), UNSTABLE)
    g = gpu_geometry(fx)
    steps = []
    with pytest.raises(abx.AbxError) as e:
        engine.simulate_hip(dev(), g, UNSTABLE, on_progress=lambda i: steps.append(i["step"]), precision=precision)
    # float storage with the reference's double arithmetic follows cpuStep bit
    # for bit, so it fails at the same probe; the FP32 kernel's rounding
    # differs at 1e-7, and a blow-up amplifies it, so its step may differ
    if precision == "f32_f64":
        assert str(e.value) == want
        assert steps == list(range(10, 40, 10))
    else:
        step = int(str(e.value).split("Schritt ")[1].split(".")[0])
        assert str(e.value) == MSG.format(step) and abs(step - 40) <= 20, str(e.value)


def test_probe_reports_a_non_finite_outflow_at_its_step():
    """NaN written into one cell of the layer above the seal (the outflow
    layer, engine.js:333-335) makes the next probe fail, naming the step it
    was called with; the oracle's probe on the same populations agrees."""
    fx = GU.load("small")
    say("probe NaN")
    og = oracle_geometry(fx)
    r = fx["desc"]["run"]
    S = O.setup_sim(og, r["K"], r["aniso"], r["u"], r["tau"])
    g = gpu_geometry(fx)
    sim = abx.Sim(dev(), g.grid(), r["K"], r["aniso"], r["u"], r["tau"], precision="f32_f64")
    sim.steps(20)
    f = sim.read_f()
    fl = np.flatnonzero(og.fluid)
    js = (fl // og.nz) % og.ny
    a = int(np.flatnonzero(js == og.top + 1)[0])
    f.reshape(19, -1)[:, a] = np.nan
    sim.write_f(f)
    with pytest.raises(abx.AbxError) as e:
        sim.probe(37)
    assert str(e.value) == MSG.format(37)
    with pytest.raises(O.OracleError) as eo:
        O.probe(S, f, 37)
    assert str(eo.value) == str(e.value)
    # a finite state probes again without error
    f.reshape(19, -1)[:, a] = S.f0().reshape(19, -1)[:, a]
    sim.write_f(f)
    gam, mr, qo = sim.probe(38)
    assert np.isfinite(qo)
    sim.destroy()


def _node_case(tmp_path, name, **extra):
    fx = GU.load(name)
    d = fx["desc"]
    stl = str(tmp_path / "box.stl")
    G.write_stl_binary(stl, G.airbox_triangles(G.spec_from_dict(d["spec"])))
    case = {"stl": stl, "prepare": d["prepare"], "run": d["run"], "precision": "f32_f64", "selftest": False}
    case.update(extra)
    cj = str(tmp_path / "case.json")
    with open(cj, "w") as fh:
        json.dump(case, fh)
    print("[gpu] node case", name, sorted(extra), flush=True)
    p = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "run_hip_case.js"), cj], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return fx, json.loads(p.stdout)


def test_node_simulate_rejects_with_the_reference_message(tmp_path):
    fx = GU.load("small")
    want = _oracle_failure(oracle_geometry(fx), UNSTABLE)
    fx, out = _node_case(tmp_path, "small", run=UNSTABLE, nanCell=0)
    # ns = NaN on one cell is ignored, as in the reference (sv > 0 is false
    # for NaN, engine.js:313), so the run fails from the unstable parameters
    assert out["simError"] == want
    assert out["progressSteps"] == [10, 20, 30]


def test_node_gate_failure_throws_the_reference_error(tmp_path):
    """runHIP with a gate tighter than the deviation throws the reference's
    "Abweichung GPU/CPU" error with the measured deviation; the default gate
    passes the same geometry."""
    fx, out = _node_case(tmp_path, "small", gateTolerance=0)
    md = out["selftest"]
    assert 0 < md < 1e-3
    err = out["runError"]
    assert err.startswith("Abweichung GPU/CPU ")
    # md.toExponential(1), e.g. "2.4e-6"
    shown = err[len("Abweichung GPU/CPU "):]
    assert "e" in shown and len(shown.split("e")[0]) == 3
    assert float(shown) == pytest.approx(md, rel=0.05)
    fx, out = _node_case(tmp_path, "small", gateTolerance=1e-3)
    assert out["runError"] is None


def test_worker_answers_a_gate_failure_with_an_error_message(tmp_path):
    fx = GU.load("small")
    d = fx["desc"]
    stl = str(tmp_path / "box.stl")
    G.write_stl_binary(stl, G.airbox_triangles(G.spec_from_dict(d["spec"])))
    cj = str(tmp_path / "case.json")
    with open(cj, "w") as fh:
        json.dump({"stl": stl, "prepare": d["prepare"], "run": dict(d["run"], gateTolerance=0)}, fh)
    print("[gpu] node worker gate failure", flush=True)
    p = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "run_worker_case.js"), cj], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = json.loads(p.stdout)
    assert "result" not in out
    assert len(out["errors"]) == 1 and out["errors"][0]["during"] == "run"
    assert out["errors"][0]["message"].startswith("Abweichung GPU/CPU ")
    assert out["status"] == ["Grafikkarte wird geprüft …"]
