"""The surrogate warm start and the steady criterion behind the C ABI and the
Node drop-in (no reference counterpart: the reference starts from rest,
engine.js:260-261, and runs a fixed number of steps, engine.js:613-630).

- abx_surrogate_warm_start (C++ over the library's kernels, no torch) seeds a
  handle bit for bit as the Python trainer's network does
  (surrogate.warm_start_init over unet.UNet), both when the box must be
  prepared again at the network's cell size and when the solve's own
  geometry is reused.
- The Node drop-in (simulateHIP with {warmStart, criterion}) seeds the same
  populations and reaches the same steady result, window by window, as the
  Python path over the same C ABI (steady.solve_to_steady).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from airboxcalc_amd import abx, steady
from airboxcalc_amd import geometry as G
from airboxcalc_amd import surrogate as S
from airboxcalc_amd import unet as U
from tests import golden_util as GU
from tests.gpu_util import dev, say

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# an unyawed box (cos and sin of the inlet yaw exact in every path), one baffle
SPEC = G.AirboxSpec(baffles=[G.Baffle("x", 120.0, 3.0, 60.0, 160.0, 70.0)])
K, ANISO = 3.5, 12.0
RUN = dict(K=K, aniso=ANISO, u=S.U0, tau=S.TAU, steps=2000, avg=500, every=100)


@pytest.mark.parametrize("dx,weights", [(2.0, S.WEIGHTS_DX2), (S.FAMILY_DX, S.WEIGHTS)])
def test_c_abi_warm_start_equals_the_python_network(dx, weights):
    say("warm start C ABI vs Python network, dx", dx)
    tris = G.airbox_triangles(SPEC)
    g = abx.Geometry(dev(), tris, G.prepare_params(SPEC, dx), max_cells=0)
    net = U.UNet(seed=0)
    norm, _ = S.load(weights, net)
    box = S.Box(0, SPEC, K, ANISO)
    a = abx.Sim(dev(), g.grid(), K, ANISO, S.U0, S.TAU)
    S.warm_start_init(net, norm, dev(), box, g, a, tris)
    sur = abx.Surrogate(dev(), weights)
    assert sur.dims == S.GRID and sur.family_dx == S.FAMILY_DX
    b = abx.Sim(dev(), g.grid(), K, ANISO, S.U0, S.TAU)
    sur.warm_start(b, g, K, ANISO)
    fa, fb = a.read_f(), b.read_f()
    assert np.array_equal(fa, fb)
    # the seed is not the cold start: the prediction carries a flow
    assert not np.array_equal(fb, np.repeat(np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float32), a.NF))
    for s in (a, b):
        s.destroy()
    sur.free()


def test_node_warm_start_and_criterion_equal_the_python_path(tmp_path):
    tris = G.airbox_triangles(SPEC)
    stl = str(tmp_path / "box.stl")
    G.write_stl_binary(stl, tris)
    crit = dict(window=500, maxSteps=4000, tolGamma=0.005, tolZeta=0.02)
    case = {"stl": stl, "prepare": G.prepare_params(SPEC, 2.0), "run": RUN, "precision": "f32",
            "warmStart": {"weights": S.WEIGHTS_DX2, "criterion": crit}}
    cj = str(tmp_path / "case.json")
    with open(cj, "w") as fh:
        json.dump(case, fh)
    print("[gpu] node warm start + criterion", flush=True)
    p = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "run_hip_case.js"), cj], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = json.loads(p.stdout)
    say("python warm start + criterion")
    g = abx.Geometry(dev(), tris, G.prepare_params(SPEC, 2.0), max_cells=0)
    sur = abx.Surrogate(dev(), S.WEIGHTS_DX2)
    s = abx.Sim(dev(), g.grid(), K, ANISO, S.U0, S.TAU)
    sur.warm_start(s, g, K, ANISO)
    assert out["sha"]["warm_f"] == GU.sha(s.read_f())
    s.destroy()
    res = steady.solve_to_steady(dev(), g, RUN, window=crit["window"], tol_gamma=crit["tolGamma"],
                                 tol_zeta=crit["tolZeta"], max_steps=crit["maxSteps"],
                                 init_fn=lambda sim: sur.warm_start(sim, g, K, ANISO))
    st = out["steady"]
    assert st["converged"] == res.converged and st["steps"] == res.steps
    assert [(w[0], w[3], w[4]) for w in st["windows"]] == res.windows
    for k, v in res.metrics.items():
        assert out["metrics"][k] == v, k
    assert [row[0] for row in out["hist"]] == [w[0] for w in res.windows]
    assert [pr["step"] for pr in out["progress"]] == [w[0] for w in res.windows]
    assert all("gammaCore" in pr and "zeta" in pr for pr in out["progress"])
    sur.free()
