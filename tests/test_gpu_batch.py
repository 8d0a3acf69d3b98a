"""Batched solves and graph replays.

A batch steps several independent simulations (different geometries, grid
sizes, K and aniso) with one launch per step.  Every member must come out
bit-identical to the same simulation stepped alone, for every launch shape
(segments per workgroup, persistent grid, HIP-graph replay) and across the AA
pattern's even/odd forms.  Graph replays are checked against eager steps after
every replay, with eager work (probes, reads) in between."""
import numpy as np
import pytest

from airboxcalc_amd import abx
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, say

pytestmark = pytest.mark.gpu

MEMBERS = [("small", 3.0, 10.0), ("small_yaw", 5.5, 18.0), ("small_noout", 1.5, 6.0), ("c1_seed3", 2.0, 12.0)]
_GEOM = {}


def _geom(name):
    if name not in _GEOM:
        _GEOM[name] = gpu_geometry(GU.load(name))
    return _GEOM[name]


def _sims(precision):
    out = []
    for name, K, aniso in MEMBERS:
        g = _geom(name)
        out.append(abx.Sim(dev(), g.grid(), K=K, aniso=aniso, u=0.06, tau=0.51, precision=precision))
    return out


def _advance(stepper, plan):
    for n, acc in plan:
        stepper(n, acc)


PLAN = [(3, False), (4, False), (5, True), (2, True)]


@pytest.mark.parametrize("shape", [(2, False, False, 4, 16), (4, False, False, 4, 0), (8, True, False, 4, 16),
                                   (4, False, True, 4, 16), (8, False, True, 4, 0), (2, True, True, 4, 8),
                                   (4, False, True, 1, 16), (2, False, False, 2, 0), (8, True, True, 8, 32),
                                   (4, False, False, 16, 16)])
def test_batch_members_match_solo_runs(shape):
    spb, pers, graph, group, occ = shape
    say("batch shape", shape)
    solo = _sims("f32")
    for s in solo:
        _advance(lambda n, acc: s.steps(n, acc), PLAN)
    members = _sims("f32")
    b = abx.Batch(dev(), members)
    b.set_xcd_group(group)
    b.set_occupancy(occ)
    b.set_launch(spb, pers, graph)
    assert b.xcd_group == group
    _advance(b.steps, PLAN)
    b.sync()
    for i, (s, m) in enumerate(zip(solo, members)):
        assert np.array_equal(m.read_f(), s.read_f()), i
        assert np.array_equal(m.read_acc(), s.read_acc()), i
    pr = b.probe(14)
    for i, s in enumerate(solo):
        assert np.array_equal(pr[i], np.array(s.probe(14)), equal_nan=True), i
    b.destroy()


def test_batch_f32_f64_members_stay_engine_exact():
    """In the reference's arithmetic the batched members still give engine.js's
    12-step populations bit for bit."""
    say("batch f32_f64 vs fixtures")
    names = ["small", "small_yaw", "c1_seed3"]
    members = []
    for name in names:
        fx = GU.load(name)
        r = fx["desc"]["run"]
        members.append(abx.Sim(dev(), _geom(name).grid(), r["K"], r["aniso"], r["u"], r["tau"], precision="f32_f64"))
    b = abx.Batch(dev(), members)
    b.set_launch(4, False, True)
    b.steps(1)
    b.sync()
    for name, m in zip(names, members):
        assert GU.sha(m.read_f()) == GU.load(name)["sha"]["f_plain_1"], name
    b.steps(1)
    b.steps(10)
    b.sync()
    for name, m in zip(names, members):
        assert GU.sha(m.read_f()) == GU.load(name)["sha"]["f_plain_12"], name
    b.destroy()


@pytest.mark.parametrize("kernel", [4, 2])
def test_single_sim_graph_replays_match_eager(kernel):
    say("graph replay kernel", kernel)
    g = _geom("small_yaw")
    eager = abx.Sim(dev(), g.grid(), precision="f32")
    graph = abx.Sim(dev(), g.grid(), precision="f32")
    for s in (eager, graph):
        s.set_launch(kernel=kernel)
    graph.set_graph(True)
    for r in range(5):
        acc = r >= 2
        eager.steps(7, acc)
        graph.steps(7, acc)  # same chunk: captured once, replayed afterwards
        assert np.array_equal(eager.probe(7 * (r + 1)), graph.probe(7 * (r + 1)), equal_nan=True)
        assert np.array_equal(graph.read_f(), eager.read_f()), r
        assert np.array_equal(graph.read_acc(), eager.read_acc()), r


def test_batch_tune_picks_a_shape_and_resets():
    say("batch tune")
    members = _sims("f32")
    b = abx.Batch(dev(), members)
    b.steps(6, True)
    res = b.tune(10)
    assert res["spb"] in (2, 4, 8) and res["ms_per_step"] > 0 and res["xcd_group"] in (1, 2, 4, 8, 16)
    assert res["even_waves_per_cu"] in (0, 16)
    with pytest.raises(abx.AbxError):
        b.set_xcd_group(3)
    for m in members:
        f = m.read_f()
        NF = m.NF
        w = np.repeat(np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float32), NF)
        assert np.array_equal(f, w)
        assert not m.read_acc().any()
    b.destroy()
