"""GPU voxelizer and geometry preparation against the reference's outputs.

The fixtures hold the SHA-256 of engine.js's own arrays (tests/golden); the
GPU path must reproduce them bit for bit."""
import numpy as np
import pytest

from airboxcalc_amd import abx
from oracle import oracle as O
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, oracle_geometry, say

pytestmark = pytest.mark.gpu

PREP = GU.names(lambda r: "prepare" in r)
ERRS = GU.names(lambda r: "errors" in r)


def test_device_is_gfx950():
    d = dev()
    say("device", d.name)
    assert "gfx950" in d.name


@pytest.mark.parametrize("name", PREP)
def test_prepare_bit_exact(name):
    fx = GU.load(name)
    say("prepare", name)
    g = gpu_geometry(fx)
    P = fx["prepare"]
    for k in ("nx", "ny", "nz", "N", "nf", "top", "jm"):
        assert getattr(g, k) == P[k], k
    assert list(g.mn) == [GU.x64(e) for e in P["mn"]]
    assert list(g.mx) == [GU.x64(e) for e in P["mx"]]
    for k in ("solid", "fluid", "opening", "outlet", "xc", "dc", "zc", "openPts", "filterPts", "outletPts"):
        assert GU.sha(g.array(k)) == fx["sha"][k], k
    if "fine" in fx["sha"]:
        assert GU.sha(g.array("fine")) == fx["sha"]["fine"]
    S = P["summary"]
    assert [g.nx, g.ny, g.nz] == S["grid"] and g.nf == S["cells"] and g.nHoles == S["nHoles"]
    assert bool(g.pleatOk) == S["pleatOk"]
    for k in ("inlet_cm2", "filter_cm2", "outlet_cm2", "outlet_cm2_real"):
        assert getattr(g, k) == GU.x64(S[k]), k
    assert g.volume_l == pytest.approx(GU.x64(S["volume_l"]), rel=1e-15)


@pytest.mark.parametrize("name", ERRS)
def test_prepare_errors(name):
    fx = GU.load(name)
    say("prepare error", name)
    with pytest.raises(abx.AbxError) as ei:
        gpu_geometry(fx)
    assert str(ei.value) == fx["errors"]["prepare"]


def _random_mesh(rng, n_boxes):
    """A few random axis-aligned and rotated closed boxes (12 triangles each)."""
    tris = []
    for _ in range(n_boxes):
        c = rng.uniform(2, 8, 3)
        e = rng.uniform(0.3, 2.5, 3)
        th = rng.uniform(0, np.pi)
        R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
        v = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], float) * e
        v = v @ R.T + c
        faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
        for a, b, cc, d in faces:
            tris.append(np.concatenate([v[a], v[b], v[cc]]))
            tris.append(np.concatenate([v[a], v[cc], v[d]]))
    return np.array(tris, np.float32).ravel()


@pytest.mark.parametrize("seed", range(6))
def test_voxelize_random_meshes_bit_exact(seed):
    rng = np.random.default_rng(seed)
    T = _random_mesh(rng, 1 + seed)
    origin = rng.uniform(-0.3, 0.3, 3)
    h = float(rng.choice([0.1, 0.13333333333333333, 0.25]))
    nx, ny, nz = int(10.5 / h), int(10.5 / h), int(10.5 / h)
    say("voxelize seed", seed, nx, ny, nz)
    a = abx.voxelize(dev(), T, origin, nx, ny, nz, h)
    b = O.voxelize(T, origin, nx, ny, nz, h)
    assert a.sum() > 0
    assert np.array_equal(a, b)


def test_voxelize_empty_and_degenerate():
    T = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2], np.float32)  # zero projected area: skipped
    say("voxelize degenerate")
    a = abx.voxelize(dev(), T, [0, 0, 0], 8, 8, 8, 0.5)
    assert a.sum() == 0
    a = abx.voxelize(dev(), np.zeros(0, np.float32), [0, 0, 0], 4, 4, 4, 1.0)
    assert a.sum() == 0


@pytest.mark.parametrize("seed", [1000000, 1000003])
def test_study_box_200k_triangles_bit_exact(seed):
    """The warm-start study's input: a box written as a binary STL of about
    200 k triangles, parsed and prepared on the GPU at dx = 2 mm.  Every array
    must equal the oracle's (itself bit-exact to engine.js on the fixtures)."""
    from airboxcalc_amd import geometry as G
    from airboxcalc_amd import surrogate as S
    box = S.family([seed])[0]
    box.spec.subdiv = S.STUDY_SUBDIV
    stl = G.stl_binary_bytes(G.airbox_triangles(box.spec))
    tris = abx.parse_stl(stl)
    assert tris.size // 9 > 150_000
    assert np.array_equal(tris.view(np.uint32), O.parse_stl(stl).view(np.uint32))
    p = G.prepare_params(box.spec, 2.0)
    say("prepare 200k-triangle study box", seed, tris.size // 9)
    g = abx.Geometry(dev(), tris, p, max_cells=0)
    o = O.prepare(tris, p, max_cells=0)
    try:
        for k in ("nx", "ny", "nz", "N", "nf", "top", "jm"):
            assert getattr(g, k) == getattr(o, k), k
        for k in ("solid", "fluid", "opening", "outlet", "fine", "xc", "dc", "zc"):
            assert np.array_equal(g.array(k), getattr(o, k)), k
    finally:
        g.free()


@pytest.mark.parametrize("seed,dx", [(1, 1.0), (2, 1.5)])
def test_flood_fill_on_large_grids_matches_oracle(seed, dx):
    """Duct, floor, flood fill and leak check run on the GPU as a
    level-synchronous breadth-first search; on the roofline boxes at dx = 1 and
    1.5 mm (up to 17.5 M cells, hundreds of levels) every array equals the
    oracle's sequential flood."""
    from airboxcalc_amd import geometry as G
    spec, _ = G.seeded_variant(seed, "roofline")
    tris = G.airbox_triangles(spec)
    p = G.prepare_params(spec, dx, outlet=True, click=True)
    say("flood fill on a large grid", seed, dx)
    g = abx.Geometry(dev(), tris, p, max_cells=0)
    o = O.prepare(tris, p, max_cells=0)
    try:
        for k in ("nx", "ny", "nz", "N", "nf", "top", "jm"):
            assert getattr(g, k) == getattr(o, k), k
        for k in ("solid", "fluid", "opening", "outlet"):
            assert np.array_equal(g.array(k), getattr(o, k)), k
    finally:
        g.free()
