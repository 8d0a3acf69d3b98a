"""The procedural airbox generator (airboxcalc_amd/geometry.py) is closed and
has the volume it should.

- Watertight: every ray along Z crosses the shell an even number of times,
  counted by the oracle's restatement of the reference voxelizer's count pass
  (engine.js:50-75).  An odd count would make the even-odd fill
  (engine.js:82-91) drop a hit and fill the wrong run.  Checked on seeded
  boxes with baffles and an oblique inlet, on the uneven batch family and on
  a 200 k-triangle box of the warm-start study.
- Volume: for a box without yaw or baffles the solid is the outer box minus
  the cavity minus the inlet hole (a polygon with its vertices on the hole's
  circle, through the inlet wall).  The number of voxel centres the fill
  marks solid must equal the count of centres inside that set, computed here
  from the box dimensions alone.
"""
import math

import numpy as np
import pytest

from airboxcalc_amd import batch as BT
from airboxcalc_amd import geometry as G
from oracle import oracle as O

JX, JY = 1.37e-4, 2.71e-4  # the voxelizer's ray jitter (engine.js:47)


def _grid(tris, h, pad=2.0):
    T = tris.reshape(-1, 3).astype(np.float64)
    lo, hi = T.min(0) - pad, T.max(0) + pad
    n = np.ceil((hi - lo) / h).astype(int)
    return lo, n


@pytest.mark.parametrize("kind,seed", [("roofline", s) for s in range(6)] + [("batch", s) for s in range(6)])
def test_every_column_has_an_even_hit_count(kind, seed):
    spec, _ = G.seeded_variant(seed, kind)
    tris = G.airbox_triangles(spec)
    for h in (0.7, 1.3):
        lo, n = _grid(tris, h)
        cnt = O.voxel_counts(tris, lo, int(n[0]), int(n[1]), h)
        assert cnt.sum() > 0
        assert (cnt % 2 == 0).all(), (kind, seed, h, int((cnt % 2).sum()))


def test_200k_triangle_study_box_is_watertight():
    spec, _ = G.seeded_variant(1_000_003, "surrogate")
    spec.subdiv = 57
    tris = G.airbox_triangles(spec)
    assert 150_000 < tris.shape[0] < 260_000
    lo, n = _grid(tris, 0.9)
    cnt = O.voxel_counts(tris, lo, int(n[0]), int(n[1]), 0.9)
    assert (cnt % 2 == 0).all()


def _hole_polygon(spec):
    """Vertices of the inlet hole in model (x, y) for a box without yaw: the
    points of the hole's circle at the angles 2 pi q / n and at the directions
    of the inlet wall's rectangle corners seen from the hole centre."""
    t, LX, LY = spec.wall, spec.LX, spec.LY
    D1 = LY - t
    cx, dh, r = 0.5 * LX, 0.5 * LY, 0.5 * spec.inlet_d
    n = max(8, int(spec.segments))
    ang = {2 * math.pi * q / n for q in range(n)}
    for x in (t, LX - t):
        for y in (0.0, D1):
            ang.add(math.atan2(y - dh, x - cx) % (2 * math.pi))
    ang = sorted(ang)
    return np.array([(cx + r * math.cos(a), dh + r * math.sin(a)) for a in ang])


def _inside_polygon(px, py, poly):
    """Even-odd point-in-polygon for arrays of points."""
    inside = np.zeros(px.shape, bool)
    for (x0, y0), (x1, y1) in zip(poly, np.roll(poly, -1, 0)):
        crosses = (y0 > py) != (y1 > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xs = x0 + (py - y0) * (x1 - x0) / (y1 - y0)
        inside ^= crosses & (px < xs)
    return inside


@pytest.mark.parametrize("dims,h", [((300.0, 150.0, 200.0), 1.3), ((220.0, 300.0, 250.0), 1.7),
                                    ((340.0, 210.0, 180.0), 2.3)])
def test_plain_box_volume_in_closed_form(dims, h):
    LX, LY, LZ = dims
    spec = G.AirboxSpec(LX=LX, LY=LY, LZ=LZ, inlet_d=min(70.0, 0.4 * min(LX, LY)))
    tris = G.airbox_triangles(spec)
    t, D1 = spec.wall, LY - spec.wall
    origin = np.array([-3.17, -2.91, -4.33])
    nx, ny, nz = (int(math.ceil((L + 8.0) / h)) for L in dims)
    vox = O.voxelize(tris, origin, nx, ny, nz, h).reshape(nx, ny, nz)
    # ray positions and voxel centres (engine.js:54, 62-64, 88)
    px = origin[0] + JX + (np.arange(nx) + 0.5) * h
    py = origin[1] + JY + (np.arange(ny) + 0.5) * h
    pz = origin[2] + (np.arange(nz) + 0.5) * h
    X, Y = np.meshgrid(px, py, indexing="ij")

    def nz_in(a, b):
        return int(((pz >= a) & (pz < b)).sum())

    outer = (X > 0) & (X < LX) & (Y > 0) & (Y < LY)
    cavity = outer & (X > t) & (X < LX - t) & (Y < D1)
    hole = cavity & _inside_polygon(X, Y, _hole_polygon(spec))
    per_col = np.where(outer, nz_in(0.0, LZ), 0)
    per_col = np.where(cavity, nz_in(0.0, t) + nz_in(LZ - t, LZ), per_col)
    per_col = np.where(hole, nz_in(LZ - t, LZ), per_col)
    assert hole.sum() > 100
    assert int(vox.sum()) == int(per_col.sum())
    # and column by column
    assert np.array_equal(vox.sum(2), per_col)


@pytest.mark.parametrize("kind,seeds,subdiv,ascii_", [("batch", [0, 1, 7], 1, False), ("roofline", [2, 5], 2, False),
                                                      ("surrogate", [1_000_003], 1, True)])
def test_variant_set_manifest_regenerates_every_file(tmp_path, kind, seeds, subdiv, ascii_):
    """A written variant set (STL files + manifest.json, the generator's
    record of seeds and parameters): every file's hash and triangle count
    match the manifest, the recorded spec regenerates the binary files bit
    for bit, the recorded run parameters are the seeded draws, a file read
    back through the oracle's parseSTL keeps every column's hit count even,
    and a changed byte is reported."""
    import json
    out = str(tmp_path / kind)
    m = G.write_variant_set(out, kind, seeds, subdiv=subdiv, ascii=ascii_)
    assert m["count"] == len(seeds) and [v["seed"] for v in m["variants"]] == seeds
    assert G.verify_variant_set(out) == []
    with open(tmp_path / kind / "manifest.json") as fh:
        assert json.load(fh) == m
    for v in m["variants"]:
        spec, extra = G.seeded_variant(v["seed"], kind)
        assert v["run"]["K"] == extra.get("K", 3.0) and v["run"]["aniso"] == extra.get("aniso", 10.0)
        assert v["run"]["dx"] == (extra["dx"] if "dx" in extra else G.FAMILY_DX[kind])
        assert v["prepare"]["dx"] == v["run"]["dx"]
    v0 = m["variants"][0]
    with open(tmp_path / kind / v0["file"], "rb") as fh:
        data = fh.read()
    tris = O.parse_stl(data).reshape(-1, 9)
    assert len(tris) == v0["triangles"]
    lo, n = _grid(tris, 1.3)
    cnt = O.voxel_counts(tris, lo, int(n[0]), int(n[1]), 1.3)
    assert cnt.sum() > 0 and (cnt % 2 == 0).all()
    # a changed byte in a file no longer matches its recorded hash
    bad = bytearray(data)
    bad[-3] ^= 0x01
    with open(tmp_path / kind / v0["file"], "wb") as fh:
        fh.write(bytes(bad))
    problems = G.verify_variant_set(out)
    assert problems and v0["file"] in problems[0]


def test_variant_set_command_line(tmp_path):
    out = str(tmp_path / "cli")
    assert G.main(["--kind", "batch", "--seeds", "3-4,9", "--out", out]) == 0
    assert G.main(["--verify", "--out", out]) == 0
    import json
    with open(tmp_path / "cli" / "manifest.json") as fh:
        assert [v["seed"] for v in json.load(fh)["variants"]] == [3, 4, 9]
