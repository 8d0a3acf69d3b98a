"""Host C/C++ under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

tests/sanitize/Makefile builds the library's host sources (abx_core.cpp: error
state, device query, STL reader; prepare_host.cpp: the host stages of
abx_prepare) and the oracle's C restatement with -fsanitize=address,undefined
into one checking program, tests/sanitize/host_checks.cpp.  Any sanitizer
report aborts it (-fno-sanitize-recover=all), so every case here fails on a
memory error, a leak or undefined behaviour as well as on a mismatch.

- prepare: the product's host stages, with CPU stand-ins for the GPU stages
  (vote, slice, duct, floor, flood fill), equal the oracle's prepare array for
  array on every golden case of the reference (including its error cases:
  engine.js:26, 124, 155, 170, 178, 195, 199, 209) and on yawed, baffled boxes;
- oracle: setup, steps with averaging, probe, evaluate, field3d, streamlines;
- fuzz: random bytes, lying binary counts and token soup into abx_parse_stl
  (against the oracle's reader), degenerate grids into prep_grid, random
  slices into prep_openings;
- device: abx_device_open without a GPU fails with a message.
"""
import os
import subprocess

import pytest

from airboxcalc_amd import geometry as G
from oracle import oracle as O
from tests import golden_util as GU

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
EXE = os.path.join(SAN, "_build", "host_checks")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def exe():
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return EXE


def _run(exe, *args, timeout=300):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    return r.stdout


def _write_case(tmp_path, name, tris, p, ascii=False, max_cells=4.5e6):
    stl = str(tmp_path / f"{name}.stl")
    (G.write_stl_ascii if ascii else G.write_stl_binary)(stl, tris)
    P = O.prep_params(p, max_cells=max_cells)
    vals = [P.dx, P.sealAxis, P.sealSign, P.seal, P.inletAxis, P.inletSign, P.yaw_deg, P.cy, P.sy, P.has_click,
            *P.click, P.pleat, P.has_outlet, *P.out_center, P.out_diameter, P.out_height, P.max_cells]
    prm = str(tmp_path / f"{name}.prm")
    with open(prm, "w") as fh:
        fh.write(" ".join(repr(float(v)) for v in vals) + "\n")
    return stl, prm


GOLD_PREP = GU.names(lambda fx: fx["desc"].get("prepare") is not None)


@pytest.mark.parametrize("name", GOLD_PREP)
def test_prepare_host_stages_equal_the_oracle(exe, tmp_path, name):
    fx = GU.load(name)
    d = fx["desc"]
    tris = G.airbox_triangles(G.spec_from_dict(d["spec"])).ravel()
    stl, prm = _write_case(tmp_path, name, tris, d["prepare"], ascii=bool(d.get("ascii")))
    out = _run(exe, "prepare", stl, prm)
    if fx.get("errors"):
        assert out.startswith("error: "), out
        assert out.strip()[len("error: "):] == fx["errors"]["prepare"]
    else:
        assert out.startswith("ok grid"), out


@pytest.mark.parametrize("seed,dx,outlet", [(1, 4.0, True), (2, 3.0, False), (5, 3.0, True), (7, 5.0, True)])
def test_prepare_host_stages_on_yawed_baffled_boxes(exe, tmp_path, seed, dx, outlet):
    spec, _ = G.seeded_variant(seed, "roofline")
    p = G.prepare_params(spec, dx, outlet=outlet)
    stl, prm = _write_case(tmp_path, f"box{seed}", G.airbox_triangles(spec), p, max_cells=0)
    out = _run(exe, "prepare", stl, prm)
    assert out.startswith("ok grid") or out.startswith("error: "), out


def test_oracle_pipeline_under_sanitizers(exe, tmp_path):
    fx = GU.load("small")
    d = fx["desc"]
    tris = G.airbox_triangles(G.spec_from_dict(d["spec"])).ravel()
    stl, prm = _write_case(tmp_path, "small", tris, d["prepare"])
    out = _run(exe, "oracle", stl, prm, 60)
    assert out.startswith("oracle NF"), out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fuzz_host_inputs(exe, seed):
    out = _run(exe, "fuzz", seed, 20000)
    assert "20000 cases" in out


def test_device_open_without_gpu_is_an_error(exe):
    out = _run(exe, "device")
    assert out.startswith("device_open -> -1"), out
