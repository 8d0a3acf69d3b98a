"""The N-API addon and the JavaScript shim load under Node without a GPU and
fail loudly (no silent fallback) when asked for a device."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")


def test_addon_and_shim_load_and_fail_loudly_without_gpu():
    from airboxcalc_amd import abx
    try:
        abx.Device(0).close()
        pytest.skip("a GPU is present")
    except abx.AbxError:
        pass
    js = ("const H=require('./airboxcalc_amd/js/engine_hip.js');"
          "const k=Object.keys(H.native).sort().join(',');"
          "let msg='';try{H.native.deviceOpen(0)}catch(e){msg=e.message}"
          "console.log(JSON.stringify({keys:k,dev:H.hipDevice(),msg}))")
    p = subprocess.run(["node", "-e", js], cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    import json
    out = json.loads(p.stdout)
    for fn in ("prepare", "simCreate", "simChunk", "evaluate", "streamlines", "voxelize", "parseSTL"):
        assert fn in out["keys"].split(",")
    assert out["dev"] is None
    assert "no HIP GPU" in out["msg"]


def test_make_frame_round_trip():
    js = ("const H=require('./airboxcalc_amd/js/engine_hip.js');"
          "const f=H.makeFrame({sealAxis:1,sealSign:-1,seal:0,inletAxis:2,inletSign:-1,inletYaw:17});"
          "const q=[12.5,-3.25,40];const L=f.toLocal(q);const m=f.toModel(L[0],L[1],L[2]);"
          "console.log(JSON.stringify(m.map((v,i)=>Math.abs(v-q[i]))))")
    p = subprocess.run(["node", "-e", js], cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    import json
    assert max(json.loads(p.stdout)) < 1e-12


def test_worker_protocol_without_gpu(tmp_path):
    """The HIP worker speaks the reference's protocol (engine.js:760-801): it
    answers gpuinfo (no device), loads an STL (transferring the buffer), and
    answers prepare with an error message rather than a CPU fallback."""
    from airboxcalc_amd import abx
    from airboxcalc_amd import geometry as G
    try:
        abx.Device(0).close()
        pytest.skip("a GPU is present")
    except abx.AbxError:
        pass
    import json
    spec, _ = G.seeded_variant(3, "roofline")
    tris = G.airbox_triangles(spec)
    stl = str(tmp_path / "box.stl")
    G.write_stl_binary(stl, tris)
    cj = tmp_path / "case.json"
    cj.write_text(json.dumps({"stl": stl, "prepare": G.prepare_params(spec, 4.0), "run": None}))
    p = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "run_worker_case.js"), str(cj)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    out = json.loads(p.stdout)
    assert out["gpuinfo"] == {"ok": False, "name": ""}
    assert out["loaded"]["type"] == "loaded" and out["loaded"]["ntri"] == len(tris)
    assert out["loaded"]["transferred"]
    assert out["prepared"] is None and out["errors"][0]["during"] == "prepare"
