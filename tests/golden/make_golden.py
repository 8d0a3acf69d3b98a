"""Write the golden fixtures of the reference compute core (src/engine.js).

Run in a container that holds the reference and Node:

    python tests/golden/make_golden.py [/root/reference/src/engine.js]

For each case below this builds a procedural airbox STL with
``airboxcalc_amd.geometry``, runs ``engine_golden.js`` on it (which loads
engine.js from the given path, never from this repository) and stores

* ``<case>.json``: the case description, scalars, exact float64 values (hex)
  and the SHA-256 of every typed array the reference produced;
* ``<case>.npz``: for the small cases, the arrays themselves (byte masks as
  packed bits) so a mismatch can be located, not only detected.

The fixtures pin what no closed form does: voxelizer bit patterns, the
opening choice, setupSim scalars, 12-step cpuStep trajectories (f32 storage,
f64 arithmetic), probe/progress reductions, evaluate/field3d/streamlines on
real fields, and the reference's error messages.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from airboxcalc_amd import geometry as G  # noqa: E402

SMALL = dict(LX=120.0, LY=60.0, LZ=90.0, wall=6.0, inlet_d=30.0, outlet_d=24.0, outlet_h=20.0, pleat=16.0)


def _run(steps, avg, every, **kw):
    return G.run_params(steps, avg, every, **kw)


def cases():
    out = []
    s = G.AirboxSpec(**SMALL)
    out.append(dict(name="small", spec=s, dx=4.0, outlet=True, click=True, voxel=True, keep_arrays=True,
                    run=_run(240, 80, 40), simulate=True, simulate_ns=True))
    out.append(dict(name="small_noout", spec=s, dx=4.0, outlet=False, click=False, voxel=True, keep_arrays=True,
                    run=_run(120, 40, 40), simulate=True))
    sy = G.AirboxSpec(**dict(SMALL, yaw_deg=20.0))
    sy.baffles = [G.Baffle("x", 70.0, 3.0, 45.0, 75.0, 32.0)]
    out.append(dict(name="small_yaw", spec=sy, dx=3.0, outlet=True, click=True, voxel=True, keep_arrays=True,
                    run=_run(120, 40, 40), simulate=True))
    out.append(dict(name="c1", spec=G.AirboxSpec(), dx=4.0, outlet=True, click=True, voxel=True,
                    run=_run(2000, 500, 100), simulate=True))
    s3, _ = G.seeded_variant(3, "roofline")
    out.append(dict(name="c1_seed3", spec=s3, dx=4.0, outlet=True, click=True, voxel=True,
                    run=_run(400, 100, 100), simulate=True))
    s5, _ = G.seeded_variant(5, "roofline")
    out.append(dict(name="seed5_dx3", spec=s5, dx=3.0, outlet=True, click=False, voxel=True,
                    run=_run(60, 20, 20), simulate=True))
    # error paths of prepare (engine.js:124, 156, 195, 200, 210)
    out.append(dict(name="err_grid", spec=G.AirboxSpec(), dx=1.0, outlet=True, click=True))
    out.append(dict(name="err_noopen", spec=G.AirboxSpec(**dict(SMALL, inlet_d=4.0)), dx=4.0, outlet=True,
                    click=True, voxel=True))
    out.append(dict(name="err_leak", spec=G.AirboxSpec(**dict(SMALL, wall=0.6)), dx=4.0, outlet=True, click=True,
                    voxel=True))
    out.append(dict(name="err_lid", spec=G.AirboxSpec(**dict(SMALL, outlet_h=1.0)), dx=4.0, outlet=True,
                    click=True))
    so = G.AirboxSpec(**SMALL)
    out.append(dict(name="err_outlet", spec=so, dx=4.0, outlet=True, click=True,
                    outlet_center=[500.0, 0.0, 45.0]))
    out.append(dict(name="ascii", spec=s, dx=4.0, outlet=True, click=True, ascii=True, voxel=True))
    return out


# arrays kept in the npz of keep_arrays cases (the rest is kept as SHA-256 only)
KEEP_U8 = ("fine", "solid", "fluid", "opening", "outlet")
KEEP_RAW = ("xc", "dc", "zc", "flag", "layerM", "layerT", "vy_plain", "vy_ns", "ns_selftest",
            "sim_fmap", "simns_fmap", "sim_lines_pts", "sim_lines_spd", "sim_lines_off",
            "sim_planInlet_speed", "sim_planLow_speed", "sim_lon_speed",
            "sim_vol_spd", "sim_vol_pst", "sim_vol_loss", "openPts", "filterPts", "outletPts")
DTYPES = {"fine": np.uint8, "solid": np.uint8, "fluid": np.uint8, "opening": np.uint8, "outlet": np.uint8,
          "xc": np.float32, "dc": np.float32, "zc": np.float32, "flag": np.uint32, "layerM": np.int32,
          "layerT": np.int32, "vy_plain": np.float32, "vy_ns": np.float32, "ns_selftest": np.float32,
          "sim_lines_off": np.uint32, "openPts": np.float64, "filterPts": np.float64, "outletPts": np.float64}


def make(engine: str, only=None) -> None:
    for c in cases():
        if only and c["name"] not in only:
            continue
        spec = c["spec"]
        tris = G.airbox_triangles(spec)
        prep = G.prepare_params(spec, c["dx"], outlet=c["outlet"], click=c["click"])
        if "outlet_center" in c:
            prep["outlet"]["center"] = c["outlet_center"]
        desc = {"name": c["name"], "spec": G.spec_dict(spec), "dx": c["dx"], "prepare": prep,
                "run": c.get("run"), "simulate": c.get("simulate", False),
                "simulate_ns": c.get("simulate_ns", False), "voxel": c.get("voxel", False),
                "keep_arrays": c.get("keep_arrays", False), "ascii": c.get("ascii", False)}
        with tempfile.TemporaryDirectory() as td:
            stl = os.path.join(td, "box.stl")
            (G.write_stl_ascii if c.get("ascii") else G.write_stl_binary)(stl, tris)
            desc["stl"] = stl
            cj = os.path.join(td, "case.json")
            with open(cj, "w") as fh:
                json.dump(desc, fh)
            od = os.path.join(td, "out")
            subprocess.run(["node", os.path.join(HERE, "engine_golden.js"), engine, cj, od], check=True)
            with open(os.path.join(od, "result.json")) as fh:
                res = json.load(fh)
            desc.pop("stl")
            res["desc"] = desc
            with open(os.path.join(HERE, c["name"] + ".json"), "w") as fh:
                json.dump(res, fh, indent=1, sort_keys=True)
            if c.get("keep_arrays"):
                arrs = {}
                for k in KEEP_U8 + KEEP_RAW:
                    p = os.path.join(od, k + ".bin")
                    if not os.path.exists(p):
                        continue
                    a = np.fromfile(p, dtype=DTYPES.get(k, np.float32))
                    if k in KEEP_U8:
                        arrs[k + "_bits"] = np.packbits(a)
                        arrs[k + "_n"] = np.array(a.size)
                    else:
                        arrs[k] = a
                np.savez_compressed(os.path.join(HERE, c["name"] + ".npz"), **arrs)
        print(c["name"], "done", res.get("errors"), flush=True)


if __name__ == "__main__":
    eng = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/engine.js"
    make(eng, set(sys.argv[2:]) or None)
