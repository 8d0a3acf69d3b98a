"""Loading the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import hashlib
import json
import os
import struct

import numpy as np

from airboxcalc_amd import geometry as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str) -> dict:
    with open(os.path.join(GOLD, name + ".json")) as fh:
        return json.load(fh)


def arrays(name: str) -> dict:
    z = np.load(os.path.join(GOLD, name + ".npz"))
    out = {}
    for k in z.files:
        if k.endswith("_bits"):
            base = k[:-5]
            out[base] = np.unpackbits(z[k])[: int(z[base + "_n"])]
        elif not k.endswith("_n"):
            out[k] = z[k]
    return out


def names(pred=None):
    fs = sorted(f[:-5] for f in os.listdir(GOLD) if f.endswith(".json"))
    return [n for n in fs if pred is None or pred(load(n))]


def x64(e) -> float:
    """Exact float64 from a fixture's {"v", "hex"} pair."""
    return struct.unpack("<d", bytes.fromhex(e["hex"]))[0]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tris_of(fx: dict) -> np.ndarray:
    """The case's triangle soup as engine.js parsed it (binary STL = the
    float32 triangles; ASCII STL = the printed values read back)."""
    d = fx["desc"]
    spec = G.spec_from_dict(d["spec"])
    t = G.airbox_triangles(spec).ravel()
    if d.get("ascii"):
        import tempfile
        from oracle import oracle as O
        with tempfile.NamedTemporaryFile(suffix=".stl") as fh:
            G.write_stl_ascii(fh.name, t)
            t = O.parse_stl(open(fh.name, "rb").read())
    return t
