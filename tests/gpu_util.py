"""Shared pieces of the GPU parity tests."""
from __future__ import annotations

import sys

import numpy as np

from airboxcalc_amd import abx
from oracle import oracle as O
from tests import golden_util as GU


def say(*a):
    """Name the case on stdout before a GPU step (a fault then names it)."""
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def dev():
    return abx.device(0)


def fixture_trig(fx):
    return GU.x64(fx["frame"]["cy"]), GU.x64(fx["frame"]["sy"])


def gpu_geometry(fx, max_cells=4.5e6):
    cy, sy = fixture_trig(fx)
    return abx.Geometry(dev(), GU.tris_of(fx), fx["desc"]["prepare"], max_cells=max_cells, cy=cy, sy=sy)


def oracle_geometry(fx, max_cells=4.5e6):
    cy, sy = fixture_trig(fx)
    return O.prepare(GU.tris_of(fx), fx["desc"]["prepare"], cy=cy, sy=sy, max_cells=max_cells)


def grid_of_oracle(g) -> abx.GridArrays:
    return abx.GridArrays(g.nx, g.ny, g.nz, g.fluid, g.opening, g.outlet, g.dc, g.top, g.jm,
                          float(g.s.pleat))


def selftest_ns(flag: np.ndarray) -> np.ndarray:
    """ns of gpuSelfTest's second variant: 1 on every 17th cell with flag 0
    (engine.js:599-600)."""
    ns = np.zeros(flag.size, np.float32)
    idx = np.arange(0, flag.size, 17)
    ns[idx[flag[idx] == 0]] = 1
    return ns
