"""Host orchestration of the gfx950 backend, mirroring engine.js's API.

prepare(dev, tris, p)            -> Geometry        (engine.js:110-234)
simulate_hip(dev, g, p, cb, ns)  -> SimResult       (simulateGPU, engine.js:613-630)

The time loop is the reference's: chunks of `every` steps split where the
averaging window starts, a probe (gamma, mass ratio, non-finite check) after
each full chunk and at the end, then the averaged acc divided by nacc.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import abx


def prepare(dev: abx.Device, tris: np.ndarray, p: dict, max_cells: float = 4.5e6, cy=None,
            sy=None) -> abx.Geometry:
    """prepare(p) (engine.js:110-234): voxelizer and vote on the GPU."""
    return abx.Geometry(dev, tris, p, max_cells=max_cells, cy=cy, sy=sy)


@dataclass
class SimResult:
    acc: np.ndarray                 # 4*NF, cid order, divided by nacc
    hist: List[list] = field(default_factory=list)
    nacc: int = 0
    NF: int = 0
    steps: int = 0
    wall_s: float = 0.0
    progress: List[dict] = field(default_factory=list)
    info: Optional[abx.SimInfo] = None


def simulate_hip(dev: abx.Device, g, p: dict, on_progress: Optional[Callable] = None, ns=None,
                 precision: str = "f32", sim: Optional[abx.Sim] = None, keep_sim: bool = False) -> SimResult:
    """simulateGPU (engine.js:613-630) on the gfx950 kernel.

    g is an abx.Geometry or an abx.GridArrays; p = {K, aniso, u, tau, steps,
    avg, every}.  Progress objects carry the reference's fields
    {step, steps, gamma, massRatio, elapsed, eta, mlups, device}."""
    grid = g.grid() if isinstance(g, abx.Geometry) else g
    own = sim is None
    if own:
        sim = abx.Sim(dev, grid, p["K"], p["aniso"], p["u"], p["tau"], precision=precision, B=1, ns=ns)
    steps, avg, every = int(p["steps"]), int(p["avg"]), int(p["every"])
    t0 = time.time()
    hist, prog = [], []
    step = nacc = 0
    start_acc = steps - avg
    while step < steps:
        n = min(every, steps - step)
        if step < start_acc < step + n:
            n = start_acc - step
        doacc = step >= start_acc
        sim.steps(n, doacc)
        step += n
        if doacc:
            nacc += n
        if step % every == 0 or step == steps:
            gam, mr, _ = sim.probe(step)
            el = max(time.time() - t0, 1e-9)
            hist.append([step, gam, mr])
            info = dict(step=step, steps=steps, gamma=gam, massRatio=mr, elapsed=el,
                        eta=el / step * (steps - step), mlups=sim.NF * step / el / 1e6, device="GPU")
            prog.append(info)
            if on_progress:
                on_progress(info)
    acc = sim.read_acc(max(nacc, 1))
    res = SimResult(acc=acc, hist=hist, nacc=nacc, NF=sim.NF, steps=steps, wall_s=time.time() - t0,
                    progress=prog, info=sim.info)
    if own and not keep_sim:
        sim.destroy()
    elif keep_sim:
        res.sim = sim
    return res
