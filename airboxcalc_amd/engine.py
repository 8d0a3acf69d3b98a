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


# ---------------------------------------------------------------------------
# Output files in the layout of the reference's command-line tool, so that its
# `--from-npz` report reads results of this backend (airbox_cfd.py:160-166,
# 287-288, 360-366).

def cli_arrays(g, acc: np.ndarray) -> dict:
    """The NPZ arrays of airbox_cfd.py:287 from a prepared geometry g and the
    averaged fields acc (4*NF, components ux, uy, uz, rho, cid order): masks
    and fields as (nx, ny, nz) arrays, fields 0 outside the fluid."""
    from .surrogate import cell_types
    grid = g.grid()
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    fluid, inlet, outlet, porous = cell_types(grid)
    NF = int(fluid.sum())
    a = np.asarray(acc, np.float64).reshape(4, NF)
    out = dict(fluid=fluid, porous=porous, solid=g.array("solid").reshape(nx, ny, nz).astype(bool),
               xc=g.array("xc").astype(np.float64), dc=g.array("dc").astype(np.float64),
               zc=g.array("zc").astype(np.float64))
    for c, k in enumerate(("ux", "uy", "uz", "rho")):
        v = np.zeros((nx, ny, nz), np.float64)
        v[fluid] = a[c]
        out[k] = v
    return out


def save_cli_npz(path: str, g, acc: np.ndarray, u: float, pleat: float, hist) -> None:
    """airbox_result*.npz of the reference's CLI (airbox_cfd.py:287-288)."""
    np.savez_compressed(path, **cli_arrays(g, acc), u=u, pleat=pleat, hist=np.array(hist, dtype=np.float64))


def cli_report(metrics: dict, K: float, dx: float, steps: int, hist) -> dict:
    """The JSON result object of the reference's CLI (airbox_cfd.py:360-366),
    from evaluate's metrics (engine.js:682-685, same definitions)."""
    zeta = float(metrics["zeta"])
    return {"gamma": round(float(metrics["gamma"]), 4), "gamma_ohne_randzellen": round(float(metrics["gammaCore"]), 4),
            "anteil_unter_50pct_kern": round(float(metrics["lowCore"]), 4),
            "anteil_ueber_150pct_kern": round(float(metrics["highCore"]), 4),
            "anteil_rueckstroemung": round(float(metrics["reverse"]), 4),
            "max_zu_mittel": round(float(metrics["maxToMean"]), 3),
            "zeta_rohluftseite": round(zeta, 3),
            "dp_rohluft_mbar_bei_30ms": round(zeta * 0.5 * 1.2 * 30 ** 2 / 100, 2),
            "K": K, "dx": dx, "steps": steps, "verlauf": [list(map(float, r)) for r in hist]}
