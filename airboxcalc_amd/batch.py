"""Batched design-variant solves (many airboxes of uneven grid size at once).

The reference runs one geometry per solve; its GpuSim batches B variants of
one geometry only (engine.js:495-545).  Here every variant has its own
geometry, cell size, K and aniso: each is prepared on the GPU, the simulation
handles join one abx.Batch (one launch per step for all of them, launch shape
chosen by the autotuner, chunks replayed from HIP graphs), the reference's
time loop runs over the batch, and every member is evaluated on the GPU.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import abx
from . import geometry as G

U0, TAU = 0.06, 0.51


@dataclass
class Variant:
    seed: int
    spec: G.AirboxSpec
    dx: float
    K: float
    aniso: float


def seeded_variants(seeds) -> List[Variant]:
    """The batched-variant family: outer dims uniform 200-360 mm per axis,
    dx in {2, 2.5, 3, 4} mm, K uniform 1.5-6, aniso uniform 5-20."""
    out = []
    for s in seeds:
        spec, ex = G.seeded_variant(int(s), "batch")
        out.append(Variant(int(s), spec, ex["dx"], ex["K"], ex["aniso"]))
    return out


def prepare_all(dev: abx.Device, variants: List[Variant]):
    geoms = []
    for v in variants:
        tris = G.airbox_triangles(v.spec)
        geoms.append(abx.Geometry(dev, tris, G.prepare_params(v.spec, v.dx), max_cells=0))
    return geoms


def solve_batch(dev: abx.Device, variants: List[Variant], geoms, steps: int, avg: int, every: int,
                precision: str = "f32", tune: bool = True, graph: bool = True, time_fn=None) -> dict:
    """simulate (engine.js:341-358) for every variant at once.  Returns the
    per-variant metrics, the probe history and the aggregate throughput."""
    sims = [abx.Sim(dev, g.grid(), K=v.K, aniso=v.aniso, u=U0, tau=TAU, precision=precision)
            for v, g in zip(variants, geoms)]
    b = abx.Batch(dev, sims)
    tuned = b.tune(20) if tune else None
    b.set_launch(tuned["spb"] if tuned else 4, tuned["persistent"] if tuned else False, graph)
    NF = sum(s.NF for s in sims)
    start_acc = steps - avg
    hist = []
    step = nacc = 0
    ms_kernel = {0: 0.0, 1: 0.0}
    n_kernel = {0: 0, 1: 0}
    b.sync()
    t0 = time.perf_counter()
    while step < steps:
        n = min(every, steps - step)
        if step < start_acc < step + n:
            n = start_acc - step
        doacc = step >= start_acc
        ms_kernel[int(doacc)] += b.time_steps(n, doacc)
        n_kernel[int(doacc)] += n
        step += n
        if doacc:
            nacc += n
        if step % every == 0 or step == steps:
            hist.append((step, b.probe(step)))
    b.sync()
    wall = time.perf_counter() - t0
    results = []
    for v, g, s in zip(variants, geoms, sims):
        ev = s.evaluate(g.eval_view(), divisor=max(nacc, 1), field3d=False)
        results.append(dict(seed=v.seed, dx=v.dx, K=v.K, aniso=v.aniso, grid=[g.nx, g.ny, g.nz], NF=s.NF,
                            metrics=ev["metrics"]))
    b.destroy()
    for s in sims:
        s.destroy()
    return dict(results=results, NF=NF, steps=steps, wall_s=wall, mlups=NF * steps / wall / 1e6,
                kernel_ms_per_step=ms_kernel[0] / max(n_kernel[0], 1),
                kernel_ms_per_step_acc=ms_kernel[1] / max(n_kernel[1], 1), tune=tuned, hist=hist)


def solve_one(dev: abx.Device, v: Variant, g, steps: int, avg: int, every: int, precision: str = "f32") -> dict:
    """The same variant solved alone (for the batch-vs-individual check)."""
    from .engine import simulate_hip
    p = dict(K=v.K, aniso=v.aniso, u=U0, tau=TAU, steps=steps, avg=avg, every=every)
    res = simulate_hip(dev, g, p, precision=precision, keep_sim=True)
    ev = res.sim.evaluate(g.eval_view(), divisor=max(res.nacc, 1), field3d=False)
    res.sim.destroy()
    return dict(metrics=ev["metrics"], hist=res.hist)
