"""Time to steady and the FP64 error report.

The reference has no convergence criterion: it always runs a fixed number of
steps and averages the last `avg` of them (engine.js:341-358, 613-630).  This
module defines the criterion the warm-start study measures against:

  run in windows of `window` steps, averaging each window on the GPU, and
  evaluate it (gamma-core and zeta, engine.js:632-691); the solve is steady
  when gamma-core moved by less than tol_gamma and zeta by less than
  tol_zeta (relative) between two consecutive windows.  The steady result is
  the mean of those two windows' metrics, a 2-window estimate like the
  reference's averaging over its last `avg` steps; zeta is linear in the
  averaged density, so its mean is exactly zeta of the 2-window average.

A cold start begins from f0 = WT (engine.js:260-261); a warm start from the
equilibrium of a predicted (rho, u) field (abx_sim_init_eq).  The FP64 report
runs the same solve in double storage and arithmetic and gives the error of
every result of the float run against it.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import abx
from .engine import simulate_hip


@dataclass
class SteadyResult:
    converged: bool
    steps: int
    wall_s: float
    metrics: dict                                        # steady: mean of the last two windows
    windows: List[tuple] = field(default_factory=list)   # (steps, gammaCore, zeta) of every window
    acc: Optional[np.ndarray] = None                     # last window's averaged fields, cid order


def rho_u_from_acc(acc: np.ndarray, NF: int) -> np.ndarray:
    """(rho, ux, uy, uz) in cid order from averaged fields (ux, uy, uz, rho)."""
    a = np.asarray(acc, np.float32).reshape(4, NF)
    return np.concatenate([a[3], a[0], a[1], a[2]]).astype(np.float32)


def solve_to_steady(dev: abx.Device, g: abx.Geometry, p: dict, window: int = 500, tol_gamma: float = 0.005,
                    tol_zeta: float = 0.02, max_steps: int = 60000, precision: str = "f32",
                    init_rho_u: Optional[np.ndarray] = None, keep_acc: bool = False,
                    init_fn: Optional[Callable[[abx.Sim], None]] = None) -> SteadyResult:
    """The criterion above, run by the library (abx_sim_solve_to_steady).
    init_rho_u: equilibrium start from (rho, u) in cid order; init_fn: any
    start set on the fresh handle (its time is not part of wall_s)."""
    sim = abx.Sim(dev, g.grid(), p["K"], p["aniso"], p["u"], p["tau"], precision=precision)
    try:
        if init_rho_u is not None:
            sim.init_eq(init_rho_u)
        if init_fn is not None:
            init_fn(sim)
        sim.sync()
        t0 = time.perf_counter()
        r = sim.solve_to_steady(g.eval_view(), window=window, tol_gamma=tol_gamma, tol_zeta=tol_zeta,
                                max_steps=max_steps)
        sim.sync()
        wall = time.perf_counter() - t0
        acc = sim.read_acc(window) if keep_acc else None
        wins = [(w[0], w[3], w[4]) for w in r["windows"]]
        return SteadyResult(r["converged"], r["steps"], wall, r["metrics"], wins, acc)
    finally:
        sim.destroy()


def fp64_errors(dev: abx.Device, g: abx.Geometry, p: dict, precision: str = "f32") -> dict:
    """Run the fixed-step solve (p = {K, aniso, u, tau, steps, avg, every}) in
    `precision` and in FP64; report the error of every result against FP64."""
    out = {}
    runs = {}
    for prec in (precision, "f64"):
        res = simulate_hip(dev, g, p, precision=prec, keep_sim=True)
        sim = res.sim
        ev = sim.evaluate(g.eval_view(), divisor=max(res.nacc, 1), field3d=True)
        acc = sim.read_acc(max(res.nacc, 1), f64=True)
        runs[prec] = dict(metrics=ev["metrics"], acc=acc.reshape(4, -1), vol=ev["vol"], hist=res.hist)
        sim.destroy()
    a, r = runs[precision], runs["f64"]
    for k, v in r["metrics"].items():
        out[f"metric_{k}"] = dict(value=a["metrics"][k], fp64=v, abs_err=abs(a["metrics"][k] - v),
                                  rel_err=abs(a["metrics"][k] - v) / abs(v) if v else None)
    u32, u64 = a["acc"][:3], r["acc"][:3]
    out["field_u_rel_l2"] = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    out["field_rho_minus_1_rel_l2"] = float(np.linalg.norm(a["acc"][3] - r["acc"][3]) /
                                           max(np.linalg.norm(r["acc"][3] - 1.0), 1e-300))
    for k in ("spd", "pst", "loss"):
        x, y = a["vol"][k].astype(np.float64), r["vol"][k].astype(np.float64)
        m = np.isfinite(y)
        out[f"field3d_{k}_rel_l2"] = float(np.linalg.norm(x[m] - y[m]) / max(np.linalg.norm(y[m]), 1e-300))
    h32 = np.array([row[1] for row in a["hist"]])
    h64 = np.array([row[1] for row in r["hist"]])
    fin = np.isfinite(h64)
    out["hist_gamma_max_abs_err"] = float(np.max(np.abs(h32[fin] - h64[fin]))) if fin.any() else None
    return out


def fixed_run(dev: abx.Device, g: abx.Geometry, p: dict, steps: int = 6000, avg: int = 2000, precision: str = "f32",
              init_rho_u: Optional[np.ndarray] = None,
              init_fn: Optional[Callable[[abx.Sim], None]] = None) -> SteadyResult:
    """The reference's protocol (simulate, engine.js:341-358 / 613-630): a
    fixed number of steps from the start state, averaging the last `avg`;
    evaluated on the GPU.  converged is True by definition.  init_fn: any
    start set on the fresh handle (its time is not part of wall_s)."""
    sim = abx.Sim(dev, g.grid(), p["K"], p["aniso"], p["u"], p["tau"], precision=precision)
    try:
        if init_rho_u is not None:
            sim.init_eq(init_rho_u)
        if init_fn is not None:
            init_fn(sim)
        sim.sync()
        t0 = time.perf_counter()
        sim.steps(steps - avg, False)
        sim.steps(avg, True)
        sim.probe(steps)
        m = sim.evaluate(g.eval_view(), divisor=avg, field3d=False)["metrics"]
        sim.sync()
        wall = time.perf_counter() - t0
        return SteadyResult(True, steps, wall, m, [(steps, m["gammaCore"], m["zeta"])], sim.read_acc(avg))
    finally:
        sim.destroy()
