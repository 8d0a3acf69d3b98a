"""Throughput of the D3Q19 stream-collide on one MI355X.

Workload: the roofline airbox (300 x 150 x 200 mm, 6 mm walls, 70 mm inlet on
the Z-min face, seal plane at Y=0 with a 40 mm pleat layer K=3 aniso=10, lid
outlet 60 mm, 50 mm lid plenum) at dx = 1 mm, seeded internal baffles and an
oblique inlet (seed 1), FP32, u = 0.06, tau = 0.51.  One step is one fused
stream-collide pass over every fluid cell.  The flow first develops from rest
for --spinup untimed steps (default 3000), then --warmup steps run untimed.
The timed region is the reference's time loop (simulateGPU,
engine.js:613-630) over --steps steps: chunks of 100 steps, a
gamma/mass-ratio probe after each chunk and at the end, the last third of the
steps accumulating the averaged fields.  It ends at the last probe, where the
reference takes its MLUPS (engine.js:338); the read-back of the averaged
fields is reported apart (readback_ms).  Inputs are resident in HBM before
timing starts.

Prints one JSON line.  MLUPS counts fluid cells only (engine.js:338).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
BYTES_PER_UPDATE = {"f32": 152, "f32_f64": 152, "f64": 304}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(dx: float, seed: int):
    from airboxcalc_amd import geometry as G
    spec, _ = G.seeded_variant(seed, "roofline")
    p = G.prepare_params(spec, dx, outlet=True, click=True)
    return spec, G.airbox_triangles(spec), p


def host_cpu() -> dict:
    """Host CPU of this machine: logical CPUs, the ones this process may use,
    and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def engine_js_rate() -> dict:
    """engine.js's own CPU backend on the 4 mm airbox (161,144 fluid cells,
    2000 steps): the wall time its simulate() recorded when the golden
    fixture tests/golden/c1.json was written (Node v12, one thread, 8 x Xeon
    @ 2.70 GHz build host; not this machine)."""
    with open(os.path.join(ROOT, "tests", "golden", "c1.json")) as fh:
        sim = json.load(fh)["sim"]["sim"]
    steps = 2000
    return {"value": round(sim["NF"] * steps / (sim["ms"] / 1e3) / 1e6, 3), "unit": "MLUPS", "cores": 1,
            "kind": "reference",
            "sample": f"engine.js simulate() of the dx=4 mm airbox, {sim['NF']} fluid cells x {steps} steps in "
                      f"{sim['ms'] / 1e3:.1f} s on one thread (Node v12, 8 x Xeon @ 2.70 GHz build host), "
                      f"from tests/golden/c1.json"}


def cpu_baseline(tris, p, run, seconds: float, threads: int):
    """The oracle's C restatement of cpuStep on the same geometry, timed on
    this host for a bounded number of steps: on `threads` OpenMP threads (the
    reported value) and on one thread."""
    from oracle import oracle as O
    g = O.prepare(tris, p, max_cells=0)
    S = O.setup_sim(g, run["K"], run["aniso"], run["u"], run["tau"])
    f = S.f0()
    f2 = np.empty_like(f)

    def timed(nt, budget):
        nonlocal f, f2
        t0 = time.perf_counter()
        n = 0
        while True:
            O.cpu_step(S, f, f2, threads=nt)
            f, f2 = f2, f
            n += 1
            el = time.perf_counter() - t0
            if el > budget or n >= 400:
                return n, el

    n1, el1 = timed(1, seconds / 3)
    nt, elt = timed(threads, 2 * seconds / 3)
    one = S.NF * n1 / el1 / 1e6
    return dict(value=S.NF * nt / elt / 1e6, unit="MLUPS", cores=threads, kind="port",
                one_core_mlups=round(one, 2), host=host_cpu(), reference_engine_js=engine_js_rate(),
                sample=f"{nt} cpuStep steps (f32 storage, f64 arithmetic) of the same {S.NF}-cell lattice on "
                       f"{threads} OpenMP threads in {elt:.1f} s, and {n1} steps on one thread in {el1:.1f} s "
                       f"({one:.2f} MLUPS); oracle/engine_oracle.c")


def bench_fp64(a, dev):
    """The roofline airbox (seed 1) at dx = 2 mm, 6000 steps averaging 2000,
    in FP32 and in FP64: the error of every FP32 result against FP64, with
    the FP64 kernel's own throughput."""
    from airboxcalc_amd import abx, steady
    spec, tris, p = workload(2.0, a.seed)
    g = abx.Geometry(dev, tris, p, max_cells=0)
    run = dict(K=3.0, aniso=10.0, u=0.06, tau=0.51, steps=a.steps, avg=a.steps // 3, every=a.every)
    t = time.perf_counter()
    e = steady.fp64_errors(dev, g, run, precision=a.precision)
    wall = time.perf_counter() - t
    sim = abx.Sim(dev, g.grid(), precision="f64", K=3.0, aniso=10.0, u=0.06, tau=0.51)
    sim.steps(20)
    ms = sim.time_steps(200) / 200
    NF = sim.NF
    sim.destroy()
    out = {"workload": f"airbox 300x150x200 mm, dx=2 mm, seed {a.seed}, {a.steps} steps averaging {a.steps // 3}",
           "grid": [g.nx, g.ny, g.nz], "fluid_cells": NF, "precision": a.precision, "wall_s_both_runs": round(wall, 2),
           "fp64_kernel_ms_per_step": round(ms, 5), "fp64_glups": round(NF / ms / 1e6, 2),
           "fp64_hbm_gbps_algorithmic": round(NF * 304 / ms / 1e6, 1), "errors": e}
    print(json.dumps(out), flush=True)


def bench_warmstart(a, dev):
    """Time to steady, warm against cold start: held-out boxes of the
    surrogate family (seeds 1000000 + i, --variants of them, default 16),
    each written as a ~200 k-triangle STL and prepared on the GPU at --dx
    (2 mm unless given), solved cold and from the surrogate's prediction
    (the 1500-box 2 mm-target network at 2 mm, the 4.8 mm network at 4.8 mm).
    Reports the wall speedup to the steady criterion (windows of 1000 steps,
    gamma-core within 0.005 and zeta within 2 % between windows) and of the
    reference's fixed protocol without its transient (1500 steps averaging
    1500 from the prediction against 6000 averaging 2000 from rest)."""
    from airboxcalc_amd import surrogate as S, unet as U
    dx = a.dx if a.dx != 1.0 else 2.0
    n = a.variants if a.variants != 64 else 16
    weights = S.WEIGHTS if abs(dx - S.FAMILY_DX) < 1e-9 else S.WEIGHTS_DX2
    net = U.UNet(seed=0)
    norm, meta = S.load(weights, net)
    boxes = S.family(range(1_000_000, 1_000_000 + n))
    t = time.perf_counter()
    rep = S.warm_start_study(dev, net, norm, boxes, dx=dx, fp64_every=0, warm_steps=1500, warm_avg=1500,
                             log=lambda m: log(m))
    wall = time.perf_counter() - t
    sm = rep["summary"]
    ok = [r for r in rep["rows"] if "error" not in r]
    ms = 1e3 * sum(r["warm"]["wall_s"] for r in ok) / max(sum(r["warm"]["steps"] for r in ok), 1)
    both = sm.get("wall_speedup_both_within")
    out = {"metric": "time-to-steady (gamma +-0.005, zeta +-2 %) warm vs cold start: wall speedup",
           "value": round(both, 3) if both is not None else None, "unit": "x", "n_gpus": 1,
           "value_definition": "total cold wall time / total warm wall time (prediction included) to the steady "
                               "criterion, over the boxes whose warm and cold answers both lie within gamma-core "
                               "+-0.005 and zeta +-2 % of the reference's fixed 6000/2000 result",
           "boxes_both_within": sm.get("n_both_within"),
           "warm_within_tol_of_fixed": sm.get("warm_within_tol_of_fixed"),
           "cold_within_tol_of_fixed": sm.get("cold_within_tol_of_fixed"),
           "wall_speedup_all_boxes": round(sm["wall_speedup_total_incl_predict"], 3),
           "steps": int(sm["steps_warm_median"]), "warmup": 0, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 (lattice), bf16 (surrogate)",
           "data": f"synthetic (procedural airbox STLs, held-out seeds 1000000..{1000000 + n - 1})",
           "config": {"workload": f"{n} held-out airboxes at dx={dx:g} mm, ~200k-triangle STL each; "
                                  f"cold and warm to the steady criterion (windows of 1000 steps)",
                      "weights": os.path.relpath(weights, ROOT), "boxes": n, "dx_mm": dx},
           "summary": sm, "wall_s": round(wall, 1), "roofline": None, "cpu_baseline": None}
    print(json.dumps(out, default=float), flush=True)


def bench_batch(a, dev):
    """Batched variants: 64 seeded airboxes (dims 200-360 mm, dx in {2, 2.5,
    3, 4} mm, K 1.5-6, aniso 5-20) solved together, 2000 steps averaging 500
    unless --steps says otherwise; two variants are solved again alone and
    must give the same metrics bit for bit."""
    from airboxcalc_amd import batch as BT
    steps = a.steps if a.steps != 6000 else 2000
    avg, every = steps // 4, a.every
    t = time.perf_counter()
    variants = BT.seeded_variants(range(a.variants))
    geoms = BT.prepare_all(dev, variants)
    log(f"prepared {len(geoms)} variants in {time.perf_counter() - t:.1f}s; cells "
        f"{min(g.nf for g in geoms)}..{max(g.nf for g in geoms)}")
    r = BT.solve_batch(dev, variants, geoms, steps, avg, every, precision=a.precision)
    log(f"batch {r['NF']} fluid cells, {steps} steps in {r['wall_s']:.2f}s, tune {r['tune']}")
    same = []
    for i in (0, len(variants) - 1):
        one = BT.solve_one(dev, variants[i], geoms[i], steps, avg, every, precision=a.precision)
        same.append(all(one["metrics"][k] == r["results"][i]["metrics"][k] for k in one["metrics"]))
    bpu = BYTES_PER_UPDATE[a.precision]
    achieved = r["NF"] * bpu / (r["kernel_ms_per_step"] * 1e-3) / 1e9
    out = {
        "metric": "MLUPS + HBM GB/s (D3Q19 stream-collide, fluid-cell updates), batched variants",
        "value": round(r["mlups"], 1), "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(r["wall_s"] * 1e3 / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": a.precision, "data": "synthetic (procedural airbox STLs, seeds 0..63)",
        "config": {"workload": f"{len(variants)} seeded airboxes, dims 200-360 mm, dx in {{2,2.5,3,4}} mm, K 1.5-6, "
                               f"aniso 5-20, one batch, {steps} steps averaging {avg}",
                   "fluid_cells_total": r["NF"], "variants": len(variants),
                   "cells_min": min(x["NF"] for x in r["results"]), "cells_max": max(x["NF"] for x in r["results"])},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "k_batch_step (AA pattern, steps without averaging)",
                     "avg_launch_ms": round(r["kernel_ms_per_step"], 5)},
        "tune": r["tune"], "batch_equals_individual": same,
        "per_variant": [dict(seed=x["seed"], dx=x["dx"], NF=x["NF"], gammaCore=round(x["metrics"]["gammaCore"], 6),
                             zeta=round(x["metrics"]["zeta"], 6)) for x in r["results"]],
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def unet_flops(net_layers, n: int, dims) -> dict:
    """Algorithmic FLOPs of one training step of the surrogate: per conv layer
    2 * voxels * co * ci * taps for the forward, the input gradient (not for
    the first layer) and the weight gradient."""
    from airboxcalc_amd import unet as U
    fwd = dgrad = wgrad = 0
    for l in net_layers:
        d, h, w = dims[U.UNet._level(l)]
        f = 2 * n * d * h * w * l.co * l.ci * l.taps
        fwd += f
        wgrad += f
        if l.name != "enc1a":
            dgrad += f
    return dict(fwd=fwd, dgrad=dgrad, wgrad=wgrad, total=fwd + dgrad + wgrad)


def bench_unet(a, dev):
    """Surrogate training step (the config-4 network: 96 x 48 x 64 volume,
    batch 8, BF16 MFMA convolutions, Adam) on synthetic inputs."""
    import torch
    from airboxcalc_amd import unet as U
    from airboxcalc_amd import surrogate as S
    n = 8
    D, H, W = S.GRID
    steps = a.steps if a.steps != 6000 else 50
    warm = min(a.warmup, 10)
    g = torch.Generator(device="cpu").manual_seed(a.seed)
    x = torch.randn((n, D, H, W, U.IN_CH), generator=g).to(torch.bfloat16).cuda()
    y = torch.randn((n, D, H, W, 4), generator=g).to(torch.bfloat16).cuda()
    m = torch.ones((n, D, H, W), dtype=torch.uint8, device="cuda")
    cw = torch.ones(4, dtype=torch.float32, device="cuda")
    net = U.UNet(seed=0)
    count = n * D * H * W   # every voxel of the mask is set
    for _ in range(warm):
        net.train_step(x, y, m, cw, count=count)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = net.train_step(x, y, m, cw, count=count)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / steps
    fl = unet_flops(net.L, n, net.dims)
    tflops = fl["total"] / (ms * 1e-3) / 1e12
    out = {
        "metric": "surrogate training steps/s (U-Net, BF16 MFMA convolutions)", "value": round(1e3 / ms, 2),
        "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)",
        "data": "synthetic (random volumes)",
        "config": {"workload": f"U-Net 16/32/64/128, batch {n}, volume {D}x{H}x{W}, forward + backward + Adam",
                   "params": U.n_params(), "flops_per_step": fl},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 5), "traffic": None,
                     "kernel": "all kernels of one training step (convolutions dominate)"},
        "loss": float(loss.item()), "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_results(a, dev):
    """The result stage of a solve on the 1 mm airbox (the default workload):
    the probe (gamma and mass ratio, engine.js:324-339, 452-462), evaluate
    with and without field3d (engine.js:632-719) and the streamlines
    (engine.js:721-758), each timed per call with its host copies (the
    reference's results are host arrays too), after --steps steps (default
    600, the last third averaged).  Beside them, the oracle's C restatement
    of the same functions on the same averaged field, one host core, timed
    once each."""
    import dataclasses
    from airboxcalc_amd import abx
    from airboxcalc_amd import geometry as G
    spec, tris, p = workload(a.dx, a.seed)
    run = dict(K=3.0, aniso=10.0, u=0.06, tau=0.51)
    # set-up stages: the STL reader on the same airbox subdivided to about 200 k triangles, prepare of both
    # meshes, the simulation handle
    big = G.airbox_triangles(dataclasses.replace(spec, subdiv=51))
    stl = G.stl_binary_bytes(big)
    t0 = time.perf_counter()
    parsed = abx.parse_stl(stl)
    parse_ms = (time.perf_counter() - t0) * 1e3
    assert parsed.size == big.size
    t0 = time.perf_counter()
    abx.Geometry(dev, parsed, p, max_cells=0)
    prep_big_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    g = abx.Geometry(dev, tris, p, max_cells=0)
    prep_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    sim = abx.Sim(dev, g.grid(), precision=a.precision, **run)
    sim.sync()
    create_ms = (time.perf_counter() - t0) * 1e3
    steps = a.steps if a.steps != 6000 else 600
    avg = steps // 3
    sim.steps(steps - avg)
    sim.steps(avg, True)
    sim.sync()
    ev = g.eval_view()

    def per_call(fn, n):
        fn()
        sim.sync()
        t0 = time.perf_counter()
        for _ in range(n):
            out = fn()
        sim.sync()
        return (time.perf_counter() - t0) * 1e3 / n, out

    probe_ms, _ = per_call(lambda: sim.probe(steps), 50)
    eval_ms, res = per_call(lambda: sim.evaluate(ev, divisor=avg, field3d=False), 10)
    eval3d_ms, res3 = per_call(lambda: sim.evaluate(ev, divisor=avg, field3d=True), 10)
    lines_ms, L = per_call(lambda: sim.streamlines(ev, divisor=avg), 10)
    acc = sim.read_acc(avg)
    out = {
        "metric": "result stage ms per call (evaluate + field3d on the averaged field)",
        "value": round(eval3d_ms, 3), "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32": "f32", "f32_f64": "f32 storage, f64 arithmetic", "f64": "f64"}[a.precision],
        "data": "synthetic (procedural airbox STL, seeded)",
        "config": {"workload": f"airbox 300x150x200 mm, dx={a.dx:g} mm, seed {a.seed}, results after {steps} steps "
                               f"averaging {avg}", "grid": [g.nx, g.ny, g.nz], "fluid_cells": sim.NF},
        "setup": {"stl_triangles": int(big.size // 9), "stl_bytes": len(stl), "parse_stl_ms": round(parse_ms, 2),
                  "prepare_ms": round(prep_ms, 1), "prepare_200k_triangles_ms": round(prep_big_ms, 1),
                  "sim_create_ms": round(create_ms, 1)},
        "probe_ms": round(probe_ms, 4), "evaluate_ms": round(eval_ms, 3), "evaluate_field3d_ms": round(eval3d_ms, 3),
        "streamlines_ms": round(lines_ms, 3), "streamline_points": int(L["spd"].size),
        "gamma": res["metrics"]["gamma"], "zeta": res["metrics"]["zeta"],
        "roofline": None, "cpu_baseline": None,
    }
    if not a.no_cpu:
        from oracle import oracle as O
        og = O.prepare(tris, p, max_cells=0)
        S = O.setup_sim(og, run["K"], run["aniso"], run["u"], run["tau"])
        f = S.f0()
        t0 = time.perf_counter()
        O.probe(S, f, steps)
        t_probe = time.perf_counter() - t0
        t0 = time.perf_counter()
        oe = O.evaluate(og, S, run["u"], acc)
        t_eval = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.field3d(og, S, run["u"], acc, oe["fluid"], oe["rhoRef"])
        t_f3 = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.streamlines(og, S, run["u"], acc, oe["fluid"])
        t_lines = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round((t_eval + t_f3) * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
            "probe_ms": round(t_probe * 1e3, 2), "evaluate_ms": round(t_eval * 1e3, 1),
            "field3d_ms": round(t_f3 * 1e3, 1), "streamlines_ms": round(t_lines * 1e3, 1),
            "same_gamma": abs(oe["metrics"]["gamma"] - res["metrics"]["gamma"]) <= 1e-6,
            "sample": "oracle/engine_oracle.c probe, evaluate, field3d and streamlines once each on the GPU's "
                      "averaged field, one thread", "host": host_cpu()}
    print(json.dumps(out), flush=True)


def bench_conv(a, dev):
    """The surrogate's 16-channel 3x3x3 convolutions (batch 8, 96 x 48 x 64):
    forward, the 16-channel kernel (its weights held in registers, read
    from LDS, or its brick and weights staged through registers) against
    the general one, and the weight
    gradient: every tap per workgroup (k_wgrad27, the default) against the
    tap-group kernels with transposing LDS reads and with gathering; each
    timed with HIP events over --steps calls."""
    import torch
    from airboxcalc_amd import abx
    L = abx.load()
    s = torch.cuda.current_stream().cuda_stream
    calls = a.steps if a.steps != 6000 else 20
    g = torch.Generator(device="cpu").manual_seed(a.seed)
    rows = []
    n, d, h, w = 8, 96, 48, 64
    for ci, co in ((16, 16), (48, 16), (32, 16), (16, 48)):
        x = torch.randn((n, d, h, w, ci), generator=g).to(torch.bfloat16).cuda()
        wf = (torch.randn((co * 27 * ci,), generator=g) / 20).float().cuda()
        wp = torch.empty(co * 27 * ci, dtype=torch.bfloat16, device="cuda")
        assert L.abx_conv3d_pack(s, wf.data_ptr(), co, 3, ci, 0, wp.data_ptr()) == 0
        y = torch.empty((n, d, h, w, co), dtype=torch.bfloat16, device="cuda")
        us = {}
        try:
            # flags: bit 0 k_conv_fwd16 (else the general kernel), bit 1 (default) transposed weight gradient,
            # bit 3 brick and weights through registers (else straight into LDS), bit 6 the weights read from
            # LDS for every MFMA (else held in registers; three workgroups per CU, two with bit 4)
            for k in (2, 3, 11, 67, 83):
                L.abx_conv3d_set_kernel(k)
                for _ in range(3):
                    assert L.abx_conv3d_forward(s, x.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), None, co, 3, 1, 0,
                                                y.data_ptr(), co) == 0
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(calls):
                    L.abx_conv3d_forward(s, x.data_ptr(), ci, n, d, h, w, ci, wp.data_ptr(), None, co, 3, 1, 0,
                                         y.data_ptr(), co)
                e1.record()
                torch.cuda.synchronize()
                us[k] = e0.elapsed_time(e1) / calls * 1e3
        finally:
            L.abx_conv3d_set_kernel(3)
        # weight gradient of the same shape: k_wgrad27 (flags 3; 35: its tiles staged through registers), the
        # tap-group kernel on transposing LDS reads (7) and gathering (5)
        dz = torch.randn((n, d, h, w, co), generator=g).to(torch.bfloat16).cuda()
        wsb = L.abx_conv3d_wgrad_workspace(n, d, h, w, ci, co, 3)
        ws = torch.empty(wsb // 4, dtype=torch.float32, device="cuda")
        dw = torch.empty(co * 27 * ci, dtype=torch.float32, device="cuda")
        db = torch.empty(co, dtype=torch.float32, device="cuda")
        wg = {}
        try:
            for k in (5, 7, 3, 35):
                L.abx_conv3d_set_kernel(k)
                for _ in range(3):
                    assert L.abx_conv3d_wgrad(s, x.data_ptr(), ci, dz.data_ptr(), co, n, d, h, w, ci, co, 3,
                                              ws.data_ptr(), wsb, dw.data_ptr(), db.data_ptr(), 0) == 0
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(calls):
                    L.abx_conv3d_wgrad(s, x.data_ptr(), ci, dz.data_ptr(), co, n, d, h, w, ci, co, 3, ws.data_ptr(),
                                       wsb, dw.data_ptr(), db.data_ptr(), 0)
                e1.record()
                torch.cuda.synchronize()
                wg[k] = e0.elapsed_time(e1) / calls * 1e3
        finally:
            L.abx_conv3d_set_kernel(3)
        fl = 2 * n * d * h * w * ci * co * 27
        rows.append(dict(ci=ci, co=co, gflop=round(fl / 1e9, 2), fwd16_us=round(us[3], 1), general_us=round(us[2], 1),
                         fwd16_staged_us=round(us[11], 1), fwd16_lds_weights_us=round(us[67], 1),
                         fwd16_lds_weights_2wg_us=round(us[83], 1),
                         fwd16_tflops=round(fl / us[3] / 1e6, 1), general_tflops=round(fl / us[2] / 1e6, 1),
                         wgrad27_us=round(wg[3], 1), wgrad27_regs_us=round(wg[35], 1), wgrad_tr_us=round(wg[7], 1),
                         wgrad_gather_us=round(wg[5], 1),
                         wgrad27_tflops=round(fl / wg[3] / 1e6, 1), wgrad_tr_tflops=round(fl / wg[7] / 1e6, 1)))
    r0 = rows[0]
    out = {
        "metric": "3x3x3 bf16 convolution TFLOP/s (16 -> 16 channels, 96x48x64, batch 8)",
        "value": r0["fwd16_tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": calls, "warmup": 3,
        "ms_per_step": round(r0["fwd16_us"] / 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic (random volumes)",
        "config": {"workload": "the surrogate's 16-channel convolutions: forward (k_conv_fwd16 with its weights in "
                               "registers, read from LDS, or staged through registers, against k_conv_fwd) "
                               "and weight gradient (k_wgrad27 against k_conv_wgrad_tr and k_conv_wgrad)",
                   "shapes": rows},
        "roofline": {"bound": "mfma", "achieved": r0["fwd16_tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(r0["fwd16_tflops"] / 2500.0, 4), "traffic": None, "kernel": "k_conv_fwd16"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--dx", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--precision", default="f32", choices=["f32", "f32_f64", "f64"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the CPU baseline (default: up to 16, the host share of one GPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", type=int, default=None, help="stream-collide scheme: 4 AA pattern (default), 2 two lattices")
    ap.add_argument("--xcd-map", type=int, default=None, help="block order across XCDs (abx_sim_set_launch)")
    ap.add_argument("--nt", type=int, default=None, help="non-temporal access bits (abx_sim_set_launch)")
    ap.add_argument("--spinup", type=int, default=3000,
                    help="untimed steps from rest before the warmup, so the timed probes read a developed flow")
    ap.add_argument("--workload", default="roofline",
                    choices=["roofline", "batch", "fp64", "unet", "conv", "warmstart", "results"],
                    help="roofline: one dx=1 mm airbox (default); batch: 64 seeded uneven variants in one batch; "
                         "fp64: the FP64 error report of the roofline airbox at dx=2 mm; "
                         "unet: surrogate training steps; conv: the surrogate's 16-channel convolutions; "
                         "warmstart: time to steady, warm against cold, on held-out boxes at dx=2 mm; "
                         "results: probe, evaluate, field3d and streamlines of the roofline airbox")
    ap.add_argument("--variants", type=int, default=64)
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the library's float4 copy kernel (HBM ceiling and counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(the single-grid solve runs on one MI355X)", file=sys.stderr)
        sys.exit(2)

    from airboxcalc_amd import abx
    dev = abx.device(0)
    log("device", dev.name)
    copy_bytes = 1_600_000_000
    if a.copy_only:
        gbps, ms = abx.copy_bandwidth(dev, copy_bytes, 20)
        print(json.dumps({"copy_bytes": copy_bytes, "gbps": round(gbps, 1), "ms_per_copy": round(ms, 5)}))
        return
    if a.workload == "batch":
        return bench_batch(a, dev)
    if a.workload == "fp64":
        return bench_fp64(a, dev)
    if a.workload == "unet":
        return bench_unet(a, dev)
    if a.workload == "conv":
        return bench_conv(a, dev)
    if a.workload == "warmstart":
        return bench_warmstart(a, dev)
    if a.workload == "results":
        return bench_results(a, dev)
    spec, tris, p = workload(a.dx, a.seed)
    run = dict(K=3.0, aniso=10.0, u=0.06, tau=0.51)
    t = time.perf_counter()
    g = abx.Geometry(dev, tris, p, max_cells=0)
    log(f"prepare {time.perf_counter() - t:.2f}s grid {g.nx}x{g.ny}x{g.nz} nf={g.nf} ntri={tris.size // 9}")
    sim = abx.Sim(dev, g.grid(), precision=a.precision, **run)
    if a.kernel is not None or a.xcd_map is not None or a.nt is not None:
        sim.set_launch(kernel=4 if a.kernel is None else a.kernel, xcd_map=4 if a.xcd_map is None else a.xcd_map,
                       nt=19 if a.nt is None else a.nt)
    NF = sim.NF
    I = sim.info
    log(f"sim NF={NF} device bytes {I.device_bytes / 2**30:.2f} GiB, lattice plane {I.plane} positions "
        f"({I.nrows} rows), {I.nseg} active segments")

    # Spin-up (untimed): the flow develops from rest until it has crossed the
    # filter layer, so that the probes of the timed steps read a real flow.
    t = time.perf_counter()
    sim.steps(a.spinup)
    sim.steps(a.warmup)
    sim.sync()
    sim.clear_acc()
    sim.sync()
    log(f"spin-up {a.spinup} + warmup {a.warmup} steps in {time.perf_counter() - t:.2f}s")

    # Timed region: the reference's time loop (simulateGPU, engine.js:613-630),
    # ending at the last probe, where the reference takes its MLUPS
    # (engine.js:338).  The read-back of the averaged fields is timed apart.
    steps, every = a.steps, a.every
    avg = steps // 3
    start_acc = steps - avg
    ev_ms = {0: 0.0, 1: 0.0}
    ev_n = {0: 0, 1: 0}
    hist = []
    sim.sync()
    t0 = time.perf_counter()
    step = nacc = 0
    while step < steps:
        n = min(every, steps - step)
        if step < start_acc < step + n:
            n = start_acc - step
        doacc = 1 if step >= start_acc else 0
        ev_ms[doacc] += sim.time_steps(n, bool(doacc))
        ev_n[doacc] += n
        step += n
        if doacc:
            nacc += n
        if step % every == 0 or step == steps:
            gam, mr, _ = sim.probe(a.spinup + a.warmup + step)
            hist.append((step, gam, mr))
    sim.sync()
    t_solve = time.perf_counter() - t0
    t1 = time.perf_counter()
    acc = sim.read_acc(max(nacc, 1))
    t_read = time.perf_counter() - t1
    log(f"timed {steps} steps in {t_solve:.3f}s (+ read-back {t_read * 1e3:.1f} ms); last gamma {hist[-1][1]:.4f} "
        f"massRatio {hist[-1][2]:.4f}")
    assert np.isfinite(acc).all()

    copy_gbps, _ = abx.copy_bandwidth(dev, copy_bytes, 10)
    # the stream-collide's own access pattern over the dense lattice (all positions active)
    aa_gbps, _ = abx.aa_pattern_bandwidth(dev, int(I.plane), 10)
    bpu = BYTES_PER_UPDATE[a.precision]
    # the AA steps' own loads and stores on this lattice, collision left out
    # (after the solve: it moves the populations)
    pattern = None
    if a.precision != "f64" and a.kernel in (None, 4):
        pe, po = sim.pattern_time(20)
        pattern = {"even_ms": round(pe, 5), "odd_ms": round(po, 5),
                   "gbps": round(NF * 152 / ((pe + po) / 2 * 1e-3) / 1e9, 1)}
    mlups = NF * steps / t_solve / 1e6
    kern_ms = ev_ms[0] / max(ev_n[0], 1)
    achieved = NF * bpu / (kern_ms * 1e-3) / 1e9
    kern_acc_ms = ev_ms[1] / max(ev_n[1], 1)
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", f"pmc_stream_collide_dx{a.dx:g}_{a.precision}.json")
    if os.path.exists(tf) and a.nt is None and a.kernel is None and a.xcd_map is None:
        with open(tf) as fh:
            pm = json.load(fh)
        traffic = pm.get("hbm_bytes_per_launch")
        traffic_src = (f"{os.path.relpath(tf, ROOT)} (rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this kernel on this "
                       f"workload, scaled by the copy-kernel calibration; not measured in this run)")
    out = {
        "metric": "MLUPS + HBM GB/s (D3Q19 stream-collide, fluid-cell updates)",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_solve * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f32_f64": "f32 storage, f64 arithmetic", "f64": "f64"}[a.precision],
        "data": "synthetic (procedural airbox STL, seeded)",
        "config": {"workload": f"airbox 300x150x200 mm, dx={a.dx:g} mm, seed {a.seed} "
                               f"(baffles {len(spec.baffles)}, inlet yaw {spec.yaw_deg} deg), "
                               f"{steps} steps averaging {avg}, probe every {every}, after {a.spinup} untimed "
                               f"spin-up steps from rest",
                   "grid": [g.nx, g.ny, g.nz], "fluid_cells": NF, "triangles": int(tris.size // 9),
                   "lattice_positions": int(I.plane), "lattice_rows": int(I.nrows),
                   "active_segments": int(I.nseg)},
        "value_definition": "fluid cells x steps / time from the start of the loop to its last probe (engine.js:338)",
        "readback_ms": round(t_read * 1e3, 2),
        "wall_s_total": round(t_solve + t_read, 4),
        "last_probe": {"step": a.spinup + a.warmup + hist[-1][0], "gamma": hist[-1][1], "massRatio": hist[-1][2]},
        "hbm_gbps": round(NF * steps * bpu / t_solve / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_stream_collide_seg, AA pattern (steps without averaging)",
                     "avg_launch_ms": round(kern_ms, 5), "bytes_per_launch": NF * bpu,
                     "avg_launch_ms_with_acc": round(kern_acc_ms, 5),
                     "copy_ceiling_gbps": round(copy_gbps, 1),
                     "frac_of_copy_ceiling": round(achieved / copy_gbps, 4),
                     "aa_pattern_ceiling_gbps": round(aa_gbps, 1),
                     "frac_of_aa_pattern_ceiling": round(achieved / aa_gbps, 4),
                     "lattice_pattern_ceiling": pattern,
                     "frac_of_lattice_pattern_ceiling": round(achieved / pattern["gbps"], 4) if pattern else None},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(tris, p, run, a.cpu_seconds, a.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
