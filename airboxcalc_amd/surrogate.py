"""Flow surrogate for warm starts: dataset from the LBM, training, inference.

The reference solves every box from rest, f = w_q (engine.js:260-261), for a
fixed number of steps.  Here a 3-D U-Net (unet.py) trained on converged LBM
solutions of seeded boxes predicts the averaged steady fields of a new box;
the prediction seeds the lattice at equilibrium (abx_sim_init_eq) and the
solve then runs to the steady criterion of steady.py.

Data layout.  A box is described on a fixed volume of GRID = 96 x 48 x 64
lattice cells (i, j, k in the reference's lattice order, k fastest,
engine.js:128-134), at the cell size FAMILY_DX that makes every box of the
family fit; cells beyond the box's own lattice are solid.  Input channels
(bf16, IN_CH wide, the rest zero):

  0 solid          1 - fluid
  1 porous         filter layer cells (setupSim flag 3, engine.js:254)
  2 inlet          inlet cells (flag 1: k = 0 and opening, engine.js:252)
  3 K              K / 6 on porous cells
  4 wall distance  signed Euclidean distance to the wall in cells / 8
                   (positive in the fluid)
  5 outlet         outlet cells (flag 2, engine.js:253)
  6 aniso          aniso / 20 on porous cells

Targets: ux/u0, uy/u0, uz/u0, (rho - 1)/RHO_SCALE of the averaged fields,
masked to fluid cells, then standardised per channel with the dataset's
fluid-cell mean and standard deviation (kept with the weights, undone by
predict).
"""
from __future__ import annotations

import copy
import os
import time
from dataclasses import dataclass
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import abx
from . import geometry as G
from . import unet as U

GRID = (96, 48, 64)
FAMILY_DX = 4.8          # mm; the largest box of the family (yaw 30 deg) is 86 x 45 x 63 cells
U0, TAU = 0.06, 0.51
RHO_SCALE = 0.1          # (rho - 1) of these boxes is of order 0.1
N_FEAT = 7


@dataclass
class Box:
    seed: int
    spec: G.AirboxSpec
    K: float
    aniso: float


def family(seeds) -> List[Box]:
    """The surrogate family: the roofline boxes (300 x 150 x 200 mm, 0-3
    baffles, inlet yaw 0-30 deg) with K uniform 1.5-6 and aniso uniform 5-20."""
    out = []
    for s in seeds:
        spec, ex = G.seeded_variant(int(s), "surrogate")
        out.append(Box(int(s), spec, ex["K"], ex["aniso"]))
    return out


def cell_types(grid: abx.GridArrays):
    """fluid, inlet, outlet, porous masks (nx, ny, nz) in setupSim's order of
    precedence (engine.js:250-255)."""
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    fluid = grid.fluid.reshape(nx, ny, nz).astype(bool)
    inlet = np.zeros_like(fluid)
    inlet[:, :, 0] = grid.opening.reshape(nx, ny).astype(bool)
    inlet &= fluid
    outlet = grid.outlet.reshape(nx, ny, nz).astype(bool) & fluid & ~inlet
    dc = grid.dc
    pj = (dc > dc[grid.top]) & (dc.astype(np.float64) <= grid.pleat)
    porous = fluid & pj[None, :, None] & ~inlet & ~outlet
    return fluid, inlet, outlet, porous


def _place(a: np.ndarray, shape, fill=0):
    out = np.full(shape + a.shape[3:], fill, a.dtype)
    nx, ny, nz = a.shape[:3]
    if nx > shape[0] or ny > shape[1] or nz > shape[2]:
        raise abx.AbxError(f"lattice {nx}x{ny}x{nz} does not fit the surrogate volume {shape}")
    out[:nx, :ny, :nz] = a
    return out


def features(grid: abx.GridArrays, K: float, aniso: float, shape=GRID) -> np.ndarray:
    """Input volume (shape + (IN_CH,)) float32."""
    from scipy import ndimage
    fluid, inlet, outlet, porous = cell_types(grid)
    fl = _place(fluid, shape, False)
    d_in = ndimage.distance_transform_edt(fl)
    d_out = ndimage.distance_transform_edt(~fl)
    x = np.zeros(shape + (U.IN_CH,), np.float32)
    x[..., 0] = ~fl
    por = _place(porous, shape, False)
    x[..., 1] = por
    x[..., 2] = _place(inlet, shape, False)
    x[..., 3] = por * (K / 6.0)
    x[..., 4] = np.clip((d_in - d_out) / 8.0, -4.0, 4.0)
    x[..., 5] = _place(outlet, shape, False)
    x[..., 6] = por * (aniso / 20.0)
    return x


def features_device(out, grid: abx.GridArrays, K: float, aniso: float, shape=GRID) -> None:
    """features() computed on the GPU (abx_surrogate_features) into `out`, a
    contiguous bf16 device tensor of shape + (IN_CH,), on torch's current
    stream: the same values as features() converted to bf16."""
    import torch
    D, H, W = shape
    if tuple(out.shape) != (D, H, W, U.IN_CH) or out.dtype != torch.bfloat16 or not out.is_contiguous():
        raise ValueError("features_device: out must be a contiguous bf16 tensor of shape GRID + (IN_CH,)")
    if grid.nx > D or grid.ny > H or grid.nz > W:
        raise abx.AbxError(f"lattice {grid.nx}x{grid.ny}x{grid.nz} does not fit the surrogate volume {shape}")
    dev = out.device
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fl, op, ou, dc = up(grid.fluid), up(grid.opening), up(grid.outlet), up(grid.dc)
    lib = abx.load()
    nb = int(lib.abx_surrogate_features_workspace(D, H, W))
    ws = torch.empty((nb + 3) // 4, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    abx._check(lib.abx_surrogate_features(stream, fl.data_ptr(), op.data_ptr(), ou.data_ptr(), dc.data_ptr(),
                                          grid.nx, grid.ny, grid.nz, grid.top, float(grid.pleat), float(K),
                                          float(aniso), D, H, W, ws.data_ptr(), nb, out.data_ptr(), U.IN_CH,
                                          U.IN_CH))


def targets(grid: abx.GridArrays, acc: np.ndarray, shape=GRID):
    """(target volume shape + (4,) float32, fluid mask shape uint8) from the
    averaged fields acc (4*NF, components ux, uy, uz, rho, cid order)."""
    fluid = grid.fluid.reshape(grid.nx, grid.ny, grid.nz).astype(bool)
    NF = int(fluid.sum())
    a = np.asarray(acc, np.float32).reshape(4, NF)
    y = np.zeros((grid.nx, grid.ny, grid.nz, 4), np.float32)
    y[fluid, 0] = a[0] / U0
    y[fluid, 1] = a[1] / U0
    y[fluid, 2] = a[2] / U0
    y[fluid, 3] = (a[3] - 1.0) / RHO_SCALE
    return _place(y, shape), _place(fluid.astype(np.uint8), shape)


def rho_u_of_prediction(pred: np.ndarray, grid: abx.GridArrays) -> np.ndarray:
    """(rho, ux, uy, uz) in cid order (abx_sim_init_eq) from a predicted volume
    on the same lattice (pred: shape + (>= 4,) float)."""
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    fluid = grid.fluid.reshape(nx, ny, nz).astype(bool)
    p = np.asarray(pred[:nx, :ny, :nz, :4], np.float32)[fluid]
    return np.concatenate([1.0 + RHO_SCALE * p[:, 3], U0 * p[:, 0], U0 * p[:, 1], U0 * p[:, 2]]).astype(np.float32)


def resample(pred: np.ndarray, gc: abx.Geometry, gf: abx.Geometry) -> np.ndarray:
    """Trilinear resampling of a prediction on the lattice of gc (same box, any
    cell size) to the fluid cells of gf, in cid order of gf: (rho, ux, uy, uz).
    Both lattices share the local frame; cell centres are origin + (i+0.5) h
    (engine.js:137-139).  Only fluid cells of gc contribute (weights
    renormalised); a cell with no fluid neighbour keeps rest (rho 1, u 0)."""
    cg, fg = gc.grid(), gf.grid()
    cflu = cg.fluid.reshape(cg.nx, cg.ny, cg.nz).astype(np.float32)
    P = np.asarray(pred[:cg.nx, :cg.ny, :cg.nz, :4], np.float32) * cflu[..., None]
    fflu = fg.fluid.reshape(fg.nx, fg.ny, fg.nz).astype(bool)
    idx = np.nonzero(fflu)
    oc, of = np.array(gc.origin), np.array(gf.origin)
    coords = []
    for ax, n in enumerate((cg.nx, cg.ny, cg.nz)):
        pos = of[ax] + (idx[ax] + 0.5) * gf.h
        coords.append(np.clip((pos - oc[ax]) / gc.h - 0.5, 0.0, n - 1.0))
    i0 = [np.minimum(np.floor(c).astype(np.int64), n - 2 if n > 1 else 0) for c, n in
          zip(coords, (cg.nx, cg.ny, cg.nz))]
    t = [c - i for c, i in zip(coords, i0)]
    acc = np.zeros((idx[0].size, 4), np.float64)
    wsum = np.zeros(idx[0].size, np.float64)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                ii, jj, kk = i0[0] + dx, i0[1] + dy, i0[2] + dz
                w = (t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2])
                w = w * cflu[ii, jj, kk]
                acc += w[:, None] * P[ii, jj, kk]
                wsum += w
    v = np.where(wsum[:, None] > 0, acc / np.maximum(wsum, 1e-30)[:, None], 0.0)
    return np.concatenate([1.0 + RHO_SCALE * v[:, 3], U0 * v[:, 0], U0 * v[:, 1], U0 * v[:, 2]]).astype(np.float32)


# ---------------------------------------------------------------------------
def restrict_targets(acc: np.ndarray, gf: abx.Geometry, gc: abx.Geometry, shape=GRID):
    """Targets on gc's lattice (as targets()) from the averaged fields acc
    (4*NF of gf, cid order) of a solve of the same box on the finer lattice
    gf: trilinear over gf's fluid cells at gc's fluid cell centres."""
    fg, cg = gf.grid(), gc.grid()
    vol, _ = targets(fg, acc, shape=(fg.nx, fg.ny, fg.nz))
    ru = resample(vol, gf, gc)                     # (rho, ux, uy, uz) on gc's fluid cells
    NF = ru.size // 4
    return targets(cg, np.concatenate([ru[NF:], ru[:NF]]), shape)


def make_dataset(dev: abx.Device, boxes: List[Box], steps: int, avg: int, chunk: int = 64,
                 log: Optional[Callable[[str], None]] = None, target_dx: Optional[float] = None,
                 norm: Optional[dict] = None) -> dict:
    """Solve every box (batched, one launch per step for `chunk` boxes, the
    last `avg` of `steps` steps averaged) and build the training volumes on
    the device: X (n, 96, 48, 64, IN_CH) bf16, Y (n, ..., 4) bf16, M uint8.
    With target_dx the solves run at that (finer) cell size and their fields
    are restricted to the FAMILY_DX lattice of the inputs.  The targets are
    standardised per channel over the fluid cells, with this dataset's mean
    and deviation or with a given norm (to continue training a network)."""
    import torch
    n = len(boxes)
    D, H, W = GRID
    cuda = torch.device("cuda", dev.ordinal)
    X = torch.zeros((n, D, H, W, U.IN_CH), dtype=torch.bfloat16, device=cuda)
    Y = torch.zeros((n, D, H, W, 4), dtype=torch.bfloat16, device=cuda)
    M = torch.zeros((n, D, H, W), dtype=torch.uint8, device=cuda)
    stats = dict(cells=0, solve_s=0.0, prep_s=0.0, lups=0.0)
    for c0 in range(0, n, chunk):
        part = boxes[c0:c0 + chunk]
        t0 = time.perf_counter()
        geoms, grids, sims, fines = [], [], [], []
        for b in part:
            tris = G.airbox_triangles(b.spec)
            g = abx.Geometry(dev, tris, G.prepare_params(b.spec, FAMILY_DX), max_cells=0)
            gr = g.grid()
            geoms.append(g)
            grids.append(gr)
            gs = g
            if target_dx is not None:
                gs = abx.Geometry(dev, tris, G.prepare_params(b.spec, target_dx), max_cells=0)
                fines.append(gs)
            sims.append(abx.Sim(dev, gs.grid(), K=b.K, aniso=b.aniso, u=U0, tau=TAU, precision="f32"))
        for k, (b, gr) in enumerate(zip(part, grids)):
            features_device(X[c0 + k], gr, b.K, b.aniso)
        t1 = time.perf_counter()
        bt = abx.Batch(dev, sims)
        tuned = bt.tune(10)
        bt.set_launch(tuned["spb"], tuned["persistent"], True)
        bt.steps(steps - avg, False)
        bt.steps(avg, True)
        bt.probe(steps)
        bt.sync()
        t2 = time.perf_counter()
        for k, (s, gr) in enumerate(zip(sims, grids)):
            if target_dx is not None:
                y, m = restrict_targets(s.read_acc(avg), fines[k], geoms[k])
            else:
                y, m = targets(gr, s.read_acc(avg))
            Y[c0 + k] = torch.from_numpy(y).to(torch.bfloat16)
            M[c0 + k] = torch.from_numpy(m)
        nf = sum(s.NF for s in sims)
        stats["cells"] += nf
        stats["prep_s"] += t1 - t0
        stats["solve_s"] += t2 - t1
        stats["lups"] += nf * steps
        bt.destroy()
        for s in sims:
            s.destroy()
        for g in geoms + fines:
            g.free()
        if log:
            log(f"dataset {c0 + len(part)}/{n}: prepare+features {t1 - t0:.1f} s, solve {t2 - t1:.1f} s "
                f"({nf * steps / max(t2 - t1, 1e-9) / 1e9:.1f} G fluid-LUPS)")
    # per-channel standardisation over the fluid cells
    s1 = torch.zeros(4, dtype=torch.float64, device=cuda)
    s2 = torch.zeros(4, dtype=torch.float64, device=cuda)
    cnt = 0
    for c0 in range(0, n, chunk):
        y, m = Y[c0:c0 + chunk].double(), M[c0:c0 + chunk].bool()
        v = y[m]
        s1 += v.sum(0)
        s2 += (v * v).sum(0)
        cnt += v.shape[0]
    mean = (s1 / max(cnt, 1)).float()
    std = torch.sqrt(torch.clamp(s2 / max(cnt, 1) - (s1 / max(cnt, 1)) ** 2, min=1e-12)).float()
    if norm is not None:
        mean = torch.as_tensor(np.asarray(norm["mean"], np.float32), device=cuda)
        std = torch.as_tensor(np.asarray(norm["std"], np.float32), device=cuda)
    for c0 in range(0, n, chunk):
        y = (Y[c0:c0 + chunk].float() - mean) / std
        Y[c0:c0 + chunk] = (y * M[c0:c0 + chunk, ..., None]).to(torch.bfloat16)
    torch.cuda.synchronize(cuda)
    return dict(X=X, Y=Y, M=M, seeds=[b.seed for b in boxes], stats=stats, steps=steps, avg=avg,
                norm=dict(mean=mean.cpu().numpy(), std=std.cpu().numpy()))


def train(net: U.UNet, data: dict, epochs: int, batch: int = 8, lr: float = 1e-3, val_frac: float = 0.05,
          seed: int = 0, log: Optional[Callable[[str], None]] = None, max_seconds: float = 1e30,
          schedule: bool = False, keep_best: bool = False) -> dict:
    """Adam on the masked MSE of the normalised targets; the last val_frac of
    the samples are held out.  Returns the per-epoch losses and timings.
    schedule: the learning rate follows a cosine from lr to lr / 10 over the
    run (its progress is the larger of the epoch and the time fraction);
    keep_best: the weights of the epoch with the lowest held-out loss are
    restored at the end."""
    import math
    import torch
    X, Y, M = data["X"], data["Y"], data["M"]
    n = X.shape[0]
    nval = max(1, int(round(n * val_frac))) if n > 1 else 0
    ntr = n - nval
    rng = np.random.default_rng(seed)
    cw = torch.ones(U.OUT_CH, dtype=torch.float32, device=X.device)
    # fluid voxels per sample, so a batch's mask count needs no device read-back
    mcount = M.reshape(n, -1).sum(dim=1, dtype=torch.int64).cpu().numpy()
    hist = []
    t_start = time.perf_counter()
    step_times = []
    best = (float("inf"), None, 0)
    for ep in range(epochs):
        progress = max(ep / epochs, (time.perf_counter() - t_start) / max_seconds)
        lr_ep = lr * (0.1 + 0.45 * (1.0 + math.cos(math.pi * min(progress, 1.0)))) if schedule else lr
        order = rng.permutation(ntr)
        tot, cnt = 0.0, 0
        losses = []
        for b0 in range(0, ntr - batch + 1, batch):
            idx = torch.as_tensor(order[b0:b0 + batch], device=X.device)
            xb, yb, mb = X[idx].contiguous(), Y[idx].contiguous(), M[idx].contiguous()
            torch.cuda.synchronize()
            ts = time.perf_counter()
            loss = net.train_step(xb, yb, mb, cw, lr=lr_ep, count=int(mcount[order[b0:b0 + batch]].sum()))
            losses.append(loss)
            torch.cuda.synchronize()
            step_times.append(time.perf_counter() - ts)
        tr = float(torch.stack(losses).mean().item()) if losses else float("nan")
        va = evaluate_loss(net, data, range(ntr, n), batch) if nval else float("nan")
        hist.append(dict(epoch=ep + 1, train=tr, val=va, lr=lr_ep, wall_s=time.perf_counter() - t_start))
        if keep_best and va < best[0]:
            best = (va, net.w.clone(), ep + 1)
        if log:
            log(f"epoch {ep + 1}: train {tr:.4f}  val {va:.4f}  lr {lr_ep:.2e}  ({time.perf_counter() - t_start:.0f} s, "
                f"{1e3 * np.median(step_times):.1f} ms/step)")
        if time.perf_counter() - t_start > max_seconds:
            break
    if keep_best and best[1] is not None:
        net.w.copy_(best[1])
        net.repack()
    return dict(history=hist, ms_per_step=1e3 * float(np.median(step_times)) if step_times else None,
                n_train=ntr, n_val=nval, best_epoch=best[2] if keep_best else None,
                best_val=best[0] if keep_best else None)


def evaluate_loss(net: U.UNet, data: dict, indices, batch: int = 8) -> float:
    import torch
    X, Y, M = data["X"], data["Y"], data["M"]
    cw = torch.ones(U.OUT_CH, dtype=torch.float32, device=X.device)
    idx = list(indices)
    tot = 0.0
    for b0 in range(0, len(idx), batch):
        sel = torch.as_tensor(idx[b0:b0 + batch], device=X.device)
        net.forward(X[sel].contiguous())
        loss, _ = net.loss_and_grad(Y[sel].contiguous(), M[sel].contiguous(), cw)
        tot += float(loss.item()) * len(sel)
    return tot / max(len(idx), 1)


def predict(net: U.UNet, x: np.ndarray, norm: dict) -> np.ndarray:
    """Forward of one input volume (GRID + (IN_CH,)) -> targets (GRID + (4,))
    float32 in the units of targets() (standardisation undone)."""
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x[None])).to(net.dev).to(torch.bfloat16)
    out = net.forward(xt)
    y = out[0, ..., :U.OUT_CH].float().cpu().numpy()
    return y * np.asarray(norm["std"], np.float32) + np.asarray(norm["mean"], np.float32)


def warm_start_init(net: U.UNet, norm: dict, dev: abx.Device, box: Box, g_fine: abx.Geometry, sim: abx.Sim,
                    tris: Optional[np.ndarray] = None, reuse: bool = True) -> None:
    """Seed sim (on g_fine's lattice) at the equilibrium of the surrogate's
    prediction without leaving the GPU: the box is prepared at FAMILY_DX, its
    features go through the network, and abx_sim_init_eq_resample maps the
    network's output (standardisation undone by the affine map) onto sim's
    fluid cells."""
    import torch
    tris = G.airbox_triangles(box.spec) if tris is None else tris
    # a solve at the network's own cell size already holds the same prepared
    # geometry (prepare is deterministic): reuse it instead of preparing twice
    same = reuse and abs(g_fine.h - FAMILY_DX) < 1e-9
    gc = g_fine if same else abx.Geometry(dev, tris, G.prepare_params(box.spec, FAMILY_DX), max_cells=0)
    try:
        grid = gc.grid()
        x = torch.empty((1,) + GRID + (U.IN_CH,), dtype=torch.bfloat16, device=net.dev)
        features_device(x[0], grid, box.K, box.aniso)
        out = net.forward(x)
        fl = torch.from_numpy(grid.fluid).to(net.dev)
        mean, std = np.asarray(norm["mean"], np.float64), np.asarray(norm["std"], np.float64)
        scale = np.array([U0, U0, U0, RHO_SCALE])
        affine = list(scale * std) + list(scale * mean + np.array([0.0, 0.0, 0.0, 1.0]))
        torch.cuda.synchronize(net.dev)
        sim.init_eq_resample(out.data_ptr(), U.HEAD_CH, GRID, (gc.nx, gc.ny, gc.nz), fl.data_ptr(), gc.origin, gc.h,
                             g_fine.origin, g_fine.h, affine)
    finally:
        if not same:
            gc.free()


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(DATA, "surrogate_unet_96x48x64.abxw")                 # trained at FAMILY_DX
WEIGHTS_DX2 = os.path.join(DATA, "surrogate_unet_96x48x64_dx2targets_1500.abxw")  # 2 mm solutions, restricted

# Inference weight file (.abxw), read by this module and by the C ABI's
# abx_surrogate_load (csrc/surrogate.cpp); little-endian:
#   char[8] "ABXUNET1"
#   int32   n_params, d, h, w, in_ch, head_ch, widths[4], meta_len, 0
#   float64 family_dx, u0_scale, rho_scale, mean[4], std[4]
#   meta_len bytes of JSON (training facts), zero padding to a multiple of 16
#   float32 w[n_params]   (layer by layer: W[co][taps][ci] then bias[co])
_ABXW_MAGIC = b"ABXUNET1"
_ABXW_HEAD = "<8s12i11d"


def write_weights(path: str, w: np.ndarray, norm: dict, meta: dict, grid=GRID, family_dx: float = FAMILY_DX) -> None:
    """The inference weight file of a trained network (FP32 weights and the
    target standardisation)."""
    import json
    import struct
    w = np.ascontiguousarray(w, np.float32)
    if w.size != U.n_params():
        raise ValueError(f"{w.size} weights, the network has {U.n_params()}")
    mb = json.dumps(meta).encode()
    head = struct.pack(_ABXW_HEAD, _ABXW_MAGIC, w.size, *grid, U.IN_CH, U.HEAD_CH, *U.WIDTHS, len(mb), 0,
                       float(family_dx), U0, RHO_SCALE, *[float(v) for v in norm["mean"]],
                       *[float(v) for v in norm["std"]])
    body = head + mb
    body += b"\0" * (-len(body) % 16)
    with open(path, "wb") as fh:
        fh.write(body)
        fh.write(w.tobytes())


def read_weights(path: str) -> dict:
    """The contents of an .abxw file: w, norm {mean, std}, meta, grid,
    family_dx, u0_scale, rho_scale."""
    import json
    import struct
    with open(path, "rb") as fh:
        data = fh.read()
    hs = struct.calcsize(_ABXW_HEAD)
    f = struct.unpack(_ABXW_HEAD, data[:hs])
    if f[0] != _ABXW_MAGIC:
        raise ValueError(f"{path}: not a surrogate weight file")
    n, d, h, w_, in_ch, head_ch = f[1:7]
    widths, meta_len = tuple(f[7:11]), f[11]
    if (in_ch, head_ch, widths) != (U.IN_CH, U.HEAD_CH, U.WIDTHS) or n != U.n_params():
        raise ValueError(f"{path}: network shape {in_ch}/{head_ch}/{widths}/{n} differs from this build's")
    dd = f[13:]
    meta = json.loads(data[hs:hs + meta_len].decode())
    off = hs + meta_len + (-(hs + meta_len) % 16)
    w = np.frombuffer(data, np.float32, n, off).copy()
    return dict(w=w, norm=dict(mean=np.array(dd[3:7], np.float32), std=np.array(dd[7:11], np.float32)), meta=meta,
                grid=(d, h, w_), family_dx=dd[0], u0_scale=dd[1], rho_scale=dd[2])


def save(path: str, net: U.UNet, norm: dict, meta: dict) -> None:
    """Training checkpoint: weights (FP32), Adam state and target
    standardisation as one .npz (write_weights makes the inference file)."""
    import json
    st = net.state()
    np.savez_compressed(path, w=st["w"], m=st["m"], v=st["v"], step=st["step"], norm_mean=norm["mean"],
                        norm_std=norm["std"], meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load(path: str, net: U.UNet) -> tuple:
    """Loads weights into net from an .abxw inference file or an .npz
    checkpoint; returns (norm, meta)."""
    import json
    if path.endswith(".abxw"):
        r = read_weights(path)
        net.load_state({"w": r["w"]})
        return r["norm"], r["meta"]
    z = np.load(path)      # plain arrays only (allow_pickle stays off)
    if "m" in z:
        net.load_state({"w": z["w"], "m": z["m"], "v": z["v"], "step": z["step"]})
    else:                  # weights-only file
        net.load_state({"w": z["w"]})
    return dict(mean=z["norm_mean"], std=z["norm_std"]), json.loads(bytes(z["meta"]).decode())


# ---------------------------------------------------------------------------
STUDY_SUBDIV = 57        # quads split 57 x 57: about 200 k triangles per box


def warm_start_study(dev: abx.Device, net: U.UNet, norm: dict, boxes: List[Box], dx: float = 2.0,
                     window: int = 1000, max_steps: int = 60000, fixed_steps: int = 6000, fixed_avg: int = 2000,
                     fp64_every: int = 4, subdiv: int = STUDY_SUBDIV, warm_steps: int = 2000, warm_avg: int = 2000,
                     ideal: Optional[Tuple[float, float]] = None, ideal_alt: bool = False,
                     log: Optional[Callable[[str], None]] = None) -> dict:
    """Cold against surrogate-warm time to steady for held-out boxes.

    Each box is written as a binary STL of about 200 k triangles, parsed and
    prepared on the GPU at dx.  Runs per box:
      fixed       from rest, the reference's fixed-step protocol (fixed_steps,
                  averaging the last fixed_avg): the result the others are
                  measured against;
      cold        from rest to the steady criterion (steady.solve_to_steady);
      warm        from the surrogate's equilibrium to the same criterion;
      warm_fixed  from the surrogate's equilibrium, the fixed protocol
                  shortened to warm_steps averaging the last warm_avg (a warm
                  start needs no transient to discard);
      cold_short  the same shortened protocol from rest (what the warm start
                  is worth at that length);
      ideal       (ideal = (a_rho, a_u)) to the criterion from the
                  equilibrium of fixed + a (prediction - fixed), a_rho for rho
                  and a_u for u: a = 0 takes the fixed run's own average, a
                  perfect prediction of the mean flow, 0 < a < 1 a prediction
                  that far off, a = 1 the network's; with ideal_alt the
                  "fixed" average is replaced by an independent one of the
                  same flow (steps fixed_steps .. fixed_steps + fixed_avg of a
                  longer run from rest), the best a predictor of the mean
                  flow could seed;
      fp64        (every fp64_every-th box) the fixed protocol in FP64.
    Velocity-field errors are relative L2 norms over the fluid cells."""
    from . import steady
    rows = []
    for bi, box in enumerate(boxes):
        spec = copy.deepcopy(box.spec)
        spec.subdiv = subdiv
        t0 = time.perf_counter()
        stl = G.stl_binary_bytes(G.airbox_triangles(spec))
        t1 = time.perf_counter()
        tris = abx.parse_stl(stl)
        g = abx.Geometry(dev, tris, G.prepare_params(spec, dx), max_cells=0)
        t2 = time.perf_counter()
        p = dict(K=box.K, aniso=box.aniso, u=U0, tau=TAU)
        try:
            try:
                fixed = steady.fixed_run(dev, g, p, fixed_steps, fixed_avg)
                cold = steady.solve_to_steady(dev, g, p, window=window, max_steps=max_steps, keep_acc=True)
                timing = {}

                def init(sim, timing=timing):
                    t = time.perf_counter()
                    warm_start_init(net, norm, dev, box, g, sim, tris)
                    timing["predict_s"] = time.perf_counter() - t
                    timing["rho_u"] = sim.read_rho_u()

                warm = steady.solve_to_steady(dev, g, p, window=window, max_steps=max_steps, init_fn=init,
                                              keep_acc=True)
                rho_u = timing["rho_u"]
                warm_fixed = steady.fixed_run(dev, g, p, warm_steps, warm_avg,
                                              init_fn=lambda sim: warm_start_init(net, norm, dev, box, g, sim, tris))
                cold_short = steady.fixed_run(dev, g, p, warm_steps, warm_avg)
                best = None
                if ideal is not None:
                    nf = fixed.acc.size // 4
                    src = steady.fixed_run(dev, g, p, fixed_steps + fixed_avg, fixed_avg) if ideal_alt else fixed
                    fru = np.concatenate([src.acc[3 * nf:], src.acc[:3 * nf]]).astype(np.float64)
                    a = np.repeat(np.array(ideal, np.float64), [nf, 3 * nf])
                    ru = fru + a * (np.asarray(timing["rho_u"], np.float64) - fru)
                    best = steady.solve_to_steady(dev, g, p, window=window, max_steps=max_steps, keep_acc=True,
                                                  init_rho_u=ru.astype(np.float32))
                f64 = steady.fixed_run(dev, g, p, fixed_steps, fixed_avg, precision="f64") \
                    if fp64_every and bi % fp64_every == 0 else None
            except abx.AbxError as e:       # e.g. the reference's instability error
                rows.append(dict(seed=box.seed, error=str(e), grid=[g.nx, g.ny, g.nz]))
                if log:
                    log(f"box {bi + 1}/{len(boxes)} seed {box.seed}: failed: {e}")
                continue
        finally:
            g.free()

        def uerr(a, b):
            a = np.asarray(a, np.float64).reshape(4, -1)[:3]
            b = np.asarray(b, np.float64).reshape(4, -1)[:3]
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

        def compare(r, ref):
            dg = abs(r.metrics["gammaCore"] - ref.metrics["gammaCore"])
            dz = abs(r.metrics["zeta"] - ref.metrics["zeta"]) / abs(ref.metrics["zeta"])
            return dict(gammaCore_abs=dg, zeta_rel=dz, within=bool(dg <= 0.005 and dz <= 0.02),
                        u_err=uerr(r.acc, ref.acc))

        def run(r):
            return dict(converged=r.converged, steps=r.steps, wall_s=r.wall_s,
                        gammaCore=r.metrics["gammaCore"], zeta=r.metrics["zeta"], windows=r.windows)

        NF = fixed.acc.size // 4
        init = np.concatenate([rho_u[NF:], rho_u[:NF]])     # (ux, uy, uz, rho), the order of acc
        row = dict(seed=box.seed, ntri=int(tris.size // 9), grid=[g.nx, g.ny, g.nz], NF=int(NF), K=box.K,
                   aniso=box.aniso, stl_s=t1 - t0, prepare_s=t2 - t1, predict_s=timing["predict_s"],
                   fixed=run(fixed), cold=run(cold), warm=run(warm),
                   cold_vs_fixed=compare(cold, fixed), warm_vs_fixed=compare(warm, fixed),
                   warm_fixed=run(warm_fixed), warm_fixed_vs_fixed=compare(warm_fixed, fixed),
                   cold_short_vs_fixed=compare(cold_short, fixed),
                   prediction_u_err_vs_fixed=uerr(init, fixed.acc))
        if best is not None:
            row["ideal"] = run(best)
            row["ideal_vs_fixed"] = compare(best, fixed)
        if f64 is not None:
            row["fp64"] = run(f64)
            row["fixed_vs_fp64"] = compare(fixed, f64)
            row["warm_vs_fp64"] = compare(warm, f64)
        rows.append(row)
        if log:
            log(f"box {bi + 1}/{len(boxes)} seed {box.seed}: {row['ntri']} tris, grid {row['grid']}, "
                f"cold {cold.steps} steps {cold.wall_s:.2f} s, warm {warm.steps} steps {warm.wall_s:.2f} s "
                f"(+{row['predict_s']:.2f} s predict); vs fixed: cold dgamma {row['cold_vs_fixed']['gammaCore_abs']:.4f} "
                f"dzeta {row['cold_vs_fixed']['zeta_rel']:.3f}, warm dgamma {row['warm_vs_fixed']['gammaCore_abs']:.4f} "
                f"dzeta {row['warm_vs_fixed']['zeta_rel']:.3f}; prediction u err {row['prediction_u_err_vs_fixed']:.3f}")
    return dict(rows=rows, summary=summarize(rows), dx=dx, window=window, fixed_steps=fixed_steps,
                fixed_avg=fixed_avg, warm_steps=warm_steps, warm_avg=warm_avg)


def summarize(rows: List[dict]) -> dict:
    ok = [r for r in rows if "error" not in r and r["cold"]["converged"] and r["warm"]["converged"]]
    out = dict(n=len(rows), n_failed=sum("error" in r for r in rows), n_converged=len(ok))
    if not ok:
        return out
    cs = np.array([r["cold"]["steps"] for r in ok], float)
    ws = np.array([r["warm"]["steps"] for r in ok], float)
    cw = np.array([r["cold"]["wall_s"] for r in ok])
    ww = np.array([r["warm"]["wall_s"] + r["predict_s"] for r in ok])
    col = lambda key, f: np.array([r[key][f] for r in ok], float)
    # the speedup over the boxes where both answers agree with the fixed result
    both = np.array([r["warm_vs_fixed"]["within"] and r["cold_vs_fixed"]["within"] for r in ok])
    out.update(n_both_within=int(both.sum()),
               wall_speedup_both_within=float(cw[both].sum() / ww[both].sum()) if both.any() else None)
    out.update(steps_cold_median=float(np.median(cs)), steps_warm_median=float(np.median(ws)),
               step_speedup_median=float(np.median(cs / ws)), step_speedup_total=float(cs.sum() / ws.sum()),
               wall_speedup_total_incl_predict=float(cw.sum() / ww.sum()),
               warm_within_tol_of_fixed=float(np.mean(col("warm_vs_fixed", "within"))),
               cold_within_tol_of_fixed=float(np.mean(col("cold_vs_fixed", "within"))),
               warm_dgamma_vs_fixed_median=float(np.median(col("warm_vs_fixed", "gammaCore_abs"))),
               cold_dgamma_vs_fixed_median=float(np.median(col("cold_vs_fixed", "gammaCore_abs"))),
               warm_dzeta_vs_fixed_median=float(np.median(col("warm_vs_fixed", "zeta_rel"))),
               cold_dzeta_vs_fixed_median=float(np.median(col("cold_vs_fixed", "zeta_rel"))),
               warm_u_err_vs_fixed_median=float(np.median(col("warm_vs_fixed", "u_err"))),
               cold_u_err_vs_fixed_median=float(np.median(col("cold_vs_fixed", "u_err"))),
               prediction_u_err_vs_fixed_median=float(np.median([r["prediction_u_err_vs_fixed"] for r in ok])))
    wf = [r for r in ok if "warm_fixed" in r]
    if wf:
        fw = np.array([r["fixed"]["wall_s"] for r in wf])
        ww = np.array([r["warm_fixed"]["wall_s"] + r["predict_s"] for r in wf])
        out.update(warm_fixed_step_ratio=float(wf[0]["fixed"]["steps"] / wf[0]["warm_fixed"]["steps"]),
                   warm_fixed_wall_speedup_total_incl_predict=float(fw.sum() / ww.sum()),
                   warm_fixed_within_tol_of_fixed=float(np.mean([r["warm_fixed_vs_fixed"]["within"] for r in wf])),
                   cold_short_within_tol_of_fixed=float(np.mean([r["cold_short_vs_fixed"]["within"] for r in wf])),
                   warm_fixed_dgamma_vs_fixed_median=float(np.median([r["warm_fixed_vs_fixed"]["gammaCore_abs"]
                                                                      for r in wf])),
                   warm_fixed_dzeta_vs_fixed_median=float(np.median([r["warm_fixed_vs_fixed"]["zeta_rel"] for r in wf])),
                   cold_short_dzeta_vs_fixed_median=float(np.median([r["cold_short_vs_fixed"]["zeta_rel"] for r in wf])),
                   warm_fixed_u_err_vs_fixed_median=float(np.median([r["warm_fixed_vs_fixed"]["u_err"] for r in wf])))
    iw = [r for r in ok if "ideal" in r]
    if iw:
        ic = np.array([r["cold"]["wall_s"] for r in iw])
        ii = np.array([r["ideal"]["wall_s"] for r in iw])
        out.update(ideal_steps_median=float(np.median([r["ideal"]["steps"] for r in iw])),
                   ideal_converged=float(np.mean([r["ideal"]["converged"] for r in iw])),
                   ideal_step_speedup_total=float(sum(r["cold"]["steps"] for r in iw) /
                                                  sum(r["ideal"]["steps"] for r in iw)),
                   ideal_wall_speedup_total=float(ic.sum() / ii.sum()),
                   ideal_within_tol_of_fixed=float(np.mean([r["ideal_vs_fixed"]["within"] for r in iw])),
                   ideal_dzeta_vs_fixed_median=float(np.median([r["ideal_vs_fixed"]["zeta_rel"] for r in iw])))
    f = [r for r in ok if "fp64" in r]
    if f:
        out.update(fp64_boxes=len(f),
                   fixed_vs_fp64_dgamma_max=float(max(r["fixed_vs_fp64"]["gammaCore_abs"] for r in f)),
                   fixed_vs_fp64_dzeta_max=float(max(r["fixed_vs_fp64"]["zeta_rel"] for r in f)),
                   fixed_vs_fp64_u_err_median=float(np.median([r["fixed_vs_fp64"]["u_err"] for r in f])),
                   warm_vs_fp64_dgamma_median=float(np.median([r["warm_vs_fp64"]["gammaCore_abs"] for r in f])),
                   warm_vs_fp64_dzeta_median=float(np.median([r["warm_vs_fp64"]["zeta_rel"] for r in f])))
    return out


def main(argv=None) -> int:
    """python -m airboxcalc_amd.surrogate train|study ...: the surrogate
    dataset + training run, and the warm-start study."""
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser(prog="python -m airboxcalc_amd.surrogate")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train", help="generate the LBM dataset on the GPU and train the U-Net")
    tr.add_argument("--boxes", type=int, default=2000)
    tr.add_argument("--first-seed", type=int, default=0)
    tr.add_argument("--steps", type=int, default=4000)
    tr.add_argument("--avg", type=int, default=1000)
    tr.add_argument("--epochs", type=int, default=40)
    tr.add_argument("--batch", type=int, default=8)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--max-train-seconds", type=float, default=600)
    tr.add_argument("--target-dx", type=float, default=None,
                    help="solve the boxes at this cell size and restrict the fields to the 96x48x64 inputs")
    tr.add_argument("--init-weights", help="continue training these weights (their standardisation is kept)")
    tr.add_argument("--cosine", action="store_true", help="cosine learning-rate decay to lr / 10")
    tr.add_argument("--keep-best", action="store_true", help="keep the weights of the best held-out epoch")
    tr.add_argument("--out", required=True, help="training checkpoint .npz; the inference file is written "
                                                 "beside it with the suffix .abxw")
    tr.add_argument("--report", help="JSON report")
    st = sub.add_parser("study", help="cold vs warm time to steady on held-out boxes")
    st.add_argument("--weights", default=WEIGHTS, help="weights .abxw or .npz (default: the trained 96x48x64 "
                                                       "surrogate)")
    st.add_argument("--boxes", type=int, default=128)
    st.add_argument("--first-seed", type=int, default=1_000_000)
    st.add_argument("--dx", type=float, default=2.0)
    st.add_argument("--window", type=int, default=1000)
    st.add_argument("--fp64-every", type=int, default=4)
    st.add_argument("--subdiv", type=int, default=STUDY_SUBDIV)
    st.add_argument("--warm-steps", type=int, default=2000, help="steps of the shortened fixed protocol")
    st.add_argument("--warm-avg", type=int, default=2000, help="steps it averages")
    st.add_argument("--ideal", type=float, nargs=2, default=None, metavar=("A_RHO", "A_U"),
                    help="also start from fixed + A (prediction - fixed), A_RHO for rho and A_U for u; A = 0: the "
                         "fixed run's own average, a perfect mean-flow prediction")
    st.add_argument("--ideal-alt", action="store_true",
                    help="--ideal blends toward an independent average of the same flow (the fixed protocol's "
                         "window moved past its end) instead of the fixed result itself")
    st.add_argument("--report", help="JSON report")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    dev = abx.device(0)
    if a.cmd == "train":
        boxes = family(range(a.first_seed, a.first_seed + a.boxes))
        net = U.UNet(seed=0)
        init_norm, init_meta = load(a.init_weights, net) if a.init_weights else (None, None)
        t0 = time.perf_counter()
        data = make_dataset(dev, boxes, a.steps, a.avg, log=log, target_dx=a.target_dx, norm=init_norm)
        t_data = time.perf_counter() - t0
        res = train(net, data, a.epochs, a.batch, a.lr, log=log, max_seconds=a.max_train_seconds,
                    schedule=a.cosine, keep_best=a.keep_best)
        meta = dict(boxes=a.boxes, first_seed=a.first_seed, steps=a.steps, avg=a.avg, dx=FAMILY_DX, grid=GRID,
                    target_dx=a.target_dx, epochs_run=len(res["history"]), batch=a.batch, lr=a.lr,
                    cosine=a.cosine, keep_best=a.keep_best, best_epoch=res.get("best_epoch"))
        if init_meta is not None:
            meta["continued_from"] = init_meta
        save(a.out, net, data["norm"], meta)
        write_weights(os.path.splitext(a.out)[0] + ".abxw", net.state()["w"], data["norm"], meta)
        rep = dict(meta=meta, dataset_s=t_data, dataset_stats=data["stats"], norm={k: v.tolist() for k, v in
                                                                                   data["norm"].items()},
                   train=res, params=U.n_params(), device=dev.name, torch=torch.__version__)
        log(json.dumps({k: rep[k] for k in ("dataset_s", "params")}))
    else:
        net = U.UNet(seed=0)
        norm, meta = load(a.weights, net)
        boxes = family(range(a.first_seed, a.first_seed + a.boxes))
        rep = warm_start_study(dev, net, norm, boxes, dx=a.dx, window=a.window, fp64_every=a.fp64_every,
                               subdiv=a.subdiv, warm_steps=a.warm_steps, warm_avg=a.warm_avg,
                               ideal=tuple(a.ideal) if a.ideal else None, ideal_alt=a.ideal_alt,
                               log=log)
        rep["weights_meta"] = meta
        rep["device"] = dev.name
        log(json.dumps(rep["summary"]))
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1, default=float)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
