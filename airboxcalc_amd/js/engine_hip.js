/* MI355X (gfx950) backend of the AirboxCalc compute core, with the API of the
   reference's src/engine.js.

   Drop-in use beside the reference's backends:
     const E = require("engine.js");            // CPU backend (reference)
     const H = require("airboxcalc_amd/js/engine_hip.js");
     const dev = H.hipDevice();                   // like gpuDevice() (engine.js:479)
     const g = H.prepare(Object.assign({ tris }, params), dev);   // like prepare (engine.js:111)
     const res = await H.simulateHIP(dev, g, runParams, onProgress, null);  // like simulateGPU (engine.js:613)

   The objects are the reference's: g carries {fr, h, nx, ny, nz, N, xc, dc,
   zc, solid, fluid, opening, outlet, top, jm, nf, mn, mx, pleat, summary}
   (engine.js:226-233); progress objects {step, steps, gamma, massRatio,
   elapsed, eta, mlups, device} (engine.js:338); results {metrics, hist, fmap,
   nx, nz, ny, xc, zc, dc, pleat, planInlet, planLow, lon, frame, vol, lines}
   (engine.js:681-690) plus device.  The work runs on the GPU through the
   N-API addon over libairbox_hip.so; there is no CPU fallback in this module
   (a caller wanting one keeps the reference's simulate).  Errors are thrown
   with the reference's messages. */
"use strict";
const path = require("path");

const A = require(path.join(__dirname, "..", "airbox_hip.node"));

const PRECISION = { f32: 0, f32_f64: 1, f64: 2 };
const CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1];

/* hipDevice(ordinal) -> {device, name, maxBinding} | null: the GPU handle, as
   gpuDevice() returns the WebGPU one (engine.js:479-491). */
function hipDevice(ordinal) {
  try {
    const d = A.deviceOpen(ordinal || 0);
    return { device: d.dev, name: d.name, maxBinding: Number.MAX_SAFE_INTEGER };
  } catch (err) {
    return null;
  }
}

let defaultDev = null;
function deviceOrDefault(dev) {
  if (dev) return dev;
  if (!defaultDev) {
    defaultDev = hipDevice(0);
    if (!defaultDev) throw new Error("Keine MI355X-Grafikkarte (gfx950) verfügbar.");
  }
  return defaultDev;
}

/* parseSTL(buf) -> Float32Array (engine.js:15-28), parsed natively. */
function parseSTL(buf) {
  return A.parseSTL(buf);
}

/* makeFrame(p) (engine.js:31-43): local frame X (free axis), D (depth from the
   seal plane), Z (from the inlet side), turned by the inlet yaw about D. */
function makeFrame(p) {
  const dAx = p.sealAxis, iAx = p.inletAxis;
  if (dAx === iAx) throw new Error("Einlassseite und Filterebene liegen auf derselben Achse.");
  const oAx = 3 - dAx - iAx, sd = p.sealSign, si = p.inletSign, seal = p.seal;
  const yawDeg = p.inletYaw || 0;
  const rad = yawDeg * Math.PI / 180, c = Math.cos(rad), s = Math.sin(rad);
  const toModel = (X, D, Z) => {
    const q = [0, 0, 0];
    q[oAx] = c * X - s * Z; q[dAx] = seal - sd * D; q[iAx] = -si * (s * X + c * Z);
    return q;
  };
  return {
    dAx, iAx, oAx, sd, si, seal, yaw: yawDeg, cy: c, sy: s,
    toLocal: (q) => {
      const a = q[oAx], b = -si * q[iAx];
      return [c * a + s * b, -sd * (q[dAx] - seal), -s * a + c * b];
    },
    toModel,
    dirToModel: (X, D, Z) => {
      const q = [0, 0, 0];
      q[oAx] = c * X - s * Z; q[dAx] = -sd * D; q[iAx] = -si * (s * X + c * Z);
      return q;
    },
  };
}

/* prepare(p, dev) -> g (engine.js:110-234).  p as the reference's prepare
   object plus an optional maxCells (the reference's cap 4.5e6 by default;
   0 lifts it for this backend). */
function prepare(p, dev) {
  const d = deviceOrDefault(dev);
  const fr = makeFrame(p);
  const r = A.prepare(d.device, p.tris, Object.assign({}, p, {
    cy: fr.cy, sy: fr.sy, maxCells: p.maxCells === undefined ? 4.5e6 : p.maxCells,
  }));
  const I = r.info, h = p.dx;
  const out = p.outlet && p.outlet.enabled ? p.outlet : null;
  return {
    // the inputs, kept for a surrogate warm start at another cell size
    source: { tris: p.tris, params: Object.assign({}, p, { tris: undefined, cy: fr.cy, sy: fr.sy }) },
    fr, h, nx: I.nx, ny: I.ny, nz: I.nz, N: I.N, xc: r.xc, dc: r.dc, zc: r.zc,
    solid: r.solid, fluid: r.fluid, opening: r.opening, outlet: r.outlet,
    top: I.top, jm: I.jm, nf: I.nf, mn: I.mn, mx: I.mx, pleat: p.pleat,
    summary: {
      grid: [I.nx, I.ny, I.nz], cells: I.nf, volume_l: I.nf * h ** 3 / 1e6, inlet_cm2: I.nOpen * h * h / 100,
      filter_cm2: I.Af * h * h / 100, outlet_cm2: I.leak * h * h / 100, nHoles: I.nHoles,
      openPts: Array.from(r.openPts), filterPts: Array.from(r.filterPts), outletPts: Array.from(r.outletPts),
      outlet_cm2_real: out ? I.nOut * h * h / 100 : 0, pleatOk: I.jm > I.top,
    },
    native: r.geom, device: d,
  };
}

/* Cell flags in the reference's compact order (engine.js:252-254):
   1 inlet, 2 outlet, 3 porous, 0 bulk. */
function cellFlags(g) {
  const { ny, nz, N, fluid, opening, outlet, top, dc, pleat } = g;
  const flags = new Uint8Array(g.nf);
  let a = 0;
  for (let c = 0; c < N; c++) {
    if (!fluid[c]) continue;
    const k = c % nz, ij = (c / nz) | 0, j = ij % ny;
    if (k === 0 && opening[ij]) flags[a] = 1;
    else if (outlet[c]) flags[a] = 2;
    else if (dc[j] > dc[top] && dc[j] <= pleat) flags[a] = 3;
    a++;
  }
  return flags;
}

/* gpuSelfTestHIP(dev, g, p) -> max |df|: the reference's GPU/CPU gate
   (engine.js:597-611): two variants (no obstacles; ns = 1 on every 17th bulk
   cell), 12 steps from f0, every 7th element compared.  The checked side is
   the production kernel: FP32 arithmetic, AA-pattern propagation in one
   lattice.  The reference side shares neither: the reference's double
   arithmetic in cpuStep's order (ABX_F32_F64, bit-identical to cpuStep) with
   two lattices and pull streaming. */
function gpuSelfTestHIP(dev, g, p) {
  const d = deviceOrDefault(dev);
  const NF = g.nf, ns = new Float32Array(2 * NF), flags = cellFlags(g);
  for (let a = 0; a < NF; a += 17) if (flags[a] === 0) ns[NF + a] = 1;
  const run = (prec) => {
    const s = A.simCreate(d.device, g.native, p, prec, 2, ns);
    if (prec === PRECISION.f32_f64) A.simSetLaunch(s.sim, 2);
    A.simSteps(s.sim, 12, false);
    const f = [A.simReadF(s.sim, 0), A.simReadF(s.sim, 1)];
    A.destroy(s.sim);
    return f;
  };
  const fg = run(PRECISION.f32), fc = run(PRECISION.f32_f64);
  let md = 0;
  for (let b = 0; b < 2; b++) for (let x = 0; x < fc[b].length; x += 7) md = Math.max(md, Math.abs(fc[b][x] - fg[b][x]));
  return md;
}

const plane = (buf, L, n1, n2, key, value) => {
  const o = { speed: buf.slice(0, L), ua: buf.slice(L, 2 * L), ub: buf.slice(2 * L, 3 * L), n1, n2 };
  o[key] = value;
  return o;
};

/* The trained surrogate networks shipped with this backend (no reference
   counterpart): "family" predicts at its own cell size of 4.8 mm, "dx2" was
   trained on 2 mm solutions (use it for solves at 2 mm). */
const SURROGATE_WEIGHTS = {
  family: path.join(__dirname, "..", "data", "surrogate_unet_96x48x64.abxw"),
  dx2: path.join(__dirname, "..", "data", "surrogate_unet_96x48x64_dx2targets_1500.abxw"),
};

/* loadSurrogate(dev, file) -> {net, dims, familyDx}: the network on the GPU
   (abx_surrogate_load), for simulateHIP's warmStart option. */
function loadSurrogate(dev, file) {
  const d = deviceOrDefault(dev);
  return A.surrogateLoad(d.device, file || SURROGATE_WEIGHTS.family);
}

/* simulateHIP(dev, g, p, onProgress, ns, opts) -> result: the reference's
   simulateGPU (engine.js:613-630) on the gfx950 kernel.  p = {K, aniso, u,
   tau, steps, avg, every}; opts (all optional):
     precision  "f32" (default), "f32_f64" (the CPU backend's arithmetic) or
                "f64";
     warmStart  a loadSurrogate() network: the lattice starts at the
                equilibrium of its prediction instead of at rest
                (engine.js:260-261);
     criterion  {window, tolGamma, tolZeta, maxSteps}: instead of p.steps /
                p.avg, run windows of `window` averaged steps until gamma-core
                moves by less than tolGamma and zeta by at most tolZeta
                (relative) between two windows (abx_sim_solve_to_steady).  The
                result's metrics are then the mean of those two windows, its
                fields the last window's average, hist one row per window, and
                result.steady = {converged, steps, windows}; a progress object
                per window adds gammaCore and zeta. */
async function simulateHIP(dev, g, p, onProgress, ns, opts) {
  const d = deviceOrDefault(dev);
  const prec = PRECISION[(opts && opts.precision) || "f32"];
  if (prec === undefined) throw new Error("unknown precision " + opts.precision);
  const c = A.simCreate(d.device, g.native, p, prec, 1, ns || null), sim = c.sim, NF = c.info.NF;
  try {
    if (opts && opts.warmStart) {
      const src = g.source || {};
      A.surrogateWarmStart(opts.warmStart.net, sim, g.native, src.tris || null, src.params || null, p.K, p.aniso);
    }
    if (opts && opts.criterion) return await steadyRun(d, g, p, onProgress, sim, NF, opts.criterion);
    const t0 = Date.now(), hist = [];
    let step = 0, nacc = 0;
    const startAcc = p.steps - p.avg;
    while (step < p.steps) {
      let n = Math.min(p.every, p.steps - step);
      if (step < startAcc && step + n > startAcc) n = startAcc - step;
      const doacc = step >= startAcc;
      const probe = (step + n) % p.every === 0 || step + n === p.steps;
      const pr = await A.simChunk(sim, n, doacc, step + n, probe);
      step += n;
      if (doacc) nacc += n;
      if (probe) {
        const el = (Date.now() - t0) / 1000;
        hist.push([step, pr.gamma, pr.massRatio]);
        const info = { step, steps: p.steps, gamma: pr.gamma, massRatio: pr.massRatio, elapsed: el,
          eta: el / step * (p.steps - step), mlups: NF * step / el / 1e6, device: "GPU" };
        if (onProgress) onProgress(info);
      }
    }
    return resultObject(d, g, sim, Math.max(nacc, 1), hist);
  } finally {
    A.destroy(sim);
  }
}

/* The result object of evaluate (engine.js:681-690) from the handle's
   averaged fields divided by div. */
function resultObject(d, g, sim, div, hist) {
  const E = A.evaluate(sim, g.native, 0, div, true), L = A.streamlines(sim, g.native, 0, div, 220);
  const { nx, ny, nz, xc, dc, zc, fr } = g, Lxz = nx * nz;
  return {
    metrics: E.metrics, hist, fmap: E.fmap, nx, nz, ny, xc: Array.from(xc), zc: Array.from(zc), dc: Array.from(dc),
    pleat: g.pleat,
    planInlet: plane(E.planInlet, Lxz, nx, nz, "depth", dc[E.jc]),
    planLow: plane(E.planLow, Lxz, nx, nz, "depth", dc[E.jlow]),
    lon: plane(E.lon, nz * ny, nz, ny, "x", xc[E.ic]),
    frame: { dAx: fr.dAx, iAx: fr.iAx, oAx: fr.oAx, sd: fr.sd, si: fr.si, seal: fr.seal, yaw: fr.yaw },
    vol: { nx, ny, nz, h: g.h, x0: xc[0], d0: dc[0], z0: zc[0], spd: E.spd, pst: E.pst, loss: E.loss, jm: g.jm,
      top: g.top, pleat: g.pleat, kB: E.kB, zIn: g.mn[2] },
    lines: L, device: "GPU · " + d.name,
  };
}

/* simulateHIP's criterion run: the library's steady solve on a pool thread,
   a progress object per window on this thread. */
async function steadyRun(d, g, p, onProgress, sim, NF, crit) {
  const window = crit.window || 1000, maxSteps = crit.maxSteps || 60000;
  const t0 = Date.now(), hist = [], windows = [];
  let allSeen = null;
  const onWindow = (w) => {
    const el = (Date.now() - t0) / 1000;
    hist.push([w.step, w.gamma, w.massRatio]);
    windows.push([w.step, w.gamma, w.massRatio, w.metrics.gammaCore, w.metrics.zeta]);
    if (onProgress) {
      onProgress({ step: w.step, steps: maxSteps, gamma: w.gamma, massRatio: w.massRatio, elapsed: el,
        eta: el / w.step * (maxSteps - w.step), mlups: NF * w.step / el / 1e6, device: "GPU",
        gammaCore: w.metrics.gammaCore, zeta: w.metrics.zeta });
    }
    if (allSeen && windows.length >= allSeen.n) allSeen.resolve();
  };
  const r = await A.solveToSteady(sim, g.native, { window, maxSteps, tolGamma: crit.tolGamma === undefined ? 0.005 :
    crit.tolGamma, tolZeta: crit.tolZeta === undefined ? 0.02 : crit.tolZeta }, onWindow);
  // the per-window calls may trail the solve's completion
  if (windows.length < r.windows) await new Promise((resolve) => { allSeen = { n: r.windows, resolve }; });
  const res = resultObject(d, g, sim, window, hist);
  res.metrics = r.metrics;
  res.steady = { converged: r.converged, steps: r.steps, windows };
  return res;
}

/* runHIP(g, p, onProgress, opts) -> result: the HIP path of the reference's
   worker runSim (engine.js:765-779): the self-test gate, then the run.  The
   gate passes when max |df| < opts.gateTolerance (default 1e-3, the
   reference's; a caller or a test may tighten it) and otherwise throws the
   reference's "Abweichung GPU/CPU" error. */
async function runHIP(g, p, onProgress, opts) {
  const d = deviceOrDefault(g.device);
  const tol = opts && opts.gateTolerance !== undefined ? opts.gateTolerance : 1e-3;
  const md = gpuSelfTestHIP(d, g, p);
  if (!(md < tol)) throw new Error("Abweichung GPU/CPU " + md.toExponential(1));
  return simulateHIP(d, g, p, onProgress, null, opts);
}

module.exports = { hipDevice, parseSTL, makeFrame, prepare, simulateHIP, gpuSelfTestHIP, runHIP, cellFlags,
  loadSurrogate, SURROGATE_WEIGHTS, native: A };
