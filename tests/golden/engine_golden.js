// Golden-vector writer: runs the reference compute core (src/engine.js) under
// Node on one case description and records its outputs.
//
// Usage: node engine_golden.js <path/to/engine.js> <case.json> <out_dir>
//
// The reference is loaded from its own file at run time (never copied); its
// non-exported helpers (voxelize, probeCPU, progressInfo) are reached by
// evaluating the file's text with one extra line that hands them out.
// Outputs: <out_dir>/result.json (scalars, SHA-256 of every typed array, and
// exact float64 values as hex), and raw little-endian array files
// <out_dir>/<name>.bin for the arrays make_golden.py keeps.
"use strict";
const fs = require("fs");
const crypto = require("crypto");
const path = require("path");

const [enginePath, casePath, outDir] = process.argv.slice(2);
const src = fs.readFileSync(enginePath, "utf8");
const mod = { exports: {} };
// eslint-disable-next-line no-new-func
new Function("module", "TextDecoder",
  src + "\n;module.internals = { voxelize, probeCPU, progressInfo };")(mod, TextDecoder);
const E = Object.assign({}, mod.exports, mod.internals);
const C = JSON.parse(fs.readFileSync(casePath, "utf8"));
fs.mkdirSync(outDir, { recursive: true });

const R = { case: C.name };
const hex64 = (x) => { const b = Buffer.alloc(8); b.writeDoubleLE(x, 0); return b.toString("hex"); };
const exact = (x) => ({ v: x, hex: hex64(x) });
const bytesOf = (ta) => Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength);
const sha = (ta) => crypto.createHash("sha256").update(bytesOf(ta)).digest("hex");
const dump = (name, ta) => { fs.writeFileSync(path.join(outDir, name + ".bin"), bytesOf(ta)); };
const keep = (name, ta) => { R.sha[name] = sha(ta); R.len[name] = ta.length; if (C.keep_arrays) dump(name, ta); };
R.sha = {}; R.len = {};

function tryCall(label, fn) {
  try { return fn(); } catch (err) { R.errors = R.errors || {}; R.errors[label] = String(err.message || err); return null; }
}

const buf = fs.readFileSync(C.stl);
const ab = buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength);
const tris = tryCall("parseSTL", () => E.parseSTL(ab));
if (!tris) { fs.writeFileSync(path.join(outDir, "result.json"), JSON.stringify(R)); process.exit(0); }
keep("tris", tris);

const P = Object.assign({ tris }, C.prepare);
// V8's own cos/sin of the yaw, exactly as makeFrame computes them (engine.js:36)
const yawRad = (P.inletYaw || 0) * Math.PI / 180;
R.frame = { cy: exact(Math.cos(yawRad)), sy: exact(Math.sin(yawRad)) };

// Local-frame triangles and the fine grid, the way prepare derives them
// (engine.js:112-126), so the fine voxel mask itself can be recorded.
if (C.voxel) {
  const fr = E.makeFrame(P), ntri = tris.length / 9, T = new Float32Array(tris.length);
  const mn = [Infinity, Infinity, Infinity], mx = [-Infinity, -Infinity, -Infinity];
  for (let v = 0; v < ntri * 3; v++) {
    const L = fr.toLocal([tris[v * 3], tris[v * 3 + 1], tris[v * 3 + 2]]);
    for (let a = 0; a < 3; a++) { T[v * 3 + a] = L[a]; if (L[a] < mn[a]) mn[a] = L[a]; if (L[a] > mx[a]) mx[a] = L[a]; }
  }
  const h = P.dx, hf = h / 3, out = P.outlet && P.outlet.enabled ? P.outlet : null;
  const origin = [mn[0] - h, (out ? Math.min(mn[1], -out.height) : mn[1]) - h, mn[2] - 10 * h];
  const nx = Math.ceil((mx[0] - origin[0] + h) / h), ny = Math.ceil((mx[1] - origin[1] + h) / h), nz = Math.ceil((mx[2] - origin[2] + h) / h);
  R.voxel = { origin: origin.map(exact), mn: mn.map(exact), mx: mx.map(exact), nxf: 3 * nx, nyf: 3 * ny, nzf: 3 * nz, hf: exact(hf) };
  keep("T", T);
  if (3 * nx * 3 * ny * 3 * nz <= (C.max_fine || 8e7)) {
    const fine = E.voxelize(T, origin, 3 * nx, 3 * ny, 3 * nz, hf);
    keep("fine", fine);
  }
}

const g = tryCall("prepare", () => E.prepare(P));
if (g) {
  R.prepare = {
    nx: g.nx, ny: g.ny, nz: g.nz, N: g.N, nf: g.nf, top: g.top, jm: g.jm, h: g.h,
    mn: g.mn.map(exact), mx: g.mx.map(exact),
    summary: Object.assign({}, g.summary, { openPts: undefined, filterPts: undefined, outletPts: undefined }),
  };
  for (const k of ["volume_l", "inlet_cm2", "filter_cm2", "outlet_cm2", "outlet_cm2_real"]) R.prepare.summary[k] = exact(g.summary[k]);
  for (const k of ["solid", "fluid", "opening", "outlet", "xc", "dc", "zc"]) keep(k, g[k]);
  for (const k of ["openPts", "filterPts", "outletPts"]) keep(k, Float64Array.from(g.summary[k]));
}

if (g && C.run) {
  const p = Object.assign({}, C.run);
  const S = E.setupSim(g, p);
  R.setup = { NF: S.NF, Ain: S.Ain, Af: S.Af, sig: exact(S.sig), cs2s: exact(S.cs2s), nM: S.layerM.length, nT: S.layerT.length };
  keep("cid", S.cid); keep("nb", S.nb); keep("flag", S.flag);
  keep("layerM", Int32Array.from(S.layerM)); keep("layerT", Int32Array.from(S.layerT));

  // cpuStep trajectories from f0 (engine.js:597-611 uses 12 steps and this ns pattern)
  const NF = S.NF;
  const nsST = new Float32Array(NF);
  for (let a = 0; a < NF; a += 17) if (S.flag[a] === 0) nsST[a] = 1;
  const variants = [["plain", null], ["ns", nsST]];
  R.traj = {};
  for (const [vn, ns] of variants) {
    let f = S.f0.slice(), f2 = new Float32Array(19 * NF);
    const acc = new Float32Array(4 * NF);
    const rec = {};
    for (let s = 1; s <= (C.traj_steps || 12); s++) {
      E.cpuStep(S, f, f2, s > 6 ? acc : null, ns);
      const t = f; f = f2; f2 = t;
      if (s === 1 || s === 2 || s === (C.traj_steps || 12)) {
        keep(`f_${vn}_${s}`, f);
      }
    }
    keep(`acc_${vn}`, acc);
    const list = S.layerM.concat(S.layerT);
    const vy = E.probeCPU(S, f, list);
    keep(`vy_${vn}`, vy);
    const info = E.progressInfo(S, vy, 12, 12, Date.now(), []);
    rec.gamma = exact(info.gamma); rec.massRatio = exact(info.massRatio);
    R.traj[vn] = rec;
  }
  if (C.keep_arrays) dump("ns_selftest", nsST);

  if (C.simulate) {
    const runs = [["sim", null]];
    if (C.simulate_ns) runs.push(["simns", nsST]);
    R.sim = {};
    for (const [rn, ns] of runs) {
      const t0 = Date.now();
      const res = E.simulate(g, p, null, ns);
      const ms = Date.now() - t0;
      const M = {};
      for (const k of Object.keys(res.metrics)) M[k] = exact(res.metrics[k]);
      R.sim[rn] = { metrics: M, hist: res.hist.map((r) => r.map(exact)), ms, NF,
        vol: { kB: res.vol.kB, x0: exact(res.vol.x0), d0: exact(res.vol.d0), z0: exact(res.vol.z0), zIn: exact(res.vol.zIn) },
        lon_x: exact(res.lon.x), planInlet_depth: exact(res.planInlet.depth), planLow_depth: exact(res.planLow.depth),
        nlines: res.lines.off.length - 1, npts: res.lines.spd.length };
      keep(`${rn}_fmap`, res.fmap);
      for (const pl of ["planInlet", "planLow", "lon"]) for (const k of ["speed", "ua", "ub"]) keep(`${rn}_${pl}_${k}`, res[pl][k]);
      for (const k of ["spd", "pst", "loss"]) keep(`${rn}_vol_${k}`, res.vol[k]);
      for (const k of ["pts", "spd", "off"]) keep(`${rn}_lines_${k}`, res.lines[k]);
    }
  }
}
fs.writeFileSync(path.join(outDir, "result.json"), JSON.stringify(R, null, 1));
