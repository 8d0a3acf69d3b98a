// Runs one case through the gfx950 backend's JavaScript API
// (airboxcalc_amd/js/engine_hip.js) and prints what the parity test compares
// with the reference's golden outputs: SHA-256 of every typed array, exact
// float64 scalars, the progress objects.
//
// Usage: node run_hip_case.js <case.json>
//   case: {stl, prepare, run, precision, selftest, nanCell, gateTolerance}
//   nanCell: run with ns = NaN on that fluid cell (cid order) and report the
//            error the run rejects with;
//   gateTolerance: call runHIP with this gate and report its error, if any;
//   warmStart: {weights, criterion}: seed a handle from the surrogate and
//            report its populations' SHA-256, then run simulateHIP with
//            {warmStart, criterion} and report its steady result.
"use strict";
const fs = require("fs");
const crypto = require("crypto");
const path = require("path");
const H = require(path.join(__dirname, "..", "..", "airboxcalc_amd", "js", "engine_hip.js"));

const C = JSON.parse(fs.readFileSync(process.argv[2], "utf8"));
const bytesOf = (ta) => Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength);
const sha = (ta) => crypto.createHash("sha256").update(bytesOf(ta)).digest("hex");
const hex64 = (x) => { const b = Buffer.alloc(8); b.writeDoubleLE(x, 0); return b.toString("hex"); };

(async () => {
  const out = { sha: {}, hex: {} };
  const dev = H.hipDevice(0);
  if (!dev) throw new Error("no gfx950 GPU");
  out.device = dev.name;
  const buf = fs.readFileSync(C.stl);
  const tris = H.parseSTL(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength));
  out.sha.tris = sha(tris);
  let g;
  try {
    g = H.prepare(Object.assign({ tris }, C.prepare), dev);
  } catch (err) {
    out.error = err.message;
    process.stdout.write(JSON.stringify(out));
    return;
  }
  for (const k of ["solid", "fluid", "opening", "outlet", "xc", "dc", "zc"]) out.sha[k] = sha(g[k]);
  for (const k of ["openPts", "filterPts", "outletPts"]) out.sha[k] = sha(Float64Array.from(g.summary[k]));
  out.g = { nx: g.nx, ny: g.ny, nz: g.nz, N: g.N, nf: g.nf, top: g.top, jm: g.jm };
  out.summary = Object.assign({}, g.summary, { openPts: undefined, filterPts: undefined, outletPts: undefined });
  for (const k of ["volume_l", "inlet_cm2", "filter_cm2", "outlet_cm2", "outlet_cm2_real"]) out.hex[k] = hex64(g.summary[k]);
  if (C.selftest) out.selftest = H.gpuSelfTestHIP(dev, g, C.run);
  if (C.nanCell !== undefined) {
    const ns = new Float32Array(g.nf);
    ns[C.nanCell] = NaN;
    const prog = [];
    try {
      await H.simulateHIP(dev, g, C.run, (info) => prog.push(info.step), ns, { precision: C.precision });
      out.simError = null;
    } catch (err) {
      out.simError = err.message;
    }
    out.progressSteps = prog;
    process.stdout.write(JSON.stringify(out));
    return;
  }
  if (C.warmStart) {
    const net = H.loadSurrogate(dev, C.warmStart.weights);
    out.net = { dims: net.dims, familyDx: net.familyDx };
    const s = H.native.simCreate(dev.device, g.native, C.run, 0, 1, null);
    H.native.surrogateWarmStart(net.net, s.sim, g.native, g.source.tris, g.source.params, C.run.K, C.run.aniso);
    out.sha.warm_f = sha(H.native.simReadF(s.sim, 0));
    H.native.destroy(s.sim);
    const prog = [];
    const res = await H.simulateHIP(dev, g, C.run, (info) => prog.push(info), null,
      { warmStart: net, criterion: C.warmStart.criterion });
    out.steady = res.steady;
    out.metrics = res.metrics;
    out.hex.metrics = {};
    for (const k of Object.keys(res.metrics)) out.hex.metrics[k] = hex64(res.metrics[k]);
    out.hist = res.hist;
    out.progress = prog;
    out.sha.vol_spd = sha(res.vol.spd);
    H.native.destroy(net.net);
    process.stdout.write(JSON.stringify(out));
    return;
  }
  if (C.gateTolerance !== undefined) {
    try {
      await H.runHIP(g, C.run, null, { gateTolerance: C.gateTolerance });
      out.runError = null;
    } catch (err) {
      out.runError = err.message;
    }
    out.selftest = H.gpuSelfTestHIP(dev, g, C.run);
    process.stdout.write(JSON.stringify(out));
    return;
  }
  if (C.run) {
    const prog = [];
    const res = await H.simulateHIP(dev, g, C.run, (info) => prog.push(info), null, { precision: C.precision });
    out.progress = prog;
    out.metrics = res.metrics;
    out.hex.metrics = {};
    for (const k of Object.keys(res.metrics)) out.hex.metrics[k] = hex64(res.metrics[k]);
    out.hist = res.hist;
    out.device_result = res.device;
    out.sha.fmap = sha(res.fmap);
    for (const pl of ["planInlet", "planLow", "lon"]) for (const k of ["speed", "ua", "ub"]) out.sha[`${pl}_${k}`] = sha(res[pl][k]);
    for (const k of ["spd", "pst", "loss"]) out.sha[`vol_${k}`] = sha(res.vol[k]);
    for (const k of ["pts", "spd", "off"]) out.sha[`lines_${k}`] = sha(res.lines[k]);
    out.shape = { planInlet: [res.planInlet.n1, res.planInlet.n2, res.planInlet.depth], lon: [res.lon.n1, res.lon.n2, res.lon.x],
      vol: [res.vol.nx, res.vol.ny, res.vol.nz, res.vol.kB, res.vol.zIn], frame: res.frame, nlines: res.lines.off.length - 1 };
  }
  process.stdout.write(JSON.stringify(out));
})().catch((err) => { process.stdout.write(JSON.stringify({ fatal: err.message })); process.exit(1); });
