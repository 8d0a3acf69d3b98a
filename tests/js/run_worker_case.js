// Drives airboxcalc_amd/js/worker_hip.js through the reference's worker
// protocol (gpuinfo, load, prepare, run) and, with a GPU, runs the same case
// through runHIP in this thread for comparison.  Prints one JSON object.
//
// Usage: node run_worker_case.js <case.json>   (case: {stl, prepare, run})
"use strict";
const fs = require("fs");
const path = require("path");
const crypto = require("crypto");
const { Worker } = require("worker_threads");
const H = require(path.join(__dirname, "..", "..", "airboxcalc_amd", "js", "engine_hip.js"));

const C = JSON.parse(fs.readFileSync(process.argv[2], "utf8"));
const sha = (ta) => crypto.createHash("sha256").update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest("hex");
const stlBuffer = () => { const b = fs.readFileSync(C.stl); return b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength); };

(async () => {
  const w = new Worker(path.join(__dirname, "..", "..", "airboxcalc_amd", "js", "worker_hip.js"));
  const out = { status: [], progress: 0, errors: [] };
  let waiting = null;
  w.on("message", (m) => {
    if (m.type === "status") out.status.push(m.text);
    else if (m.type === "progress") out.progress++;
    else {
      if (m.type === "error") out.errors.push({ message: m.message, during: m.during });
      if (waiting) { const r = waiting; waiting = null; r(m); }
    }
  });
  const ask = (m, tr) => new Promise((res) => { waiting = res; w.postMessage(m, tr || []); });
  const gi = await ask({ type: "gpuinfo" });
  out.gpuinfo = { ok: gi.ok, name: gi.name };
  const buf = stlBuffer();
  const ld = await ask({ type: "load", buffer: buf }, [buf]);
  out.loaded = { type: ld.type, ntri: ld.ntri, mn: ld.mn, mx: ld.mx, transferred: buf.byteLength === 0 };
  const pr = await ask({ type: "prepare", params: C.prepare });
  out.prepared = pr.type === "prepared" ? { grid: pr.summary.grid, cells: pr.summary.cells } : null;
  if (C.run && out.prepared) {
    const rs = await ask({ type: "run", params: C.run });
    if (rs.type === "result") {
      const r = rs.res;
      out.result = { metrics: r.metrics, hist: r.hist, device: r.device, steady: r.steady || null, sha: {} };
      for (const k of ["spd", "pst", "loss"]) out.result.sha["vol_" + k] = sha(r.vol[k]);
      for (const k of ["pts", "spd", "off"]) out.result.sha["lines_" + k] = sha(r.lines[k]);
      // the direct call in this thread, for comparison
      const tris = H.parseSTL(stlBuffer());
      const g = H.prepare(Object.assign({ tris }, C.prepare), H.hipDevice(0));
      const opts = {};
      if (C.run.warmStart)
        opts.warmStart = H.loadSurrogate(H.hipDevice(0), H.SURROGATE_WEIGHTS[C.run.warmStart] || C.run.warmStart);
      if (C.run.criterion) opts.criterion = C.run.criterion;
      const d = await H.runHIP(g, C.run, null, opts);
      out.direct = { metrics: d.metrics, hist: d.hist, steady: d.steady || null, sha: {} };
      for (const k of ["spd", "pst", "loss"]) out.direct.sha["vol_" + k] = sha(d.vol[k]);
      for (const k of ["pts", "spd", "off"]) out.direct.sha["lines_" + k] = sha(d.lines[k]);
    }
  }
  await w.terminate();
  process.stdout.write(JSON.stringify(out));
})().catch((e) => { console.error(e); process.exit(1); });
