/* worker_hip.js - the reference's worker protocol (engine.js:760-801) served
   by the HIP backend, for Node hosts (worker_threads; Electron's Node
   integration):

     const { Worker } = require("worker_threads");
     const w = new Worker("airboxcalc_amd/js/worker_hip.js");
     w.postMessage({ type: "load", buffer }, [buffer]);   // -> {type: "loaded", ...}
     w.postMessage({ type: "prepare", params });          // -> {type: "prepared", summary, ms}
     w.postMessage({ type: "run", params });              // -> status / progress ... {type: "result", res}

   Messages in and out are the reference's: "load" / "gpuinfo" / "prepare" /
   "run" in; "loaded" / "gpuinfo" / "prepared" / "status" / "progress" /
   "result" / "error" out, with the same fields and the same transfer lists
   (the result's volume and streamline arrays move, engine.js:797-798).  The
   one difference: there is no CPU fallback in this backend.  Without a gfx950
   GPU, or when the self-test gate fails, "run" answers with an error message
   and the host keeps the reference's own worker for the CPU path. */
"use strict";
const { parentPort } = require("worker_threads");
const H = require("./engine_hip.js");

let tris = null, geo = null, dev = null, devTried = false;
const getDev = () => {
  if (!devTried) {
    devTried = true;
    dev = H.hipDevice(0);
  }
  return dev;
};
const post = (m, tr) => parentPort.postMessage(m, tr || []);

/* The trained networks loaded so far, by weight file. */
const nets = new Map();
const surrogate = (d, which) => {
  const file = H.SURROGATE_WEIGHTS[which] || which;
  if (!nets.has(file)) nets.set(file, H.loadSurrogate(d, file));
  return nets.get(file);
};

/* The run parameters are the reference's (engine.js:765-779); a caller may
   add gateTolerance to tighten the self-test gate (default 1e-3), warmStart
   ("family", "dx2" or a weight file: start from the surrogate's prediction)
   and criterion ({window, tolGamma, tolZeta, maxSteps}: run to the steady
   criterion instead of steps / avg; the result then carries .steady). */
async function runSim(p) {
  const prog = (info) => post({ type: "progress", info });
  const d = getDev();
  if (!d) throw new Error("Keine MI355X-Grafikkarte (gfx950) verfügbar.");
  post({ type: "status", text: "Grafikkarte wird geprüft …" });
  const tol = p.gateTolerance !== undefined ? p.gateTolerance : 1e-3;
  const md = H.gpuSelfTestHIP(d, geo, p);
  if (!(md < tol)) throw new Error("Abweichung GPU/CPU " + md.toExponential(1));
  post({ type: "status", text: "Rechnet auf der Grafikkarte (" + d.name + ")" });
  let opts = null;
  if (p.warmStart || p.criterion) {
    opts = {};
    if (p.warmStart) opts.warmStart = surrogate(d, p.warmStart);
    if (p.criterion) opts.criterion = p.criterion;
  }
  const res = await H.simulateHIP(d, geo, p, prog, null, opts);
  res.device = "GPU · " + d.name;
  return res;
}

parentPort.on("message", async (m) => {
  try {
    if (m.type === "load") {
      tris = H.parseSTL(m.buffer);
      geo = null;
      const mn = [Infinity, Infinity, Infinity], mx = [-Infinity, -Infinity, -Infinity];
      for (let v = 0; v < tris.length; v += 3)
        for (let a = 0; a < 3; a++) {
          mn[a] = Math.min(mn[a], tris[v + a]);
          mx[a] = Math.max(mx[a], tris[v + a]);
        }
      const copy = tris.slice();
      post({ type: "loaded", tris: copy, ntri: tris.length / 9, mn, mx }, [copy.buffer]);
    } else if (m.type === "gpuinfo") {
      const d = getDev();
      post({ type: "gpuinfo", ok: !!d, name: d ? d.name : "" });
    } else if (m.type === "prepare") {
      const t = Date.now();
      geo = H.prepare(Object.assign({ tris }, m.params), getDev());
      post({ type: "prepared", summary: geo.summary, ms: Date.now() - t });
    } else if (m.type === "run") {
      if (!geo) throw new Error("Bitte zuerst die Geometrie prüfen.");
      const res = await runSim(m.params), v = res.vol, L = res.lines;
      post({ type: "result", res }, [v.spd.buffer, v.pst.buffer, v.loss.buffer, L.pts.buffer, L.spd.buffer, L.off.buffer]);
    }
  } catch (err) {
    post({ type: "error", message: err.message || String(err), during: m.type });
  }
});
