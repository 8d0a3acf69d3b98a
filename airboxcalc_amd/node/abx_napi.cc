// N-API addon: the thin binding between Node.js and the C ABI of
// libairbox_hip.so (include/airbox_hip.h).  It marshals typed arrays and
// plain objects, owns nothing but the opaque handles (wrapped as externals
// with finalizers) and turns every failure into a thrown Error carrying the
// library's message.  The JavaScript shim (../js/engine_hip.js) builds the
// reference's API (engine.js: prepare / simulate / evaluate objects) on it.
//
// Long work (stream-collide chunks followed by the probe) runs as
// napi_async_work on the libuv pool and resolves a Promise, so the JS thread
// stays free as with the reference's mapAsync readbacks (engine.js:566-572).
#include <node_api.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/airbox_hip.h"

namespace {

#define NAPI_OK(call)                                      \
  do {                                                     \
    if ((call) != napi_ok) {                               \
      napi_throw_error(env, nullptr, "N-API call failed"); \
      return nullptr;                                      \
    }                                                      \
  } while (0)

napi_value throw_lib(napi_env env) {
  napi_throw_error(env, nullptr, abx_last_error());
  return nullptr;
}

napi_value throw_msg(napi_env env, const std::string& m) {
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

// ---- handles -----------------------------------------------------------------
enum Kind { K_DEV = 1, K_GEOM = 2, K_SIM = 3, K_NET = 4 };
// busy: a simChunk of this simulation handle is running on a pool thread.
// While it runs, synchronous calls on the handle throw and destroy() is
// deferred to the chunk's completion (the handle's stream and state belong to
// the running chunk).
struct Handle {
  int kind;
  void* p;
  int busy;
  bool destroy_pending;
};

void release(Handle* h) {
  if (h->p) {
    if (h->kind == K_SIM) abx_sim_destroy((abx_sim*)h->p);
    if (h->kind == K_GEOM) abx_geom_free((abx_geom*)h->p);
    if (h->kind == K_DEV) abx_device_close((abx_dev*)h->p);
    if (h->kind == K_NET) abx_surrogate_free((abx_surrogate*)h->p);
    h->p = nullptr;
  }
}

void finalize_handle(napi_env, void* data, void*) {
  Handle* h = (Handle*)data;
  release(h);  // a running chunk holds a reference, so busy is 0 here
  delete h;
}

napi_value wrap(napi_env env, int kind, void* p) {
  napi_value v;
  Handle* h = new Handle{kind, p, 0, false};
  if (napi_create_external(env, h, finalize_handle, nullptr, &v) != napi_ok) {
    delete h;
    return nullptr;
  }
  return v;
}

Handle* unwrap(napi_env env, napi_value v, int kind) {
  void* data = nullptr;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_external || napi_get_value_external(env, v, &data) != napi_ok) {
    napi_throw_type_error(env, nullptr, "expected a native handle");
    return nullptr;
  }
  Handle* h = (Handle*)data;
  if (h->kind != kind || !h->p || h->destroy_pending) {
    napi_throw_type_error(env, nullptr,
                          h->p && !h->destroy_pending ? "wrong kind of native handle" : "native handle already destroyed");
    return nullptr;
  }
  return h;
}

// ---- argument helpers ----------------------------------------------------------
bool get_args(napi_env env, napi_callback_info info, size_t need, std::vector<napi_value>& argv) {
  size_t argc = 16;
  argv.assign(16, nullptr);
  napi_get_cb_info(env, info, &argc, argv.data(), nullptr, nullptr);
  argv.resize(argc);
  if (argc < need) {
    napi_throw_type_error(env, nullptr, "too few arguments");
    return false;
  }
  return true;
}

// An integer argument in [lo, hi]; throws a RangeError otherwise (NaN, a
// fraction or a value outside the range).
bool int_arg(napi_env env, napi_value v, int lo, int hi, const char* what, int* out) {
  napi_valuetype t;
  double d = NAN;
  if (v && napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_double(env, v, &d);
  if (!(d >= lo && d <= hi) || d != std::floor(d)) {
    const std::string m = std::string(what) + ": expected an integer in [" + std::to_string(lo) + ", " +
                          std::to_string(hi) + "]";
    napi_throw_range_error(env, nullptr, m.c_str());
    return false;
  }
  *out = (int)d;
  return true;
}

double num(napi_env env, napi_value v, double dflt = NAN) {
  napi_valuetype t;
  if (!v || napi_typeof(env, v, &t) != napi_ok || t != napi_number) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

// num() as an int: dflt unless v is a number in the int range (so no NaN or
// out-of-range value reaches a cast); fractions truncate toward zero.
int inum(napi_env env, napi_value v, int dflt) {
  const double d = num(env, v, NAN);
  return d >= (double)INT32_MIN && d <= (double)INT32_MAX ? (int)d : dflt;
}

napi_value prop(napi_env env, napi_value obj, const char* key) {
  napi_valuetype t;
  if (!obj || napi_typeof(env, obj, &t) != napi_ok || t != napi_object) return nullptr;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) return nullptr;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  return v;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  return !v || napi_typeof(env, v, &t) != napi_ok || t == napi_null || t == napi_undefined;
}

// A typed array's data, element type and length; or an ArrayBuffer's bytes.
struct View {
  void* data = nullptr;
  size_t length = 0;  // elements
  napi_typedarray_type type = napi_uint8_array;
  bool ok = false;
};

View view(napi_env env, napi_value v) {
  View r;
  bool is_ta = false, is_ab = false;
  napi_is_typedarray(env, v, &is_ta);
  if (is_ta) {
    napi_value ab;
    size_t off;
    napi_get_typedarray_info(env, v, &r.type, &r.length, &r.data, &ab, &off);
    r.ok = true;
    return r;
  }
  napi_is_arraybuffer(env, v, &is_ab);
  if (is_ab) {
    napi_get_arraybuffer_info(env, v, &r.data, &r.length);
    r.type = napi_uint8_array;
    r.ok = true;
  }
  return r;
}

bool want(napi_env env, const View& v, napi_typedarray_type t, size_t n, const char* what) {
  if (!v.ok || v.type != t || (n != (size_t)-1 && v.length != n)) {
    std::string m = std::string(what) + ": wrong typed array (type or length)";
    napi_throw_type_error(env, nullptr, m.c_str());
    return false;
  }
  return true;
}

template <typename T>
napi_value typed(napi_env env, napi_typedarray_type t, const T* src, size_t n, T** out = nullptr) {
  napi_value ab, ta;
  void* data = nullptr;
  if (napi_create_arraybuffer(env, n * sizeof(T), &data, &ab) != napi_ok) return nullptr;
  if (src && n) memcpy(data, src, n * sizeof(T));
  if (napi_create_typedarray(env, t, n, ab, 0, &ta) != napi_ok) return nullptr;
  if (out) *out = (T*)data;
  return ta;
}

void set(napi_env env, napi_value obj, const char* key, napi_value v) { napi_set_named_property(env, obj, key, v); }
void setd(napi_env env, napi_value obj, const char* key, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  set(env, obj, key, v);
}
void seti(napi_env env, napi_value obj, const char* key, int64_t i) {
  napi_value v;
  napi_create_int64(env, i, &v);
  set(env, obj, key, v);
}
napi_value obj(napi_env env) {
  napi_value o;
  napi_create_object(env, &o);
  return o;
}

// ---- functions -----------------------------------------------------------------
napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  napi_create_string_utf8(env, abx_version(), NAPI_AUTO_LENGTH, &s);
  return s;
}

napi_value DeviceOpen(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  get_args(env, info, 0, a);
  const int ord = a.empty() ? 0 : inum(env, a[0], 0);
  abx_dev* d = nullptr;
  char name[256];
  if (abx_device_open(ord, &d, name, sizeof name)) return throw_lib(env);
  napi_value o = obj(env), s;
  set(env, o, "dev", wrap(env, K_DEV, d));
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s);
  set(env, o, "name", s);
  return o;
}

napi_value ParseSTL(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 1, a)) return nullptr;
  View v = view(env, a[0]);
  if (!v.ok) return throw_msg(env, "parseSTL: expected an ArrayBuffer or Uint8Array");
  size_t bytes = v.length;
  if (v.type != napi_uint8_array) return throw_msg(env, "parseSTL: expected an ArrayBuffer or Uint8Array");
  int64_t n = 0;
  if (abx_parse_stl((const uint8_t*)v.data, (int64_t)bytes, nullptr, 0, &n)) return throw_lib(env);
  float* out = nullptr;
  napi_value r = typed<float>(env, napi_float32_array, nullptr, (size_t)n, &out);
  if (abx_parse_stl((const uint8_t*)v.data, (int64_t)bytes, out, n, &n)) return throw_lib(env);
  return r;
}

napi_value Voxelize(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 7, a)) return nullptr;
  Handle* d = unwrap(env, a[0], K_DEV);
  if (!d) return nullptr;
  View T = view(env, a[1]);
  if (!want(env, T, napi_float32_array, (size_t)-1, "voxelize T")) return nullptr;
  double origin[3];
  for (uint32_t k = 0; k < 3; ++k) {
    napi_value e;
    napi_get_element(env, a[2], k, &e);
    origin[k] = num(env, e);
  }
  const int nx = inum(env, a[3], 0), ny = inum(env, a[4], 0), nz = inum(env, a[5], 0);
  const double h = num(env, a[6]);
  uint8_t* out = nullptr;
  napi_value r = typed<uint8_t>(env, napi_uint8_array, nullptr, (size_t)nx * ny * nz, &out);
  if (abx_voxelize((abx_dev*)d->p, (const float*)T.data, (int64_t)(T.length / 9), origin, nx, ny, nz, h, out))
    return throw_lib(env);
  return r;
}

napi_value geom_object(napi_env env, abx_geom* g) {
  abx_geom_info I;
  abx_geom_info_get(g, &I);
  napi_value o = obj(env), info = obj(env);
  set(env, o, "geom", wrap(env, K_GEOM, g));
  seti(env, info, "nx", I.nx);
  seti(env, info, "ny", I.ny);
  seti(env, info, "nz", I.nz);
  seti(env, info, "N", I.N);
  seti(env, info, "nf", I.nf);
  seti(env, info, "top", I.top);
  seti(env, info, "jm", I.jm);
  seti(env, info, "kw0", I.kw0);
  setd(env, info, "h", I.h);
  const char* vec3[3] = {"origin", "mn", "mx"};
  const double* src3[3] = {I.origin, I.mn, I.mx};
  for (int k = 0; k < 3; ++k) {
    napi_value arr;
    napi_create_array_with_length(env, 3, &arr);
    for (uint32_t c = 0; c < 3; ++c) {
      napi_value e;
      napi_create_double(env, src3[k][c], &e);
      napi_set_element(env, arr, c, e);
    }
    set(env, info, vec3[k], arr);
  }
  seti(env, info, "nHoles", I.nHoles);
  seti(env, info, "nOpen", I.nOpen);
  seti(env, info, "nOut", I.nOut);
  seti(env, info, "Af", I.Af);
  seti(env, info, "leak", I.leak);
  seti(env, info, "pleatOk", I.pleatOk);
  seti(env, info, "hasOutlet", I.has_outlet);
  setd(env, info, "volume_l", I.volume_l);
  setd(env, info, "inlet_cm2", I.inlet_cm2);
  setd(env, info, "filter_cm2", I.filter_cm2);
  setd(env, info, "outlet_cm2", I.outlet_cm2);
  setd(env, info, "outlet_cm2_real", I.outlet_cm2_real);
  set(env, o, "info", info);
  const int64_t N = I.N;
  struct {
    const char* key;
    int what;
    int64_t n;
    napi_typedarray_type t;
    size_t esz;
  } arrs[] = {{"solid", ABX_G_SOLID, N, napi_uint8_array, 1},
              {"fluid", ABX_G_FLUID, N, napi_uint8_array, 1},
              {"opening", ABX_G_OPENING, (int64_t)I.nx * I.ny, napi_uint8_array, 1},
              {"outlet", ABX_G_OUTLET, N, napi_uint8_array, 1},
              {"xc", ABX_G_XC, I.nx, napi_float32_array, 4},
              {"dc", ABX_G_DC, I.ny, napi_float32_array, 4},
              {"zc", ABX_G_ZC, I.nz, napi_float32_array, 4},
              {"openPts", ABX_G_OPENPTS, 3 * I.nOpenPts, napi_float64_array, 8},
              {"filterPts", ABX_G_FILTERPTS, 3 * I.nFilterPts, napi_float64_array, 8},
              {"outletPts", ABX_G_OUTLETPTS, 3 * I.nOutletPts, napi_float64_array, 8}};
  for (auto& A : arrs) {
    napi_value ab, ta;
    void* data = nullptr;
    napi_create_arraybuffer(env, (size_t)A.n * A.esz, &data, &ab);
    if (A.n && abx_geom_read(g, A.what, data, A.n)) return throw_lib(env);
    napi_create_typedarray(env, A.t, (size_t)A.n, ab, 0, &ta);
    set(env, o, A.key, ta);
  }
  return o;
}

// prepare's parameter object (ui.html:477-479) plus cy/sy and maxCells.
abx_prep_params prep_params(napi_env env, napi_value p) {
  abx_prep_params P;
  memset(&P, 0, sizeof P);
  P.dx = num(env, prop(env, p, "dx"));
  P.sealAxis = inum(env, prop(env, p, "sealAxis"), 0);
  P.sealSign = inum(env, prop(env, p, "sealSign"), 1);
  P.seal = num(env, prop(env, p, "seal"), 0);
  P.inletAxis = inum(env, prop(env, p, "inletAxis"), 0);
  P.inletSign = inum(env, prop(env, p, "inletSign"), 1);
  P.inletYaw = num(env, prop(env, p, "inletYaw"), 0);
  P.cy = num(env, prop(env, p, "cy"));
  P.sy = num(env, prop(env, p, "sy"));
  P.pleat = num(env, prop(env, p, "pleat"));
  P.max_cells = num(env, prop(env, p, "maxCells"), 4.5e6);
  napi_value click = prop(env, p, "click");
  if (!is_nullish(env, click)) {
    P.has_click = 1;
    for (uint32_t k = 0; k < 3; ++k) {
      napi_value e;
      napi_get_element(env, click, k, &e);
      P.click[k] = num(env, e);
    }
  }
  napi_value out = prop(env, p, "outlet");
  if (!is_nullish(env, out)) {
    bool en = false;
    napi_value ev = prop(env, out, "enabled");
    if (ev) napi_get_value_bool(env, ev, &en);
    if (en) {
      P.has_outlet = 1;
      napi_value c = prop(env, out, "center");
      for (uint32_t k = 0; k < 3; ++k) {
        napi_value e;
        napi_get_element(env, c, k, &e);
        P.out_center[k] = num(env, e);
      }
      P.out_diameter = num(env, prop(env, out, "diameter"));
      P.out_height = num(env, prop(env, out, "height"));
    }
  }
  return P;
}

napi_value Prepare(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 3, a)) return nullptr;
  Handle* d = unwrap(env, a[0], K_DEV);
  if (!d) return nullptr;
  View T = view(env, a[1]);
  if (!want(env, T, napi_float32_array, (size_t)-1, "prepare tris")) return nullptr;
  const abx_prep_params P = prep_params(env, a[2]);
  abx_geom* g = nullptr;
  if (abx_prepare((abx_dev*)d->p, (const float*)T.data, (int64_t)T.length, &P, &g)) return throw_lib(env);
  return geom_object(env, g);
}

napi_value sim_info(napi_env env, abx_sim* s) {
  abx_sim_info I;
  abx_sim_info_get(s, &I);
  napi_value o = obj(env);
  seti(env, o, "NF", I.NF);
  seti(env, o, "Ain", I.Ain);
  seti(env, o, "Af", I.Af);
  seti(env, o, "Lp", I.Lp);
  seti(env, o, "nM", I.nM);
  seti(env, o, "nT", I.nT);
  setd(env, o, "u0", I.u0);
  setd(env, o, "tau0", I.tau0);
  setd(env, o, "sig", I.sig);
  setd(env, o, "aniso", I.aniso);
  setd(env, o, "cs2s", I.cs2s);
  seti(env, o, "precision", I.precision);
  seti(env, o, "B", I.B);
  seti(env, o, "deviceBytes", I.device_bytes);
  return o;
}

napi_value SimCreate(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 3, a)) return nullptr;
  Handle* d = unwrap(env, a[0], K_DEV);
  if (!d) return nullptr;
  Handle* g = unwrap(env, a[1], K_GEOM);
  if (!g) return nullptr;
  abx_run_params rp;
  rp.K = num(env, prop(env, a[2], "K"));
  rp.aniso = num(env, prop(env, a[2], "aniso"));
  rp.u = num(env, prop(env, a[2], "u"));
  rp.tau = num(env, prop(env, a[2], "tau"));
  const int prec = a.size() > 3 ? inum(env, a[3], 0) : 0;
  const int B = a.size() > 4 ? inum(env, a[4], 1) : 1;
  abx_grid grid;
  abx_geom_grid((abx_geom*)g->p, &grid);
  const float* ns = nullptr;
  if (a.size() > 5 && !is_nullish(env, a[5])) {
    View v = view(env, a[5]);
    if (!want(env, v, napi_float32_array, (size_t)-1, "simCreate ns")) return nullptr;
    ns = (const float*)v.data;
  }
  abx_sim* s = nullptr;
  if (abx_sim_create((abx_dev*)d->p, &grid, &rp, prec, B, ns, &s)) return throw_lib(env);
  napi_value o = obj(env);
  set(env, o, "sim", wrap(env, K_SIM, s));
  set(env, o, "info", sim_info(env, s));
  return o;
}

// The simulation of a handle that no chunk is running on (synchronous calls).
abx_sim* sim_arg(napi_env env, napi_value v) {
  Handle* h = unwrap(env, v, K_SIM);
  if (!h) return nullptr;
  if (h->busy) {
    napi_throw_error(env, nullptr, "simulation handle busy: a simChunk is still running on it");
    return nullptr;
  }
  return (abx_sim*)h->p;
}

napi_value SimSteps(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 3, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  int n;
  if (!int_arg(env, a[1], 0, 1 << 30, "simSteps n", &n)) return nullptr;
  bool doacc = false;
  napi_get_value_bool(env, a[2], &doacc);
  if (abx_sim_steps(s, n, doacc ? 1 : 0) || abx_sim_sync(s)) return throw_lib(env);
  return nullptr;
}

// simSetLaunch(sim, kernel, xcdMap, nt): abx_sim_set_launch.
napi_value SimSetLaunch(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 2, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  int kernel, xm = -1, nt = 19;
  if (!int_arg(env, a[1], 0, 16, "simSetLaunch kernel", &kernel)) return nullptr;
  if (a.size() > 2 && !int_arg(env, a[2], -1, 4096, "simSetLaunch xcdMap", &xm)) return nullptr;
  if (a.size() > 3 && !int_arg(env, a[3], 0, 65535, "simSetLaunch nt", &nt)) return nullptr;
  if (abx_sim_set_launch(s, kernel, xm, nt)) return throw_lib(env);
  return nullptr;
}

napi_value probe_object(napi_env env, const double out[3]) {
  napi_value o = obj(env);
  setd(env, o, "gamma", out[0]);
  setd(env, o, "massRatio", out[1]);
  setd(env, o, "qo", out[2]);
  return o;
}

napi_value SimProbe(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 2, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  int step;
  if (!int_arg(env, a[1], 0, INT32_MAX, "simProbe step", &step)) return nullptr;
  double out[3];
  if (abx_sim_probe(s, step, out)) return throw_lib(env);
  return probe_object(env, out);
}

// simChunk(sim, n, doacc, step[, probe = true]) -> Promise<{gamma,
// massRatio, qo} | undefined>: n steps, then the probe, on a pool thread.
struct ChunkWork {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref simref;
  Handle* h;
  abx_sim* sim;
  int n, doacc, step, probe, rc;
  double out[3];
  std::string err;
};

void chunk_exec(napi_env, void* data) {
  ChunkWork* w = (ChunkWork*)data;
  w->rc = abx_sim_steps(w->sim, w->n, w->doacc);
  if (!w->rc) w->rc = w->probe ? abx_sim_probe(w->sim, w->step, w->out) : abx_sim_sync(w->sim);
  if (w->rc) w->err = abx_last_error();
}

void chunk_done(napi_env env, napi_status, void* data) {
  ChunkWork* w = (ChunkWork*)data;
  if (w->rc) {
    napi_value msg, err;
    napi_create_string_utf8(env, w->err.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else if (w->probe) {
    napi_resolve_deferred(env, w->deferred, probe_object(env, w->out));
  } else {
    napi_value u;
    napi_get_undefined(env, &u);
    napi_resolve_deferred(env, w->deferred, u);
  }
  if (--w->h->busy == 0 && w->h->destroy_pending) release(w->h);
  napi_delete_reference(env, w->simref);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value SimChunk(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 4, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  int n, step;
  if (!int_arg(env, a[1], 0, 1 << 30, "simChunk n", &n) ||
      !int_arg(env, a[3], 0, INT32_MAX, "simChunk step", &step))
    return nullptr;
  ChunkWork* w = new ChunkWork();
  w->h = unwrap(env, a[0], K_SIM);
  w->h->busy++;
  w->sim = s;
  w->n = n;
  bool doacc = false;
  napi_get_value_bool(env, a[2], &doacc);
  w->doacc = doacc ? 1 : 0;
  w->step = step;
  bool probe = true;
  if (a.size() > 4) napi_get_value_bool(env, a[4], &probe);
  w->probe = probe ? 1 : 0;
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_reference(env, a[0], 1, &w->simref);  // keep the handle alive
  napi_create_string_utf8(env, "abx_sim_chunk", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, chunk_exec, chunk_done, w, &w->work);
  napi_queue_async_work(env, w->work);
  return promise;
}

napi_value SimReadAcc(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 3, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  abx_sim_info I;
  abx_sim_info_get(s, &I);
  float* out = nullptr;
  napi_value r = typed<float>(env, napi_float32_array, nullptr, (size_t)(4 * I.NF), &out);
  if (abx_sim_read_acc(s, inum(env, a[1], -1), num(env, a[2], 1.0), out)) return throw_lib(env);
  return r;
}

napi_value SimReadF(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 2, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  abx_sim_info I;
  abx_sim_info_get(s, &I);
  const size_t n = (size_t)(19 * I.NF);
  napi_value r;
  void* out = nullptr;
  if (I.precision == ABX_F64)
    r = typed<double>(env, napi_float64_array, nullptr, n, (double**)&out);
  else
    r = typed<float>(env, napi_float32_array, nullptr, n, (float**)&out);
  if (abx_sim_read_f(s, inum(env, a[1], -1), out)) return throw_lib(env);
  return r;
}

napi_value SimWriteF(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 3, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  abx_sim_info I;
  abx_sim_info_get(s, &I);
  View v = view(env, a[2]);
  if (!want(env, v, I.precision == ABX_F64 ? napi_float64_array : napi_float32_array, (size_t)(19 * I.NF),
            "simWriteF"))
    return nullptr;
  if (abx_sim_write_f(s, inum(env, a[1], -2), v.data)) return throw_lib(env);
  return nullptr;
}

napi_value SimInitEq(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 2, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  abx_sim_info I;
  abx_sim_info_get(s, &I);
  View v = view(env, a[1]);
  if (!want(env, v, napi_float32_array, (size_t)(4 * I.NF), "simInitEq")) return nullptr;
  if (abx_sim_init_eq(s, (const float*)v.data)) return throw_lib(env);
  return nullptr;
}

napi_value SimSetNS(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 2, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  const float* ns = nullptr;
  if (!is_nullish(env, a[1])) {
    abx_sim_info I;
    abx_sim_info_get(s, &I);
    View v = view(env, a[1]);
    if (!want(env, v, napi_float32_array, (size_t)(I.B * I.NF), "simSetNS")) return nullptr;
    ns = (const float*)v.data;
  }
  if (abx_sim_set_ns(s, ns)) return throw_lib(env);
  return nullptr;
}

napi_value SimClearAcc(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 1, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  if (abx_sim_clear_acc(s) || abx_sim_sync(s)) return throw_lib(env);
  return nullptr;
}

napi_value SimInfo(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 1, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  return sim_info(env, s);
}

napi_value metrics_object(napi_env env, const abx_metrics& m) {
  napi_value met = obj(env);
  setd(env, met, "gamma", m.gamma);
  setd(env, met, "gammaCore", m.gammaCore);
  setd(env, met, "lowCore", m.lowCore);
  setd(env, met, "highCore", m.highCore);
  setd(env, met, "reverse", m.reverse);
  setd(env, met, "maxToMean", m.maxToMean);
  setd(env, met, "zeta", m.zeta);
  setd(env, met, "inlet_cm2", m.inlet_cm2);
  setd(env, met, "filter_cm2", m.filter_cm2);
  return met;
}

napi_value Evaluate(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 4, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  Handle* g = unwrap(env, a[1], K_GEOM);
  if (!g) return nullptr;
  const int b = inum(env, a[2], 0);
  const double div = num(env, a[3], 1.0);
  bool f3 = true;
  if (a.size() > 4) napi_get_value_bool(env, a[4], &f3);
  abx_eval_geom eg;
  abx_geom_eval_view((abx_geom*)g->p, &eg);
  const size_t L = (size_t)eg.nx * eg.nz, C = (size_t)eg.nz * eg.ny, N = (size_t)eg.nx * eg.ny * eg.nz;
  float *fmap, *pI, *pL, *lon, *spd = nullptr, *pst = nullptr, *loss = nullptr;
  napi_value o = obj(env);
  set(env, o, "fmap", typed<float>(env, napi_float32_array, nullptr, L, &fmap));
  set(env, o, "planInlet", typed<float>(env, napi_float32_array, nullptr, 3 * L, &pI));
  set(env, o, "planLow", typed<float>(env, napi_float32_array, nullptr, 3 * L, &pL));
  set(env, o, "lon", typed<float>(env, napi_float32_array, nullptr, 3 * C, &lon));
  if (f3) {
    set(env, o, "spd", typed<float>(env, napi_float32_array, nullptr, N, &spd));
    set(env, o, "pst", typed<float>(env, napi_float32_array, nullptr, N, &pst));
    set(env, o, "loss", typed<float>(env, napi_float32_array, nullptr, N, &loss));
  }
  abx_metrics m;
  if (abx_evaluate(s, &eg, b, div, &m, fmap, pI, pL, lon, spd, pst, loss)) return throw_lib(env);
  set(env, o, "metrics", metrics_object(env, m));
  seti(env, o, "ic", m.ic);
  seti(env, o, "jc", m.jc);
  seti(env, o, "jlow", m.jlow);
  seti(env, o, "jb", m.jb);
  seti(env, o, "kB", m.kB);
  setd(env, o, "rhoRef", m.rhoRef);
  setd(env, o, "pin", m.pin);
  return o;
}

napi_value Streamlines(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 4, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  Handle* g = unwrap(env, a[1], K_GEOM);
  if (!g) return nullptr;
  const int b = inum(env, a[2], 0);
  const double div = num(env, a[3], 1.0);
  const int maxSeeds = a.size() > 4 ? inum(env, a[4], 220) : 220;
  abx_eval_geom eg;
  abx_geom_eval_view((abx_geom*)g->p, &eg);
  if (maxSeeds < 1 || maxSeeds > 100000) return throw_msg(env, "streamlines: maxSeeds out of range");
  // at most maxSeeds + 1 lines (the fractional seed stride) of at most 2001
  // points (every 3rd of 6000 RK2 steps plus the last one), engine.js:738-748
  const int64_t cap = 2002 * ((int64_t)maxSeeds + 1) + 16;
  std::vector<float> pts((size_t)(3 * cap)), spd((size_t)cap);
  std::vector<uint32_t> off((size_t)maxSeeds + 3);
  int64_t npts = 0, nl = 0;
  if (abx_streamlines(s, &eg, b, div, maxSeeds, pts.data(), (int64_t)pts.size(), spd.data(), (int64_t)spd.size(),
                      off.data(), (int64_t)off.size(), &npts, &nl))
    return throw_lib(env);
  napi_value o = obj(env);
  set(env, o, "pts", typed<float>(env, napi_float32_array, pts.data(), (size_t)(3 * npts)));
  set(env, o, "spd", typed<float>(env, napi_float32_array, spd.data(), (size_t)npts));
  set(env, o, "off", typed<uint32_t>(env, napi_uint32_array, off.data(), (size_t)(nl + 1)));
  return o;
}

// surrogateLoad(dev, path) -> {net, dims, familyDx}: abx_surrogate_load.
napi_value SurrogateLoad(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 2, a)) return nullptr;
  Handle* d = unwrap(env, a[0], K_DEV);
  if (!d) return nullptr;
  size_t len = 0;
  if (napi_get_value_string_utf8(env, a[1], nullptr, 0, &len) != napi_ok) return throw_msg(env, "surrogateLoad: path");
  std::string path(len, '\0');
  napi_get_value_string_utf8(env, a[1], &path[0], len + 1, &len);
  abx_surrogate* n = nullptr;
  if (abx_surrogate_load((abx_dev*)d->p, path.c_str(), &n)) return throw_lib(env);
  int dims[3];
  double fdx = 0;
  abx_surrogate_info(n, dims, &fdx);
  napi_value o = obj(env), da;
  set(env, o, "net", wrap(env, K_NET, n));
  napi_create_array_with_length(env, 3, &da);
  for (uint32_t k = 0; k < 3; ++k) {
    napi_value e;
    napi_create_int32(env, dims[k], &e);
    napi_set_element(env, da, k, e);
  }
  set(env, o, "dims", da);
  setd(env, o, "familyDx", fdx);
  return o;
}

// surrogateWarmStart(net, sim, geom, tris, prepParams, K, aniso):
// abx_surrogate_warm_start (synchronous; a few ms).
napi_value SurrogateWarmStart(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 7, a)) return nullptr;
  Handle* n = unwrap(env, a[0], K_NET);
  if (!n) return nullptr;
  abx_sim* s = sim_arg(env, a[1]);
  if (!s) return nullptr;
  Handle* g = unwrap(env, a[2], K_GEOM);
  if (!g) return nullptr;
  const float* tris = nullptr;
  int64_t nf = 0;
  abx_prep_params P;
  const abx_prep_params* pp = nullptr;
  if (!is_nullish(env, a[3])) {
    View T = view(env, a[3]);
    if (!want(env, T, napi_float32_array, (size_t)-1, "surrogateWarmStart tris")) return nullptr;
    tris = (const float*)T.data;
    nf = (int64_t)T.length;
    P = prep_params(env, a[4]);
    pp = &P;
  }
  if (abx_surrogate_warm_start((abx_surrogate*)n->p, s, (abx_geom*)g->p, tris, nf, pp, num(env, a[5]),
                               num(env, a[6])))
    return throw_lib(env);
  return nullptr;
}

// solveToSteady(sim, geom, {window, maxSteps, step0, tolGamma, tolZeta},
// onWindow) -> Promise<{converged, steps, windows, metrics}>: the steady
// criterion (abx_sim_solve_to_steady) on a pool thread; onWindow gets
// {step, gamma, massRatio, metrics} of every window on the JS thread.
struct SteadyWork {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref simref, geomref;
  Handle *hs, *hg;
  abx_sim* sim;
  abx_eval_geom eg;
  abx_steady_params p;
  napi_threadsafe_function tsfn = nullptr;
  int rc = 0;
  std::string err;
  abx_steady_result res;
};

int steady_window_cb(void* user, const abx_steady_window* w) {
  SteadyWork* sw = (SteadyWork*)user;
  if (sw->tsfn) napi_call_threadsafe_function(sw->tsfn, new abx_steady_window(*w), napi_tsfn_blocking);
  return 0;
}

void steady_call_js(napi_env env, napi_value cb, void*, void* data) {
  abx_steady_window* w = (abx_steady_window*)data;
  if (env && cb) {
    napi_value o = obj(env), u;
    seti(env, o, "step", w->step);
    setd(env, o, "gamma", w->gamma);
    setd(env, o, "massRatio", w->massRatio);
    set(env, o, "metrics", metrics_object(env, w->metrics));
    napi_get_undefined(env, &u);
    napi_call_function(env, u, cb, 1, &o, nullptr);
  }
  delete w;
}

void steady_exec(napi_env, void* data) {
  SteadyWork* w = (SteadyWork*)data;
  w->rc = abx_sim_solve_to_steady(w->sim, &w->eg, &w->p, steady_window_cb, w, &w->res);
  if (w->rc) w->err = abx_last_error();
}

void steady_done(napi_env env, napi_status, void* data) {
  SteadyWork* w = (SteadyWork*)data;
  if (w->tsfn) napi_release_threadsafe_function(w->tsfn, napi_tsfn_release);
  if (w->rc) {
    napi_value msg, err;
    napi_create_string_utf8(env, w->err.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_value o = obj(env), b;
    napi_get_boolean(env, w->res.converged != 0, &b);
    set(env, o, "converged", b);
    seti(env, o, "steps", w->res.steps);
    seti(env, o, "windows", w->res.windows);
    set(env, o, "metrics", metrics_object(env, w->res.metrics));
    napi_resolve_deferred(env, w->deferred, o);
  }
  if (--w->hs->busy == 0 && w->hs->destroy_pending) release(w->hs);
  if (--w->hg->busy == 0 && w->hg->destroy_pending) release(w->hg);
  napi_delete_reference(env, w->simref);
  napi_delete_reference(env, w->geomref);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value SolveToSteady(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 3, a)) return nullptr;
  abx_sim* s = sim_arg(env, a[0]);
  if (!s) return nullptr;
  Handle* hg = unwrap(env, a[1], K_GEOM);
  if (!hg) return nullptr;
  abx_steady_params P;
  if (!int_arg(env, prop(env, a[2], "window"), 1, 1 << 30, "solveToSteady window", &P.window) ||
      !int_arg(env, prop(env, a[2], "maxSteps"), 1, INT32_MAX, "solveToSteady maxSteps", &P.max_steps))
    return nullptr;
  P.step0 = 0;
  napi_value s0 = prop(env, a[2], "step0");
  if (s0 && !int_arg(env, s0, 0, INT32_MAX, "solveToSteady step0", &P.step0)) return nullptr;
  P.tol_gamma = num(env, prop(env, a[2], "tolGamma"), 0.005);
  P.tol_zeta = num(env, prop(env, a[2], "tolZeta"), 0.02);
  SteadyWork* w = new SteadyWork();
  w->hs = unwrap(env, a[0], K_SIM);
  w->hg = hg;
  w->sim = s;
  w->p = P;
  abx_geom_eval_view((abx_geom*)hg->p, &w->eg);
  napi_value name, promise;
  napi_create_string_utf8(env, "abx_solve_to_steady", NAPI_AUTO_LENGTH, &name);
  napi_valuetype t = napi_undefined;
  if (a.size() > 3) napi_typeof(env, a[3], &t);
  if (t == napi_function &&
      napi_create_threadsafe_function(env, a[3], nullptr, name, 0, 1, nullptr, nullptr, nullptr, steady_call_js,
                                      &w->tsfn) != napi_ok) {
    delete w;
    return throw_msg(env, "solveToSteady: progress callback");
  }
  w->hs->busy++;
  w->hg->busy++;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_reference(env, a[0], 1, &w->simref);  // keep the handles alive
  napi_create_reference(env, a[1], 1, &w->geomref);
  napi_create_async_work(env, nullptr, name, steady_exec, steady_done, w, &w->work);
  napi_queue_async_work(env, w->work);
  return promise;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  std::vector<napi_value> a;
  if (!get_args(env, info, 1, a)) return nullptr;
  void* data = nullptr;
  if (napi_get_value_external(env, a[0], &data) != napi_ok || !data) return nullptr;
  Handle* h = (Handle*)data;
  if (h->busy)
    h->destroy_pending = true;  // released when the running chunk completes
  else
    release(h);
  return nullptr;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback cb;
  } fns[] = {{"version", Version},       {"deviceOpen", DeviceOpen}, {"parseSTL", ParseSTL},
             {"voxelize", Voxelize},     {"prepare", Prepare},       {"simCreate", SimCreate},
             {"simInfo", SimInfo},       {"simSteps", SimSteps},     {"simChunk", SimChunk},
             {"simProbe", SimProbe},     {"simReadAcc", SimReadAcc}, {"simReadF", SimReadF},
             {"simWriteF", SimWriteF},   {"simInitEq", SimInitEq},   {"simSetNS", SimSetNS},
             {"simClearAcc", SimClearAcc}, {"evaluate", Evaluate},   {"streamlines", Streamlines},
             {"simSetLaunch", SimSetLaunch}, {"surrogateLoad", SurrogateLoad},
             {"surrogateWarmStart", SurrogateWarmStart}, {"solveToSteady", SolveToSteady}, {"destroy", Destroy}};
  for (auto& f : fns) {
    napi_value fn;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn);
    napi_set_named_property(env, exports, f.name, fn);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
