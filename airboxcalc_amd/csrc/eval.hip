// Post-processing on the GPU: evaluate (engine.js:632-691), field3d
// (engine.js:698-719) and streamlines (engine.js:721-758), read directly from
// the dense averaged fields of a simulation handle.
//
// Every per-cell output is an elementwise map and matches the reference bit
// for bit.  The scalar sums (mean filter velocity, gamma, gamma-core, the
// zeta densities, the inlet total pressure) are block reductions with wave
// shuffles in double, so they agree with the reference's sequential sums to
// rounding (relative 1e-12), and quantities normalised by them (fmap, pst,
// loss) to within one float32 ulp.
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "abx_common.hpp"
#include "sim_internal.hpp"

namespace abx {

struct EvalDev {
  const void* acc;  // [4][plane] of the variant
  const float* ns;  // [plane] of the variant or null
  const uint32_t* fluid_bits;
  const int32_t* rowid;   // nx*ny: row of (i, j), -1 without fluid
  const int4* rowinfo;    // per row {rowoff, kmin, kmax, i*ny + j}
  int64_t plane;
  int nx, ny, nz;
  double div;
};

// Lattice position of cell (i, j, k) of the grid, -1 where no position is
// stored (a row without fluid, or k outside its row's span).
__device__ __forceinline__ int64_t dpos(const EvalDev& e, int64_t i, int64_t j, int64_t k) {
  const int r = e.rowid[i * e.ny + j];
  if (r < 0) return -1;
  const int4 ri = e.rowinfo[r];
  return k < ri.y || k > ri.z ? -1 : (int64_t)ri.x + k;
}
// fluid of evaluate: g.fluid minus cells with ns > 0.5 (engine.js:634)
__device__ __forceinline__ bool efluid(const EvalDev& e, int64_t d) {
  if (d < 0 || !((e.fluid_bits[d >> 5] >> (d & 31)) & 1u)) return false;
  return !(e.ns && e.ns[d] > 0.5f);
}
// averaged component value as the reference's acc holds it (acc / nacc)
template <typename S>
__device__ __forceinline__ double U(const EvalDev& e, int comp, int64_t d) {
  const double v = (double)((const S*)e.acc)[comp * e.plane + d] / e.div;
  return sizeof(S) == 4 ? (double)(float)v : v;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o, 64);
    v = (u > v || isnan(u)) ? u : v;
  }
  return v;
}
// block (1024 threads) sum / max via wave shuffles and LDS
__device__ double bsum(double v, double* red) {
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = threadIdx.x < 16 ? red[threadIdx.x] : 0.0;
  if (threadIdx.x < 64) s = wsum(s);
  __syncthreads();
  if (threadIdx.x == 0) red[0] = s;
  __syncthreads();
  return red[0];
}
__device__ double bmax(double v, double* red) {
  v = wmax(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = threadIdx.x < 16 ? red[threadIdx.x] : -INFINITY;
  if (threadIdx.x < 64) s = wmax(s);
  __syncthreads();
  if (threadIdx.x == 0) red[0] = s;
  __syncthreads();
  return red[0];
}

struct EvalOut {
  double sw, n, csum, cn, wm, cm, dev, cdev, low, high, rev, mxr, pin, pinN, pb, pbN, pinf, winf;
};

// Filter map on layer jm and the raw per-cell w = -u_y (NaN outside fluid).
template <typename S>
__global__ void k_layer(EvalDev e, int jm, float* fmap, uint8_t* core) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (int64_t)e.nx * e.nz) return;
  const int i = (int)(x / e.nz), k = (int)(x % e.nz);
  const int64_t d = dpos(e, i, jm, k);
  fmap[x] = efluid(e, d) ? (float)(-U<S>(e, 1, d)) : NAN;
  uint8_t c = 0;
  if (i >= 1 && i < e.nx - 1 && k >= 1 && k < e.nz - 1)
    c = efluid(e, d) && efluid(e, dpos(e, i - 1, jm, k)) && efluid(e, dpos(e, i + 1, jm, k)) &&
        efluid(e, dpos(e, i, jm, k - 1)) && efluid(e, dpos(e, i, jm, k + 1));
  core[x] = c;
}

// Scalar reductions of evaluate and field3d (engine.js:639-660, 704-710) in
// one workgroup.  The filter map is normalised in place.
template <typename S>
__global__ __launch_bounds__(1024) void k_eval_scalars(EvalDev e, int jb, int kB, const uint8_t* opening,
                                                       float* fmap, const uint8_t* core, EvalOut* out) {
  __shared__ double red[16];
  const int64_t L = (int64_t)e.nx * e.nz;
  double sw = 0, n = 0, csum = 0, cn = 0;
  for (int64_t x = threadIdx.x; x < L; x += 1024) {
    const float w = fmap[x];
    if (!isnan(w)) {
      sw += w;
      n += 1;
    }
    if (core[x]) {
      csum += w;
      cn += 1;
    }
  }
  sw = bsum(sw, red);
  n = bsum(n, red);
  csum = bsum(csum, red);
  cn = bsum(cn, red);
  const double wm = sw / n, cm = csum / cn;
  double dev = 0, cdev = 0, low = 0, high = 0, rev = 0, mxr = 0;
  for (int64_t x = threadIdx.x; x < L; x += 1024) {
    const double w = fmap[x];
    if (isnan(w)) continue;
    dev += fabs(w - wm);
    if (w < 0) rev += 1;
    const double r = w / wm;
    mxr = (r > mxr || isnan(r)) ? r : mxr;
    if (core[x]) {
      cdev += fabs(w - cm);
      if (w < 0.5 * cm) low += 1;
      if (w > 1.5 * cm) high += 1;
    }
    fmap[x] = (float)(w / wm);
  }
  dev = bsum(dev, red);
  cdev = bsum(cdev, red);
  low = bsum(low, red);
  high = bsum(high, red);
  rev = bsum(rev, red);
  mxr = bmax(mxr, red);
  // zeta: density over the inlet opening at k = 1 and over row jb
  double pin = 0, pinN = 0, pb = 0, pbN = 0, pinf = 0, winf = 0;
  for (int64_t c = threadIdx.x; c < (int64_t)e.nx * e.ny; c += 1024) {
    if (!opening[c]) continue;
    const int i = (int)(c / e.ny), j = (int)(c % e.ny);
    const int64_t d = dpos(e, i, j, 1);
    if (efluid(e, d)) {
      pin += U<S>(e, 3, d);
      pinN += 1;
    }
    // mass-flow-weighted total pressure at kB (field3d, engine.js:704-710)
    const int64_t dB = dpos(e, i, j, kB);
    if (efluid(e, dB)) {
      const double ux = U<S>(e, 0, dB), uy = U<S>(e, 1, dB), uz = U<S>(e, 2, dB);
      const double w = fmax(uz, 0.0);
      pinf += w * (U<S>(e, 3, dB) / 3 + 0.5 * (ux * ux + uy * uy + uz * uz));
      winf += w;
    }
  }
  for (int64_t x = threadIdx.x; x < L; x += 1024) {
    const int i = (int)(x / e.nz), k = (int)(x % e.nz);
    const int64_t d = dpos(e, i, jb, k);
    if (efluid(e, d)) {
      pb += U<S>(e, 3, d);
      pbN += 1;
    }
  }
  pin = bsum(pin, red);
  pinN = bsum(pinN, red);
  pb = bsum(pb, red);
  pbN = bsum(pbN, red);
  pinf = bsum(pinf, red);
  winf = bsum(winf, red);
  if (threadIdx.x == 0) {
    EvalOut o;
    o.sw = sw;
    o.n = n;
    o.csum = csum;
    o.cn = cn;
    o.wm = wm;
    o.cm = cm;
    o.dev = dev;
    o.cdev = cdev;
    o.low = low;
    o.high = high;
    o.rev = rev;
    o.mxr = mxr;
    o.pin = pin;
    o.pinN = pinN;
    o.pb = pb;
    o.pbN = pbN;
    o.pinf = pinf;
    o.winf = winf;
    *out = o;
  }
}

// Math.hypot as V8 computes it: scale by the largest magnitude, compensated
// sum of squares, square root, rescale.
__device__ __host__ inline double js_hypot3(double a, double b, double c) {
  const double v[3] = {fabs(a), fabs(b), fabs(c)};
  if (isinf(a) || isinf(b) || isinf(c)) return INFINITY;
  if (isnan(a) || isnan(b) || isnan(c)) return NAN;
  double mx = v[0] > v[1] ? v[0] : v[1];
  mx = mx > v[2] ? mx : v[2];
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  for (int i = 0; i < 3; ++i) {
    const double nn = v[i] / mx;
    const double summand = nn * nn - comp;
    const double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return sqrt(sum) * mx;
}

// Speed and in-plane components on a layer (engine.js:666-674) or on the
// longitudinal cut (engine.js:675-680).  mode 0: plane at j; mode 1: cut at i.
template <typename S>
__global__ void k_plane(EvalDev e, int mode, int at, double u0, float* sp, float* a, float* b) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t d;
  int c0, c1;
  if (mode == 0) {
    if (x >= (int64_t)e.nx * e.nz) return;
    const int i = (int)(x / e.nz), k = (int)(x % e.nz);
    d = dpos(e, i, at, k);
    c0 = 0;
    c1 = 2;
  } else {
    if (x >= (int64_t)e.nz * e.ny) return;
    const int k = (int)(x / e.ny), j = (int)(x % e.ny);
    d = dpos(e, at, j, k);
    c0 = 2;
    c1 = 1;
  }
  if (!efluid(e, d)) {
    sp[x] = NAN;
    a[x] = 0;
    b[x] = 0;
    return;
  }
  const double ux = U<S>(e, 0, d) / u0, uy = U<S>(e, 1, d) / u0, uz = U<S>(e, 2, d) / u0;
  const double u3[3] = {ux, uy, uz};
  sp[x] = (float)js_hypot3(ux, uy, uz);
  a[x] = (float)u3[c0];
  b[x] = (float)u3[c1];
}

// field3d per cell (engine.js:711-717), outputs in the reference's flat order.
template <typename S>
__global__ void k_field3d(EvalDev e, int kB, double u0, double q, double rhoRef, double pin, float* spd, float* pst,
                          float* loss) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)e.nx * e.ny * e.nz) return;
  const int k = (int)(c % e.nz);
  const int64_t ij = c / e.nz;
  const int j = (int)(ij % e.ny), i = (int)(ij / e.ny);
  const int64_t d = dpos(e, i, j, k);
  if (!efluid(e, d) || k < kB) {
    spd[c] = NAN;
    pst[c] = NAN;
    loss[c] = NAN;
    return;
  }
  const double ux = U<S>(e, 0, d), uy = U<S>(e, 1, d), uz = U<S>(e, 2, d), r = U<S>(e, 3, d);
  spd[c] = (float)(js_hypot3(ux, uy, uz) / u0);
  pst[c] = (float)((r - rhoRef) / 3 / q);
  loss[c] = (float)((pin - (r / 3 + 0.5 * (ux * ux + uy * uy + uz * uz))) / q);
}

// ---- streamlines (engine.js:721-758), 8 lanes per seed ------------------

// Trilinear velocity at (x, y, z) from the fluid corners (engine.js:724-737),
// by the 8 lanes of a group, lane g taking corner (a, b, c) = (g >> 2, g >> 1
// & 1, g & 1): the corners' row lookups, fluid tests and loads run at once,
// and every lane then adds the corners' products in the reference's order (a,
// then b, then c), so all lanes hold the same sums, those of one thread
// walking the corners.  round_ok: the cell of the rounded position (JS
// Math.round per axis: floor(x) + 1 where x - floor(x) >= 0.5) lies in the
// grid and is fluid, the loop's first test (engine.js:741); that cell is the
// corner whose offsets are the fractions >= 0.5, whose lane has made the same
// test.
template <typename S>
__device__ double sl_vel8(const EvalDev& e, int g, double x, double y, double z, double out[3], bool* round_ok) {
  const double i0 = floor(x), j0 = floor(y), k0 = floor(z), fx = x - i0, fy = y - j0, fz = z - k0;
  const int a = g >> 2, b = (g >> 1) & 1, c = g & 1;
  const double i = i0 + a, j = j0 + b, k = k0 + c;
  bool ok = !(i < 0 || j < 0 || k < 0 || i >= e.nx || j >= e.ny || k >= e.nz);
  double w = 0, p0 = 0, p1 = 0, p2 = 0;
  if (ok) {
    w = (a ? fx : 1 - fx) * (b ? fy : 1 - fy) * (c ? fz : 1 - fz);
    const int64_t d = dpos(e, (int64_t)i, (int64_t)j, (int64_t)k);
    ok = efluid(e, d);
    if (ok) {
      p0 = w * U<S>(e, 0, d);
      p1 = w * U<S>(e, 1, d);
      p2 = w * U<S>(e, 2, d);
    }
  }
  *round_ok = __shfl((int)ok, (fx >= 0.5 ? 4 : 0) | (fy >= 0.5 ? 2 : 0) | (fz >= 0.5 ? 1 : 0), 8) != 0;
  out[0] = out[1] = out[2] = 0;
  double ws = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bool okq = __shfl((int)ok, q, 8) != 0;
    const double wq = __shfl(w, q, 8), a0 = __shfl(p0, q, 8), a1 = __shfl(p1, q, 8), a2 = __shfl(p2, q, 8);
    if (okq) {
      out[0] += a0;
      out[1] += a1;
      out[2] += a2;
      ws += wq;
    }
  }
  return ws;
}

// One streamline per group of 8 lanes (engine.js:721-758); every lane of a
// group follows the line, lane 0 writes its points.
template <typename S>
__global__ __launch_bounds__(64) void k_streamlines(EvalDev e, const double* seeds, int nlines, double u0, double top,
                                                    int maxp, double* P, double* V, int* np_out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = t >> 3, g = t & 7;
  // whole groups only: the shuffles read all 8 lanes of the group
  const bool live = l < nlines;
  double x = 0, y = 0, z = 0;
  if (live) {
    x = seeds[3 * l];
    y = seeds[3 * l + 1];
    z = seeds[3 * l + 2];
  }
  double* Pl = P + (int64_t)l * 3 * maxp;
  double* Vl = V + (int64_t)l * maxp;
  int np = 0;
  const double ds = 0.35;
  double v1[3], v2[3];
  bool rok;
  for (int st = 0; live && st < 6000; ++st) {
    // the rounded cell's test (out of the grid or not fluid: stop) comes from
    // the corner lanes of the interpolation at the same point
    const double ws1 = sl_vel8<S>(e, g, x, y, z, v1, &rok);
    if (!rok) break;
    if (ws1 < 0.3) break;
    const double m1 = js_hypot3(v1[0], v1[1], v1[2]);
    if (m1 < 0.01 * u0) break;
    if (st % 3 == 0 && np < maxp) {
      if (g == 0) {
        Pl[3 * np] = x;
        Pl[3 * np + 1] = y;
        Pl[3 * np + 2] = z;
        Vl[np] = m1 / u0;
      }
      np++;
    }
    const double xm = x + 0.5 * ds * v1[0] / m1, ym = y + 0.5 * ds * v1[1] / m1, zm = z + 0.5 * ds * v1[2] / m1;
    if (sl_vel8<S>(e, g, xm, ym, zm, v2, &rok) < 0.3) break;
    const double m2 = js_hypot3(v2[0], v2[1], v2[2]);
    if (m2 < 0.01 * u0) break;
    x += ds * v2[0] / m2;
    y += ds * v2[1] / m2;
    z += ds * v2[2] / m2;
    if (y < top + 0.5) {
      if (np < maxp) {
        if (g == 0) {
          Pl[3 * np] = x;
          Pl[3 * np + 1] = y;
          Pl[3 * np + 2] = z;
          Vl[np] = m2 / u0;
        }
        np++;
      }
      break;
    }
  }
  if (live && g == 0) np_out[l] = np;
}

inline unsigned nbk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
      set_errorf("hipMalloc of %zu bytes failed in evaluate", bytes);
      p = nullptr;
      return -1;
    }
    return 0;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

int first_ge_f(const float* a, int n, double v, bool or_last) {
  for (int i = 0; i < n; ++i)
    if ((double)a[i] >= v) return i;
  return or_last ? n - 1 : -1;
}

EvalDev eval_dev(abx_sim* s, int b, double div) {
  const SimView v = sim_view(s);
  EvalDev e;
  e.acc = (const char*)v.acc + (size_t)b * 4 * v.plane * v.ssz;
  e.ns = v.ns ? v.ns + (int64_t)b * v.plane : nullptr;
  e.fluid_bits = v.fluid_bits;
  e.plane = v.plane;
  e.nx = v.nx;
  e.ny = v.ny;
  e.nz = v.nz;
  e.rowid = v.rowid;
  e.rowinfo = v.rowinfo;
  e.div = div;
  return e;
}

int check_geom(abx_sim* s, const abx_eval_geom* g, int b) {
  const SimView v = sim_view(s);
  if (!g || g->nx != v.nx || g->ny != v.ny || g->nz != v.nz || !g->opening || !g->xc || !g->dc || !g->zc) {
    set_error("evaluate: geometry does not match the simulation grid");
    return -1;
  }
  if (b < 0 || b >= v.B) {
    set_error("evaluate: variant out of range");
    return -1;
  }
  return 0;
}

}  // namespace abx

using namespace abx;

namespace {

// Section indices of evaluate and field3d (engine.js:657, 662-665, 704).
struct Sections {
  int jb, ic, jc, jlow, kB;
};

int sections(const abx_eval_geom* g, Sections* sc) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const double h = g->h;
  sc->jb = first_ge_f(g->dc, ny, g->pleat + 2 * h, true);
  double si = 0, sj = 0, so = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      if (g->opening[(int64_t)i * ny + j]) {
        si += i;
        sj += j;
        so++;
      }
  auto jsround = [](double x) {
    const double r = floor(x);
    return (x - r >= 0.5) ? r + 1 : r;
  };
  if (!(so > 0)) {
    set_error("evaluate: no inlet opening");
    return -1;
  }
  sc->ic = (int)jsround(si / so);
  sc->jc = (int)jsround(sj / so);
  sc->jlow = first_ge_f(g->dc, ny, g->pleat + 12, true);
  const int kw = first_ge_f(g->zc, nz, g->mn_z, false);
  sc->kB = std::max(2, kw - 2);
  return 0;
}

// The filter map, core mask and scalar sums of evaluate on the handle's
// stream, into device buffers fmap (nx*nz floats), core (nx*nz bytes) and
// out; opening (nx*ny bytes) is copied to dopen first.
int launch_metrics(abx_sim* s, const abx_eval_geom* g, const EvalDev& e, const Sections& sc, float* fmap,
                   uint8_t* core, uint8_t* dopen, EvalOut* out) {
  const SimView v = sim_view(s);
  hipStream_t st = v.stream;
  const int64_t L = (int64_t)g->nx * g->nz;
  ABX_HIP(hipMemcpyAsync(dopen, g->opening, (size_t)g->nx * g->ny, hipMemcpyHostToDevice, st));
  if (v.ssz == 8) {
    hipLaunchKernelGGL(k_layer<double>, dim3(nbk(L, 256)), dim3(256), 0, st, e, g->jm, fmap, core);
    hipLaunchKernelGGL(k_eval_scalars<double>, dim3(1), dim3(1024), 0, st, e, sc.jb, sc.kB, dopen, fmap, core, out);
  } else {
    hipLaunchKernelGGL(k_layer<float>, dim3(nbk(L, 256)), dim3(256), 0, st, e, g->jm, fmap, core);
    hipLaunchKernelGGL(k_eval_scalars<float>, dim3(1), dim3(1024), 0, st, e, sc.jb, sc.kB, dopen, fmap, core, out);
  }
  ABX_HIP(hipGetLastError());
  return 0;
}

// evaluate's metrics from the scalar sums (engine.js:639-660, 682-685).
void fill_metrics(abx_sim* s, const abx_eval_geom* g, const Sections& sc, const EvalOut& o, abx_metrics* m) {
  const SimView v = sim_view(s);
  const double h = g->h, u0 = v.u0;
  m->gamma = 1 - o.dev / (2 * o.n * o.wm);
  m->gammaCore = 1 - o.cdev / (2 * o.cn * o.cm);
  m->lowCore = o.low / o.cn;
  m->highCore = o.high / o.cn;
  m->reverse = o.rev / o.n;
  m->maxToMean = o.mxr;
  m->zeta = ((o.pin / o.pinN - o.pb / o.pbN) / 3) / (0.5 * u0 * u0);
  m->inlet_cm2 = v.Ain * h * h / 100;
  m->filter_cm2 = v.Af * h * h / 100;
  m->ic = sc.ic;
  m->jc = sc.jc;
  m->jlow = sc.jlow;
  m->jb = sc.jb;
  m->kB = sc.kB;
  m->rhoRef = o.pb / o.pbN;
  m->pin = o.pinf / std::max(o.winf, 1e-12);
  m->n = (int)o.n;
  m->cn = (int)o.cn;
}

}  // namespace

extern "C" int abx_evaluate(abx_sim* s, const abx_eval_geom* g, int b, double divisor, abx_metrics* m,
                            float* fmap, float* planInlet, float* planLow, float* lon, float* spd, float* pst,
                            float* loss) {
  if (!s || !m) {
    set_error("abx_evaluate: null argument");
    return -1;
  }
  if (check_geom(s, g, b)) return -1;
  const SimView v = sim_view(s);
  ABX_HIP(hipSetDevice(v.ordinal));
  hipStream_t st = v.stream;
  EvalDev e = eval_dev(s, b, divisor);
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t L = (int64_t)nx * nz, N = (int64_t)nx * ny * nz;
  const double u0 = v.u0;
  Sections sc;
  if (sections(g, &sc)) return -1;
  Dev dfmap, dcore, dout, dopen, dplanes;
  if (dfmap.alloc(4 * (size_t)L) || dcore.alloc((size_t)L) || dout.alloc(sizeof(EvalOut)) ||
      dopen.alloc((size_t)nx * ny) || dplanes.alloc(4 * (size_t)(6 * L + 3 * (int64_t)nz * ny)))
    return -1;
  if (launch_metrics(s, g, e, sc, dfmap.as<float>(), dcore.as<uint8_t>(), dopen.as<uint8_t>(), dout.as<EvalOut>()))
    return -1;
  float* pl = dplanes.as<float>();
  const bool f64 = v.ssz == 8;
  if (f64) {
    hipLaunchKernelGGL(k_plane<double>, dim3(nbk(L, 256)), dim3(256), 0, st, e, 0, sc.jc, u0, pl, pl + L,
                       pl + 2 * L);
    hipLaunchKernelGGL(k_plane<double>, dim3(nbk(L, 256)), dim3(256), 0, st, e, 0, sc.jlow, u0, pl + 3 * L,
                       pl + 4 * L, pl + 5 * L);
    hipLaunchKernelGGL(k_plane<double>, dim3(nbk((int64_t)nz * ny, 256)), dim3(256), 0, st, e, 1, sc.ic, u0,
                       pl + 6 * L, pl + 6 * L + (int64_t)nz * ny, pl + 6 * L + 2 * (int64_t)nz * ny);
  } else {
    hipLaunchKernelGGL(k_plane<float>, dim3(nbk(L, 256)), dim3(256), 0, st, e, 0, sc.jc, u0, pl, pl + L,
                       pl + 2 * L);
    hipLaunchKernelGGL(k_plane<float>, dim3(nbk(L, 256)), dim3(256), 0, st, e, 0, sc.jlow, u0, pl + 3 * L,
                       pl + 4 * L, pl + 5 * L);
    hipLaunchKernelGGL(k_plane<float>, dim3(nbk((int64_t)nz * ny, 256)), dim3(256), 0, st, e, 1, sc.ic, u0,
                       pl + 6 * L, pl + 6 * L + (int64_t)nz * ny, pl + 6 * L + 2 * (int64_t)nz * ny);
  }
  ABX_HIP(hipGetLastError());
  EvalOut o;
  ABX_HIP(hipMemcpyAsync(&o, dout.p, sizeof o, hipMemcpyDeviceToHost, st));
  if (fmap) ABX_HIP(hipMemcpyAsync(fmap, dfmap.p, 4 * (size_t)L, hipMemcpyDeviceToHost, st));
  if (planInlet) ABX_HIP(hipMemcpyAsync(planInlet, pl, 12 * (size_t)L, hipMemcpyDeviceToHost, st));
  if (planLow) ABX_HIP(hipMemcpyAsync(planLow, pl + 3 * L, 12 * (size_t)L, hipMemcpyDeviceToHost, st));
  if (lon) ABX_HIP(hipMemcpyAsync(lon, pl + 6 * L, 12 * (size_t)nz * ny, hipMemcpyDeviceToHost, st));
  ABX_HIP(hipStreamSynchronize(st));
  fill_metrics(s, g, sc, o, m);
  const double q = 0.5 * u0 * u0;
  if (spd || pst || loss) {
    if (!spd || !pst || !loss) {
      set_error("abx_evaluate: spd, pst and loss go together");
      return -1;
    }
    Dev d3;
    if (d3.alloc(12 * (size_t)N)) return -1;
    float* f3 = d3.as<float>();
    if (f64)
      hipLaunchKernelGGL(k_field3d<double>, dim3(nbk(N, 256)), dim3(256), 0, st, e, sc.kB, u0, q, m->rhoRef, m->pin,
                         f3, f3 + N, f3 + 2 * N);
    else
      hipLaunchKernelGGL(k_field3d<float>, dim3(nbk(N, 256)), dim3(256), 0, st, e, sc.kB, u0, q, m->rhoRef, m->pin,
                         f3, f3 + N, f3 + 2 * N);
    ABX_HIP(hipGetLastError());
    ABX_HIP(hipMemcpyAsync(spd, f3, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
    ABX_HIP(hipMemcpyAsync(pst, f3 + N, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
    ABX_HIP(hipMemcpyAsync(loss, f3 + 2 * N, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
    ABX_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

extern "C" int abx_streamlines(abx_sim* s, const abx_eval_geom* g, int b, double divisor, int maxSeeds, float* pts,
                               int64_t pts_cap, float* spd, int64_t spd_cap, uint32_t* off, int64_t off_cap,
                               int64_t* npts_out, int64_t* nlines_out) {
  if (!s || !npts_out || !nlines_out || maxSeeds < 1) {
    set_error("abx_streamlines: bad argument");
    return -1;
  }
  if (check_geom(s, g, b)) return -1;
  const SimView v = sim_view(s);
  ABX_HIP(hipSetDevice(v.ordinal));
  hipStream_t st = v.stream;
  EvalDev e = eval_dev(s, b, divisor);
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  // seeds: 4 per opening cell at kS, a fractional stride down to maxSeeds (engine.js:736-741)
  const int kw = first_ge_f(g->zc, nz, g->mn_z, false);
  const int kS = std::max(2, kw - 3);
  std::vector<double> seeds;
  const double offs[4][2] = {{-0.25, -0.25}, {0.25, -0.25}, {-0.25, 0.25}, {0.25, 0.25}};
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      if (g->opening[(int64_t)i * ny + j])
        for (auto& o : offs) {
          seeds.push_back(i + o[0]);
          seeds.push_back(j + o[1]);
          seeds.push_back(kS);
        }
  const int64_t nseed = (int64_t)seeds.size() / 3;
  const double stride = std::max(1.0, (double)nseed / maxSeeds);
  std::vector<double> sel;
  for (double sd = 0; sd < nseed; sd += stride) {
    const int64_t k = (int64_t)floor(sd);
    sel.insert(sel.end(), &seeds[(size_t)(3 * k)], &seeds[(size_t)(3 * k)] + 3);
  }
  const int nl = (int)(sel.size() / 3);
  const int maxp = 2002;
  *npts_out = 0;
  *nlines_out = 0;
  if (off && off_cap > 0) off[0] = 0;
  if (nl == 0) return 0;
  Dev dseeds, dP, dV, dn;
  if (dseeds.alloc(8 * sel.size()) || dP.alloc(8 * 3 * (size_t)maxp * nl) || dV.alloc(8 * (size_t)maxp * nl) ||
      dn.alloc(4 * (size_t)nl))
    return -1;
  ABX_HIP(hipMemcpyAsync(dseeds.p, sel.data(), 8 * sel.size(), hipMemcpyHostToDevice, st));
  if (v.ssz == 8)
    hipLaunchKernelGGL(k_streamlines<double>, dim3(nbk((int64_t)nl * 8, 64)), dim3(64), 0, st, e, dseeds.as<double>(),
                       nl, v.u0, (double)g->top, maxp, dP.as<double>(), dV.as<double>(), dn.as<int>());
  else
    hipLaunchKernelGGL(k_streamlines<float>, dim3(nbk((int64_t)nl * 8, 64)), dim3(64), 0, st, e, dseeds.as<double>(),
                       nl, v.u0, (double)g->top, maxp, dP.as<double>(), dV.as<double>(), dn.as<int>());
  ABX_HIP(hipGetLastError());
  std::vector<double> P((size_t)3 * maxp * nl), V((size_t)maxp * nl);
  std::vector<int> np((size_t)nl);
  ABX_HIP(hipMemcpyAsync(P.data(), dP.p, 8 * P.size(), hipMemcpyDeviceToHost, st));
  ABX_HIP(hipMemcpyAsync(V.data(), dV.p, 8 * V.size(), hipMemcpyDeviceToHost, st));
  ABX_HIP(hipMemcpyAsync(np.data(), dn.p, 4 * np.size(), hipMemcpyDeviceToHost, st));
  ABX_HIP(hipStreamSynchronize(st));
  // model coordinates (toModel, engine.js:40, 753)
  int64_t npts = 0, nlines = 0;
  const double h = g->h;
  for (int l = 0; l < nl; ++l) {
    if (np[(size_t)l] < 2) continue;
    for (int q = 0; q < np[(size_t)l]; ++q) {
      const double* p = &P[(size_t)(3 * ((int64_t)l * maxp + q))];
      const double X = g->xc[0] + p[0] * h, D = g->dc[0] + p[1] * h, Z = g->zc[0] + p[2] * h;
      double mq[3];
      mq[g->oAx] = g->cy * X - g->sy * Z;
      mq[g->dAx] = g->seal - g->sd * D;
      mq[g->iAx] = -g->si * (g->sy * X + g->cy * Z);
      if (pts && 3 * npts + 2 < pts_cap)
        for (int a = 0; a < 3; ++a) pts[3 * npts + a] = (float)mq[a];
      if (spd && npts < spd_cap) spd[npts] = (float)V[(size_t)((int64_t)l * maxp + q)];
      npts++;
    }
    nlines++;
    if (off && nlines < off_cap) off[nlines] = (uint32_t)npts;
  }
  *npts_out = npts;
  *nlines_out = nlines;
  if ((pts && 3 * npts > pts_cap) || (spd && npts > spd_cap) || (off && nlines + 1 > off_cap)) {
    set_errorf("abx_streamlines: output buffers too small for %lld points on %lld lines", (long long)npts,
               (long long)nlines);
    return -1;
  }
  return 0;
}

// ---- evaluate's metrics without device allocations ------------------------
// The buffers abx_evaluate allocates per call for its metrics (the filter map,
// the core mask, the inlet opening and the scalar sums) live here in one
// fixed device array per GPU.  The host never passes `scratch` itself to a
// kernel or a copy: it takes the array's device address once per device with
// hipGetSymbolAddress and uses that.
constexpr size_t kScratchBytes = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Device address of this GPU's scratch array (the device is current).
int scratch_address(int ordinal, unsigned char** out) {
  if (ordinal < 0 || ordinal >= kMaxDevices) {
    set_errorf("abx_evaluate_metrics_noalloc: device %d out of range", ordinal);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[ordinal]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_errorf("abx_evaluate_metrics_noalloc: scratch address: %s", hipGetErrorString(e));
      return -1;
    }
    g_scratch_ptr[ordinal] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[ordinal];
  return 0;
}

}  // namespace

extern "C" int64_t abx_evaluate_metrics_noalloc_bytes(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const size_t L = (size_t)nx * nz;
  return (int64_t)(align256(sizeof(EvalOut)) + align256(4 * L) + align256(L) + align256((size_t)nx * ny));
}

extern "C" int abx_evaluate_metrics_noalloc(abx_sim* s, const abx_eval_geom* g, int b, double divisor,
                                            abx_metrics* m, float* fmap) {
  if (!s || !m) {
    set_error("abx_evaluate_metrics_noalloc: null argument");
    return -1;
  }
  if (check_geom(s, g, b)) return -1;
  const int64_t need = abx_evaluate_metrics_noalloc_bytes(g->nx, g->ny, g->nz);
  if (need < 0 || (size_t)need > kScratchBytes) {
    set_errorf("abx_evaluate_metrics_noalloc: a %d x %d x %d grid needs %lld bytes of scratch, more than the "
               "%zu bytes available",
               g->nx, g->ny, g->nz, (long long)need, kScratchBytes);
    return -1;
  }
  const SimView v = sim_view(s);
  ABX_HIP(hipSetDevice(v.ordinal));
  unsigned char* base = nullptr;
  if (scratch_address(v.ordinal, &base)) return -1;
  Sections sc;
  if (sections(g, &sc)) return -1;
  const size_t L = (size_t)g->nx * g->nz;
  // every byte used below is written before it is read in this call
  EvalOut* dout = (EvalOut*)base;
  float* dfmap = (float*)(base + align256(sizeof(EvalOut)));
  uint8_t* dcore = (uint8_t*)dfmap + align256(4 * L);
  uint8_t* dopen = dcore + align256(L);
  const EvalDev e = eval_dev(s, b, divisor);
  if (launch_metrics(s, g, e, sc, dfmap, dcore, dopen, dout)) return -1;
  EvalOut o;
  ABX_HIP(hipMemcpyAsync(&o, dout, sizeof o, hipMemcpyDeviceToHost, v.stream));
  if (fmap) ABX_HIP(hipMemcpyAsync(fmap, dfmap, 4 * L, hipMemcpyDeviceToHost, v.stream));
  ABX_HIP(hipStreamSynchronize(v.stream));
  fill_metrics(s, g, sc, o, m);
  return 0;
}
