// D3Q19 stream-collide on gfx950: the device side of the reference's
// setupSim / cpuStep / GpuSim / probe (engine.js:239-321, 378-462, 493-594).
//
// Lattice layout (HBM).  The reference stores the compacted fluid list with a
// 19-entry neighbour table (engine.js:241-251), 76 B of indices per cell that
// every step re-reads.  Here every grid row (i, j) that holds fluid stores
// the span of k from its first to its last fluid cell, and the rows follow
// one another without gaps (in the order i, j), so
// nearly every stored position is a fluid cell and every 128-byte line of a
// population plane is almost entirely fluid: cell (i, j, k) sits at position
// rowoff(i, j) + k of each plane (structure of arrays, f[q*plane + d]).  The
// neighbour x - c_q lies in row (i - cx, j - cy) at position
// d + delta(row, -cx, -cy) - cz, with the nine deltas of each row in a table;
// one meta word per position holds the neighbours' fluid bits and the row
// (relative to its 64-position segment).  A neighbour that is not fluid (a
// wall, or outside the grid) turns into halfway bounce-back f_opp(q)(x) of
// the previous step, which is exactly the reference's "nb = -1" case.
//
// Per cell the even AA step reads only bits: a fluid bit (32 cells per word)
// and three type bits (inlet / outlet / porous, engine.js:252-254); the odd
// step also reads the cell's meta word and its row's deltas.  A wavefront
// owns 64 consecutive positions (a segment); only segments that hold fluid
// are launched (the segment list).  Algorithmic traffic per fluid cell
// update: 19 loads + 19 stores of the storage type (152 B in float).
#include <math.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <cstring>
#include <vector>

#include "abx_common.hpp"
#include "lbm_device.hpp"
#include "sim_internal.hpp"

namespace abx {

// Segment kernel: one position per lane, a workgroup of 256 lanes takes
// four independent active 64-position segments (one per wavefront), so only
// segments holding fluid occupy lanes.
//
// MODE selects the propagation scheme:
//  0  two lattices, pull: read fin[q](x - c_q) (or fin[opp q](x) behind a
//     wall), write fout[q](x);
//  1  AA pattern, even step, one lattice in place: read F[q](x), write the
//     post-collision f_q to F[opp q](x);
//  2  AA pattern, odd step: read F[opp q](x - c_q) (behind a wall: F[q](x)),
//     write f_q to F[q](x + c_q) (towards a wall: F[opp q](x)).
// Both schemes apply the same halfway bounce-back and give identical values;
// the AA steps touch each population slot once for reading and once for
// writing in place, which halves the lattice footprint and streams far more
// efficiently (even steps are fully aligned).  NT: non-temporal population
// loads (bit 0) / stores (bit 1).
template <typename S, typename M, bool ACC, bool NS, int MODE, int NT>
__global__ __launch_bounds__(256) void k_stream_collide_seg(StepArgs a) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int nblk = (a.nchunks + 3) >> 2;
  // Block order across the 8 XCDs (workgroup x runs on XCD x % 8 as the
  // s-th of that XCD's workgroups, s = x / 8):
  //   xcd_map 0: block x (round robin);
  //   xcd_map 1: XCD x % 8 owns the contiguous range [(x % 8) per_xcd, ...);
  //   xcd_map G > 1: groups of G consecutive blocks per XCD, all XCDs moving
  //   through the lattice together (block (s / G) 8G + (x % 8) G + s % G).
  int blk = (int)blockIdx.x;
  if (a.xcd_map == 1) blk = (int)(blockIdx.x & 7) * a.per_xcd + (int)(blockIdx.x >> 3);
  if (a.xcd_map > 1) {
    const int sq = (int)(blockIdx.x >> 3), G = a.xcd_map;
    blk = (sq / G) * 8 * G + (int)(blockIdx.x & 7) * G + sq % G;
  }
  if (blk >= nblk) return;
  const int sg = 4 * blk + wv;
  if (sg < a.nchunks) seg_cell<S, M, ACC, NS, MODE, NT>(a, a.chunks[sg], l, blockIdx.y);
}

// The AA steps' memory pattern alone on a handle's lattice (abx_sim_pattern_time):
// k_stream_collide_seg's loads and stores with the collision left out.
template <int MODE, int NT>
__global__ __launch_bounds__(256) void k_pattern_seg(StepArgs a) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int nblk = (a.nchunks + 3) >> 2;
  int blk = (int)blockIdx.x;
  if (a.xcd_map == 1) blk = (int)(blockIdx.x & 7) * a.per_xcd + (int)(blockIdx.x >> 3);
  if (a.xcd_map > 1) {
    const int sq = (int)(blockIdx.x >> 3), G = a.xcd_map;
    blk = (sq / G) * 8 * G + (int)(blockIdx.x & 7) * G + sq % G;
  }
  if (blk >= nblk) return;
  const int sg = 4 * blk + wv;
  if (sg < a.nchunks) seg_cell<float, float, false, false, MODE, NT, true>(a, a.chunks[sg], l, 0);
}

template <typename S, typename M>
__global__ void k_probe(const S* f, int64_t P, const int32_t* plist, int np, double* vy, LatView L, int form) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const int d = plist[n];
  M r = 0, my = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const M v = (M)post_f(f, P, L, d, q, form);
    r += v;
    my += (M)CY[q] * v;
  }
  const M u = my / r;
  vy[n] = (sizeof(S) == 4) ? (double)(float)u : (double)u;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum over 1024 threads (16 waves): wave shuffles, then LDS.
__device__ double block_sum_1024(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x < 64) {
    s = threadIdx.x < 16 ? red[threadIdx.x] : 0.0;
    s = wave_sum(s);
  }
  __syncthreads();
  if (threadIdx.x == 0) red[0] = s;
  __syncthreads();
  return red[0];
}

// gamma / mass ratio of progressInfo (engine.js:330-336) in one workgroup:
// w = -u_y on layer jm (nM cells) and on layer top+1 (nT cells).
// out = {gamma, massRatio, qo, wm}.
__global__ __launch_bounds__(1024) void k_gamma(const double* vy, int nM, int nT, double ain_u0, double* out) {
  __shared__ double red[16];
  double sw = 0, qo = 0;
  for (int n = threadIdx.x; n < nM; n += 1024) sw += -vy[n];
  for (int n = nM + threadIdx.x; n < nM + nT; n += 1024) qo += -vy[n];
  sw = block_sum_1024(sw, red);
  qo = block_sum_1024(qo, red);
  const double wm = sw / nM;
  double dev = 0;
  for (int n = threadIdx.x; n < nM; n += 1024) dev += fabs(-vy[n] - wm);
  dev = block_sum_1024(dev, red);
  if (threadIdx.x == 0) {
    out[0] = 1 - dev / (2 * (double)nM * wm);
    out[1] = qo / ain_u0;
    out[2] = qo;
    out[3] = wm;
  }
}

// out[q*NF + a] = post-collision f_q of fluid cell a (the reference's gather
// kernel, engine.js:468-476, over whole components), one thread per cell.
template <typename S>
__global__ void k_gather_f(const S* F, int64_t P, const int32_t* cells, int64_t NF, S* out, LatView L, int form) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= NF) return;
  const int d = cells[a];
#pragma unroll
  for (int q = 0; q < 19; ++q) out[q * NF + a] = post_f(F, P, L, d, q, form);
}

// acc / divisor in double, stored in T (engine.js:628: Float32Array /= nacc).
template <typename S, typename T>
__global__ void k_gather_div(const S* src, int64_t P, const int32_t* cells, int64_t NF, double div, T* out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NF * 4) return;
  const int64_t c = n / NF, a = n - c * NF;
  out[n] = (T)((double)src[c * P + cells[a]] / div);
}

// Store post-collision populations given in cid order in the stored form.
template <typename S>
__global__ void k_scatter_f(const S* in, int64_t P, const int32_t* cells, int64_t NF, S* F, LatView L, int form) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= NF) return;
  const int d = cells[a];
#pragma unroll
  for (int q = 0; q < 19; ++q) F[post_slot(P, L, d, q, form)] = in[q * NF + a];
}

// f_q = feq_q(rho, u) at the fluid cells (feq as engine.js:296).
template <typename S, typename M>
__global__ void k_init_eq(const float* rho_u, const int32_t* cells, int64_t NF, int64_t P, S* f, LatView L, int form) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= NF) return;
  const M rho = (M)rho_u[a], ux = (M)rho_u[NF + a], uy = (M)rho_u[2 * NF + a], uz = (M)rho_u[3 * NF + a];
  const M usq = (M)1.5 * (ux * ux + uy * uy + uz * uz);
  const int d = cells[a];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const M cu = (M)3 * ((M)CX[q] * ux + (M)CY[q] * uy + (M)CZ[q] * uz);
    f[post_slot(P, L, d, q, form)] = (S)((M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq));
  }
}

// Resampling of a coarse volume (bf16, channels (ux, uy, uz, rho) mapped to
// lattice units by phys_c = a_c * v_c + b_c) onto the fluid cells, in cid
// order, as (rho, ux, uy, uz) for k_init_eq.  Both lattices share the local
// frame; cell centres are origin + (i + 0.5) h (engine.js:137-139).
// Trilinear weights over the coarse fluid cells only, renormalised; a cell
// with no fluid coarse neighbour gets rest (rho 1, u 0).
struct ResampleArgs {
  const uint16_t* vol;
  const uint8_t* cfluid;  // nxc * nyc * nzc, k fastest
  int cs, VD, VH, VW, nxc, nyc, nzc;
  double oc[3], hc, of[3], hf;
  float a[4], b[4];
};

__global__ void k_resample_rho_u(ResampleArgs r, const int32_t* cells, int64_t NF, LatView L, const int4* rowinfo,
                                 int ny, float* rho_u) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NF) return;
  const int d = cells[e];
  const int4 ri = rowinfo[lat_row(L, d, L.meta[d])];  // {rowoff, kmin, kmax, i ny + j}
  const int k = d - ri.x, i = ri.w / ny, j = ri.w - i * ny;
  const int n3[3] = {r.nxc, r.nyc, r.nzc};
  const int idx[3] = {i, j, k};
  int i0[3];
  double w1[3];
  for (int ax = 0; ax < 3; ++ax) {
    const double pos = r.of[ax] + ((double)idx[ax] + 0.5) * r.hf;
    double c = (pos - r.oc[ax]) / r.hc - 0.5;
    c = fmin(fmax(c, 0.0), (double)(n3[ax] - 1));
    int f0 = (int)floor(c);
    f0 = min(f0, n3[ax] > 1 ? n3[ax] - 2 : 0);
    i0[ax] = f0;
    w1[ax] = c - (double)f0;
  }
  double acc[4] = {0, 0, 0, 0}, ws = 0;
  for (int o = 0; o < 8; ++o) {
    const int ci = i0[0] + (o & 1), cj = i0[1] + ((o >> 1) & 1), ck = i0[2] + (o >> 2);
    if (ci >= r.nxc || cj >= r.nyc || ck >= r.nzc) continue;
    if (!r.cfluid[((int64_t)ci * r.nyc + cj) * r.nzc + ck]) continue;
    const double w = ((o & 1) ? w1[0] : 1.0 - w1[0]) * (((o >> 1) & 1) ? w1[1] : 1.0 - w1[1]) *
                     ((o >> 2) ? w1[2] : 1.0 - w1[2]);
    const uint16_t* v = r.vol + (((int64_t)ci * r.VH + cj) * r.VW + ck) * r.cs;
    for (int c = 0; c < 4; ++c) acc[c] += w * (double)__uint_as_float((uint32_t)v[c] << 16);
    ws += w;
  }
  float rho = 1.f, u[3] = {0.f, 0.f, 0.f};
  if (ws > 0) {
    for (int c = 0; c < 3; ++c) u[c] = (float)((double)r.a[c] * (acc[c] / ws) + (double)r.b[c]);
    rho = (float)((double)r.a[3] * (acc[3] / ws) + (double)r.b[3]);
  }
  rho_u[e] = rho;
  rho_u[NF + e] = u[0];
  rho_u[2 * NF + e] = u[1];
  rho_u[3 * NF + e] = u[2];
}

// Cold start f_q = WT[q] on every position of every plane (f0, engine.js:260-261).
template <typename S>
__global__ void k_fill_wt(S* f, int64_t P, int nplanes) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= P * nplanes) return;
  const int q = (int)((n / P) % 19);
  f[n] = (S)wt(q);
}

template <typename S>
__global__ void k_fill_zero(S* x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (S)0;
}

}  // namespace abx

// ------------------------------------------------------------------------ host
using namespace abx;

struct abx_sim {
  abx_dev* dev = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // default launch: groups of 4 consecutive blocks per XCD, all XCDs sweeping
  // the lattice together
  int prec = ABX_F32, B = 1, xcd_map = 4;
  int nx = 0, ny = 0, nz = 0;
  int nrows = 0;
  int64_t plane = 0, NF = 0;
  size_t ssz = 4;  // bytes per stored value
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  void* acc = nullptr;
  float* ns = nullptr;
  uint32_t* fluid_bits = nullptr;
  uint32_t* type_bits = nullptr;
  // the row-span layout (LatView): per-position meta words, the first row of
  // every segment, per-row neighbour deltas; per-row {rowoff, kmin, kmax,
  // i ny + j} and the row of every (i, j) (-1 without fluid)
  uint32_t* meta = nullptr;
  int32_t* seg_row0 = nullptr;
  int32_t* rdelta = nullptr;
  int4* rowinfo = nullptr;
  int32_t* rowid = nullptr;
  int32_t* chunks64 = nullptr;    // active 64-position segments
  int nchunks64 = 0;
  std::vector<int32_t> seg_host;  // host copy of chunks64 (ascending positions)
  // kernel 4: AA pattern in one lattice (the default); 2: two lattices, pull.
  // nt: non-temporal accesses of the AA steps, seg_cell's NT flags of the even
  // step in bits 0-7 and of the odd step in bits 8-15.
  int kernel = 4, nt = 19;
  // waves per CU of the AA even and odd steps (0: as many as registers
  // allow).  The even step, a pure stream of 38 accesses per lane, runs
  // faster with fewer requests in flight: 16 waves (4 workgroups) per CU.
  int occ_even = 16, occ_odd = 0;
  // stored form of the populations: 0 plain (f[cur][q](x) = f_q), 1 AA
  // swapped (F[opp q](x) = f_q), 2 AA streamed (after an odd AA step)
  int form = 1;
  int32_t* cells = nullptr;
  int32_t* plist = nullptr;
  double* vy = nullptr;
  double* red = nullptr;
  double* red_host = nullptr;
  int nM = 0, nT = 0, Ain = 0, Af = 0, Lp = 0;
  double u0 = 0, tau0 = 0, sig = 0, aniso = 0, cs2s = 0;
  int64_t device_bytes = 0;
  int use_graph = 0;
  // captured chunks of steps: (n, doacc, form, cur, kernel, nt, xcd_map) -> executable graph
  std::map<std::tuple<int, int, int, int, int, int, int>, hipGraphExec_t> graphs;
};

namespace {

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
int dalloc(abx_sim* s, T** p, size_t count) {
  size_t bytes = std::max<size_t>(count * sizeof(T), 256);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    set_errorf("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    *p = nullptr;
    return -1;
  }
  s->device_bytes += (int64_t)bytes;
  return 0;
}

int ensure_second_lattice(abx_sim* s);

LatView lat_view(const abx_sim* s) {
  LatView L;
  L.meta = s->meta;
  L.seg_row0 = s->seg_row0;
  L.rdelta = s->rdelta;
  return L;
}

StepArgs step_args(abx_sim* s, int doacc) {
  StepArgs a;
  a.fin = s->f[s->cur];
  a.fout = s->kernel == 4 ? s->f[s->cur] : s->f[s->cur ^ 1];
  a.acc = doacc ? s->acc : nullptr;
  a.ns = s->ns;
  a.fluid_bits = s->fluid_bits;
  a.type_bits = s->type_bits;
  a.chunks = s->chunks64;
  a.nchunks = s->nchunks64;
  a.xcd_map = s->xcd_map;
  a.per_xcd = (a.nchunks + 7) / 8;
  a.lat = lat_view(s);
  a.plane = s->plane;
  a.u0 = s->u0;
  a.tau0 = s->tau0;
  a.sig = s->sig;
  a.aniso = s->aniso;
  a.cs2s = s->cs2s;
  return a;
}

// Dynamic LDS that caps a launch of `waves_per_wg`-wave workgroups at
// `waves` waves per CU (0: no cap): one more workgroup than allowed can
// never fit beside the others.  The stream-collide itself uses no LDS.
size_t occupancy_lds(const abx_dev* d, int waves, int waves_per_wg) {
  if (waves <= 0) return 0;
  const int wgs = std::max(1, waves / waves_per_wg);
  return (size_t)(d->lds_per_cu / (wgs + 1) + 1024);
}

// Grid of the segment kernels for a's block order (sets a.per_xcd).
unsigned seg_grid_x(StepArgs& a) {
  const int nblk = (a.nchunks + 3) / 4;
  a.per_xcd = (nblk + 7) / 8;
  unsigned gx = a.xcd_map == 1 ? (unsigned)(8 * a.per_xcd) : (unsigned)nblk;
  if (a.xcd_map > 1) gx = (unsigned)((nblk + 8 * a.xcd_map - 1) / (8 * a.xcd_map) * (8 * a.xcd_map));
  return gx;
}

template <typename S, typename M, int MODE, int NT>
void launch_seg(abx_sim* s, StepArgs a, int doacc) {
  dim3 grid(seg_grid_x(a), (unsigned)s->B), block(256);
  const bool ns = a.ns != nullptr;
  const size_t lds = occupancy_lds(s->dev, MODE == 1 ? s->occ_even : (MODE == 2 ? s->occ_odd : 0), 4);
  if (doacc) {
    if (ns)
      hipLaunchKernelGGL((k_stream_collide_seg<S, M, true, true, MODE, NT>), grid, block, lds, s->stream, a);
    else
      hipLaunchKernelGGL((k_stream_collide_seg<S, M, true, false, MODE, NT>), grid, block, lds, s->stream, a);
  } else {
    if (ns)
      hipLaunchKernelGGL((k_stream_collide_seg<S, M, false, true, MODE, NT>), grid, block, lds, s->stream, a);
    else
      hipLaunchKernelGGL((k_stream_collide_seg<S, M, false, false, MODE, NT>), grid, block, lds, s->stream, a);
  }
}

// Non-temporal flag sets of the AA steps (seg_cell's NT) that are compiled:
// the even step (mode 1) takes the first list, the odd step (mode 2) the
// second.  Double storage always uses cached accesses.
constexpr int kNtEven[] = {0, 3, 19};
constexpr int kNtOdd[] = {0, 1, 2, 3, 8, 11, 16, 19, 32, 48};

bool nt_supported(int nt) {
  const int ev = nt & 0xff, od = nt >> 8;
  bool e = false, o = false;
  for (int v : kNtEven) e |= v == ev;
  for (int v : kNtOdd) o |= v == od;
  return e && o && nt >= 0 && nt < (1 << 16);
}

template <typename S, typename M, int MODE>
void launch_aa(abx_sim* s, const StepArgs& a, int doacc, int nt) {
  if constexpr (sizeof(S) == 4) {
    if constexpr (MODE == 1) {
      switch (nt) {
        case 3:
          return launch_seg<S, M, MODE, 3>(s, a, doacc);
        case 19:
          return launch_seg<S, M, MODE, 19>(s, a, doacc);
        default:
          break;
      }
    } else {
      switch (nt) {
        case 1:
          return launch_seg<S, M, MODE, 1>(s, a, doacc);
        case 2:
          return launch_seg<S, M, MODE, 2>(s, a, doacc);
        case 3:
          return launch_seg<S, M, MODE, 3>(s, a, doacc);
        case 8:
          return launch_seg<S, M, MODE, 8>(s, a, doacc);
        case 11:
          return launch_seg<S, M, MODE, 11>(s, a, doacc);
        case 16:
          return launch_seg<S, M, MODE, 16>(s, a, doacc);
        case 19:
          return launch_seg<S, M, MODE, 19>(s, a, doacc);
        case 32:
          return launch_seg<S, M, MODE, 32>(s, a, doacc);
        case 48:
          return launch_seg<S, M, MODE, 48>(s, a, doacc);
        default:
          break;
      }
    }
  }
  launch_seg<S, M, MODE, 0>(s, a, doacc);
}

template <typename S, typename M>
void launch_step_t(abx_sim* s, const StepArgs& a, int doacc) {
  if (s->kernel == 4) {
    // AA pattern: an odd step follows the swapped form, an even one the streamed form
    if (s->form == 2)
      launch_aa<S, M, 1>(s, a, doacc, s->nt & 0xff);
    else
      launch_aa<S, M, 2>(s, a, doacc, s->nt >> 8);
  } else {
    launch_seg<S, M, 0, 0>(s, a, doacc);
  }
}

int launch_step(abx_sim* s, int doacc) {
  if (s->nchunks64 == 0) return 0;
  if (s->kernel != 4 && ensure_second_lattice(s)) return -1;
  StepArgs a = step_args(s, doacc);
  switch (s->prec) {
    case ABX_F32:
      launch_step_t<float, float>(s, a, doacc);
      break;
    case ABX_F32_F64:
      launch_step_t<float, double>(s, a, doacc);
      break;
    default:
      launch_step_t<double, double>(s, a, doacc);
  }
  ABX_HIP(hipGetLastError());
  if (s->kernel == 4)
    s->form = s->form == 2 ? 1 : 2;
  else
    s->cur ^= 1;
  return 0;
}

int launch_steps(abx_sim* s, int n, int doacc) {
  for (int i = 0; i < n; ++i)
    if (launch_step(s, doacc)) return -1;
  return 0;
}

int fill_cold(abx_sim* s) {
  const int64_t n = s->plane * 19 * s->B;
  for (int l = 0; l < 2; ++l) {
    if (!s->f[l]) continue;
    if (s->prec == ABX_F64)
      hipLaunchKernelGGL(k_fill_wt<double>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (double*)s->f[l],
                         s->plane, (int)(19 * s->B));
    else
      hipLaunchKernelGGL(k_fill_wt<float>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (float*)s->f[l],
                         s->plane, (int)(19 * s->B));
  }
  ABX_HIP(hipGetLastError());
  // WT is symmetric under q -> opp q, so the fill is valid in either form
  s->form = s->kernel == 4 ? 1 : 0;
  s->cur = 0;
  return 0;
}

// The second lattice, needed only by the two-lattice kernels.
int ensure_second_lattice(abx_sim* s) {
  if (s->f[1]) return 0;
  const size_t fcount = (size_t)(s->plane * 19 * s->B);
  if (s->prec == ABX_F64) return dalloc(s, (double**)&s->f[1], fcount);
  return dalloc(s, (float**)&s->f[1], fcount);
}

int zero_acc(abx_sim* s) {
  const int64_t n = s->plane * 4 * s->B;
  if (s->prec == ABX_F64)
    hipLaunchKernelGGL(k_fill_zero<double>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (double*)s->acc, n);
  else
    hipLaunchKernelGGL(k_fill_zero<float>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (float*)s->acc, n);
  ABX_HIP(hipGetLastError());
  return 0;
}

// Dense ns [B][plane] from B*NF values in cid order (NULL: all zero).
int upload_ns(abx_sim* s, const float* ns_cid, const std::vector<int32_t>& cells_host) {
  if (!ns_cid) {
    if (s->ns) {
      (void)hipFree(s->ns);
      s->device_bytes -= (int64_t)(s->plane * s->B * sizeof(float));
      s->ns = nullptr;
    }
    return 0;
  }
  if (!s->ns && dalloc(s, &s->ns, (size_t)(s->plane * s->B))) return -1;
  std::vector<float> dense((size_t)(s->plane * s->B), 0.0f);
  for (int b = 0; b < s->B; ++b)
    for (int64_t a = 0; a < s->NF; ++a) dense[(size_t)(b * s->plane + cells_host[(size_t)a])] = ns_cid[b * s->NF + a];
  ABX_HIP(hipMemcpyAsync(s->ns, dense.data(), dense.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
  ABX_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

void sim_drop_graphs(abx_sim* s) {
  for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
  s->graphs.clear();
}

void sim_free(abx_sim* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  sim_drop_graphs(s);
  void* ptrs[] = {s->f[0],  s->f[1],     s->acc,      s->ns,    s->fluid_bits, s->type_bits, s->meta,
                  s->seg_row0, s->rdelta, s->rowinfo, s->rowid, s->chunks64,  s->cells,     s->plist,
                  s->vy,     s->red};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->red_host) (void)hipHostFree(s->red_host);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // namespace

extern "C" int abx_sim_create(abx_dev* dev, const abx_grid* g, const abx_run_params* rp, int precision, int B,
                              const float* ns, abx_sim** out) {
  if (!dev || !g || !rp || !out) {
    set_error("abx_sim_create: null argument");
    return -1;
  }
  *out = nullptr;
  if (precision < ABX_F32 || precision > ABX_F64) {
    set_errorf("abx_sim_create: unknown precision %d", precision);
    return -1;
  }
  if (B < 1 || B > 64) {
    set_errorf("abx_sim_create: B = %d variants (1..64 supported)", B);
    return -1;
  }
  if (g->nx < 1 || g->ny < 1 || g->nz < 1 || !g->fluid || !g->opening || !g->outlet || !g->dc) {
    set_error("abx_sim_create: incomplete grid");
    return -1;
  }
  if (g->top < 0 || g->top >= g->ny || g->jm < 0 || g->jm >= g->ny) {
    set_error("abx_sim_create: top/jm outside the grid");
    return -1;
  }
  ABX_HIP(hipSetDevice(dev->ordinal));
  abx_sim* s = new abx_sim();
  s->dev = dev;
  s->prec = precision;
  s->B = B;
  s->ssz = precision == ABX_F64 ? 8 : 4;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  // rows (i, j) that hold fluid, and the span [kmin, kmax] of each
  const int64_t N = (int64_t)nx * ny * nz, NR = (int64_t)nx * ny;
  std::vector<int> kmin((size_t)NR, -1), kmax((size_t)NR, -1);
  for (int64_t ij = 0; ij < NR; ++ij) {
    const uint8_t* fr = g->fluid + ij * nz;
    for (int k = 0; k < nz; ++k)
      if (fr[k]) {
        if (kmin[(size_t)ij] < 0) kmin[(size_t)ij] = k;
        kmax[(size_t)ij] = k;
      }
  }
  // rows stored in the order i, j without gaps
  std::vector<int32_t> rowid((size_t)NR, -1);
  std::vector<int4> rows;  // {rowoff, kmin, kmax, i ny + j}: cell k of the row at position rowoff + k
  int64_t pos = 0;
  for (int64_t ij = 0; ij < NR; ++ij) {
    if (kmin[(size_t)ij] < 0) continue;
    rowid[(size_t)ij] = (int32_t)rows.size();
    rows.push_back(make_int4((int)(pos - kmin[(size_t)ij]), kmin[(size_t)ij], kmax[(size_t)ij], (int)ij));
    pos += kmax[(size_t)ij] - kmin[(size_t)ij] + 1;
  }
  // at least one segment without fluid after the last row
  pos = (pos + 63) / 64 * 64;
  s->plane = ((pos + 64 + CHUNK - 1) / CHUNK) * CHUNK;
  if ((int64_t)s->plane * 19 * B > ((int64_t)1 << 40) || s->plane > INT32_MAX / 2 || NR > INT32_MAX) {
    set_error("abx_sim_create: grid too large for 32-bit lattice positions");
    delete s;
    return -1;
  }
  s->nrows = (int)rows.size();
  const int64_t nseg = s->plane / 64;
  std::vector<int32_t> seg_row0((size_t)nseg, 0);
  {
    std::vector<uint8_t> set((size_t)nseg, 0);
    for (size_t r = 0; r < rows.size(); ++r)
      for (int64_t sg = ((int64_t)rows[r].x + rows[r].y) >> 6; sg <= ((int64_t)rows[r].x + rows[r].z) >> 6; ++sg)
        if (!set[(size_t)sg]) {
          set[(size_t)sg] = 1;
          seg_row0[(size_t)sg] = (int32_t)r;
        }
  }
  // meta words: the row of every stored position relative to its segment's
  // first row (at most 63: a segment holds at most 64 rows)
  std::vector<uint32_t> meta((size_t)s->plane, 0u);
  for (size_t r = 0; r < rows.size(); ++r)
    for (int k = rows[r].y; k <= rows[r].z; ++k) {
      const int64_t p = (int64_t)rows[r].x + k;
      meta[(size_t)p] = (uint32_t)((int64_t)r - seg_row0[(size_t)(p >> 6)]) << 19;
    }
  // position deltas to the nine neighbour rows
  std::vector<int32_t> rdelta(12 * rows.size(), 0);
  for (size_t r = 0; r < rows.size(); ++r) {
    const int i = rows[r].w / ny, j = rows[r].w % ny;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj) {
        const int ni = i + di, nj = j + dj;
        if (ni < 0 || ni >= nx || nj < 0 || nj >= ny) continue;
        const int32_t r2 = rowid[(size_t)((int64_t)ni * ny + nj)];
        if (r2 >= 0) rdelta[12 * r + (size_t)((di + 1) * 3 + (dj + 1))] = rows[(size_t)r2].x - rows[r].x;
      }
  }
  auto fluid_at = [&](int64_t i, int64_t j, int64_t k) {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz && g->fluid[(i * ny + j) * nz + k];
  };

  // setupSim (engine.js:239-264): compact order, flags, Lp, Ain, layers, sigma
  const float* dc = g->dc;
  const int top = g->top;
  int Lp = 0;
  for (int j = 0; j < ny; ++j)
    if (dc[j] > dc[top] && (double)dc[j] <= g->pleat) Lp++;
  std::vector<int32_t> cells;
  std::vector<int32_t> lM, lT;
  const int64_t nwords = s->plane / 32;
  std::vector<uint32_t> typ((size_t)(3 * nwords), 0u);
  std::vector<uint32_t> fbits((size_t)nwords, 0u);
  int Ain = 0;
  for (int64_t c = 0; c < N; ++c) {
    if (!g->fluid[c]) continue;
    const int64_t k = c % nz, ij = c / nz, j = ij % ny, i = ij / ny;
    const int64_t d = (int64_t)rows[(size_t)rowid[(size_t)ij]].x + k;
    if ((int64_t)cells.size() >= INT32_MAX) {
      set_error("abx_sim_create: too many fluid cells");
      delete s;
      return -1;
    }
    const int32_t a = (int32_t)cells.size();
    cells.push_back((int32_t)d);
    fbits[(size_t)(d >> 5)] |= 1u << (d & 31);
    for (int q = 0; q < 19; ++q)
      if (fluid_at(i - CX[q], j - CY[q], k - CZ[q])) meta[(size_t)d] |= 1u << q;
    int t = -1;
    if (k == 0 && g->opening[ij]) {
      t = 0;
      Ain++;
    } else if (g->outlet[c]) {
      t = 1;
    } else if (dc[j] > dc[top] && (double)dc[j] <= g->pleat) {
      t = 2;
    }
    if (t >= 0) typ[(size_t)(3 * (d >> 5) + t)] |= 1u << (d & 31);
    if (j == g->jm) lM.push_back(a);
    if (j == top + 1) lT.push_back(a);
  }
  s->NF = (int64_t)cells.size();
  s->Ain = Ain;
  s->Lp = Lp;
  s->nM = (int)lM.size();
  s->nT = (int)lT.size();
  s->Af = s->nM;
  s->u0 = rp->u;
  s->tau0 = rp->tau;
  s->aniso = rp->aniso;
  const double uface = s->u0 * Ain / s->Af;
  s->sig = rp->K * 0.5 * s->u0 * s->u0 / (uface * std::max(Lp, 1));
  s->cs2s = 18 * 1.4142135623730951 * 0.17 * 0.17;

  // the segments that hold fluid
  std::vector<int32_t> chunks64;
  for (int64_t sg = 0; sg < nseg; ++sg)
    if (fbits[(size_t)(2 * sg)] | fbits[(size_t)(2 * sg + 1)]) chunks64.push_back((int32_t)(sg * 64));
  s->nchunks64 = (int)chunks64.size();
  s->seg_host = chunks64;
  std::vector<int32_t> plist((size_t)(s->nM + s->nT));
  for (int n = 0; n < s->nM; ++n) plist[(size_t)n] = cells[(size_t)lM[(size_t)n]];
  for (int n = 0; n < s->nT; ++n) plist[(size_t)(s->nM + n)] = cells[(size_t)lT[(size_t)n]];

  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
    set_error("abx_sim_create: stream/event creation failed");
    sim_free(s);
    return -1;
  }
  const size_t fcount = (size_t)(s->plane * 19 * B);
  int rc = 0;
  if (precision == ABX_F64) {
    rc |= dalloc(s, (double**)&s->f[0], fcount);
    rc |= rc ? 0 : dalloc(s, (double**)&s->acc, (size_t)(s->plane * 4 * B));
  } else {
    rc |= dalloc(s, (float**)&s->f[0], fcount);
    rc |= rc ? 0 : dalloc(s, (float**)&s->acc, (size_t)(s->plane * 4 * B));
  }
  rc |= rc ? 0 : dalloc(s, &s->fluid_bits, fbits.size());
  rc |= rc ? 0 : dalloc(s, &s->type_bits, typ.size());
  rc |= rc ? 0 : dalloc(s, &s->meta, meta.size());
  rc |= rc ? 0 : dalloc(s, &s->seg_row0, seg_row0.size());
  rc |= rc ? 0 : dalloc(s, &s->rdelta, std::max<size_t>(rdelta.size(), 12));
  rc |= rc ? 0 : dalloc(s, &s->rowinfo, std::max<size_t>(rows.size(), 1));
  rc |= rc ? 0 : dalloc(s, &s->rowid, rowid.size());
  rc |= rc ? 0 : dalloc(s, &s->chunks64, std::max<size_t>(chunks64.size(), 1));
  rc |= rc ? 0 : dalloc(s, &s->cells, std::max<size_t>(cells.size(), 1));
  rc |= rc ? 0 : dalloc(s, &s->plist, std::max<size_t>(plist.size(), 1));
  rc |= rc ? 0 : dalloc(s, &s->vy, std::max<size_t>(plist.size(), 1));
  rc |= rc ? 0 : dalloc(s, &s->red, 8);
  if (rc) {
    sim_free(s);
    return -1;
  }
  if (hipHostMalloc((void**)&s->red_host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    set_error("abx_sim_create: pinned host buffer");
    sim_free(s);
    return -1;
  }
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream) : hipSuccess;
  };
  if (up(s->fluid_bits, fbits.data(), fbits.size() * 4) != hipSuccess ||
      up(s->type_bits, typ.data(), typ.size() * 4) != hipSuccess || up(s->meta, meta.data(), meta.size() * 4) ||
      up(s->seg_row0, seg_row0.data(), seg_row0.size() * 4) || up(s->rdelta, rdelta.data(), rdelta.size() * 4) ||
      up(s->rowinfo, rows.data(), rows.size() * sizeof(int4)) || up(s->rowid, rowid.data(), rowid.size() * 4) ||
      up(s->chunks64, chunks64.data(), chunks64.size() * 4) || up(s->cells, cells.data(), cells.size() * 4) ||
      up(s->plist, plist.data(), plist.size() * 4)) {
    set_error("abx_sim_create: upload failed");
    sim_free(s);
    return -1;
  }
  if (fill_cold(s) || zero_acc(s) || upload_ns(s, ns, cells)) {
    sim_free(s);
    return -1;
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess) {
    set_error("abx_sim_create: initialisation failed");
    sim_free(s);
    return -1;
  }
  *out = s;
  return 0;
}

extern "C" int abx_sim_info_get(const abx_sim* s, abx_sim_info* o) {
  if (!s || !o) {
    set_error("abx_sim_info_get: null argument");
    return -1;
  }
  o->NF = s->NF;
  o->Ain = s->Ain;
  o->Af = s->Af;
  o->Lp = s->Lp;
  o->nM = s->nM;
  o->nT = s->nT;
  o->u0 = s->u0;
  o->tau0 = s->tau0;
  o->sig = s->sig;
  o->aniso = s->aniso;
  o->cs2s = s->cs2s;
  o->precision = s->prec;
  o->B = s->B;
  o->device_bytes = s->device_bytes;
  o->plane = s->plane;
  o->nrows = s->nrows;
  o->nseg = s->nchunks64;
  return 0;
}

extern "C" int abx_sim_steps(abx_sim* s, int n, int doacc) {
  if (!s || n < 0) {
    set_error("abx_sim_steps: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  doacc = doacc ? 1 : 0;
  if (!s->use_graph || n < 2) return launch_steps(s, n, doacc);
  // a chunk of n steps depends only on (n, doacc, stored form, current
  // lattice, kernel): capture it once, replay it afterwards
  if (s->kernel != 4 && ensure_second_lattice(s)) return -1;
  const int form0 = s->form, cur0 = s->cur;
  const auto key = std::make_tuple(n, doacc, form0, cur0, s->kernel, s->nt, s->xcd_map);
  auto it = s->graphs.find(key);
  if (it == s->graphs.end()) {
    hipGraph_t g;
    ABX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = launch_steps(s, n, doacc);
    const hipError_t e = hipStreamEndCapture(s->stream, &g);
    s->form = form0;
    s->cur = cur0;
    if (rc || e != hipSuccess) {
      if (!rc) set_errorf("abx_sim_steps: graph capture failed: %s", hipGetErrorString(e));
      return -1;
    }
    hipGraphExec_t ex;
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      set_errorf("abx_sim_steps: graph instantiation failed: %s", hipGetErrorString(e2));
      return -1;
    }
    it = s->graphs.emplace(key, ex).first;
  }
  ABX_HIP(hipGraphLaunch(it->second, s->stream));
  if (n % 2) {
    if (s->kernel == 4)
      s->form = form0 == 2 ? 1 : 2;
    else
      s->cur = cur0 ^ 1;
  }
  return 0;
}

extern "C" int abx_sim_set_graph(abx_sim* s, int enable) {
  if (!s) {
    set_error("abx_sim_set_graph: null argument");
    return -1;
  }
  s->use_graph = enable ? 1 : 0;
  return 0;
}

extern "C" int abx_sim_time_steps(abx_sim* s, int n, int doacc, double* ms_total) {
  if (!s || n < 0 || !ms_total) {
    set_error("abx_sim_time_steps: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  ABX_HIP(hipEventRecord(s->ev0, s->stream));
  if (launch_steps(s, n, doacc ? 1 : 0)) return -1;
  ABX_HIP(hipEventRecord(s->ev1, s->stream));
  ABX_HIP(hipEventSynchronize(s->ev1));
  float ms = 0;
  ABX_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  *ms_total = ms;
  return 0;
}

extern "C" int abx_sim_probe(abx_sim* s, int step, double out[3]) {
  if (!s || !out) {
    set_error("abx_sim_probe: null argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  const int np = s->nM + s->nT;
  if (np > 0) {
    if (s->prec == ABX_F32)
      hipLaunchKernelGGL((k_probe<float, float>), dim3(nblk(np, 256)), dim3(256), 0, s->stream,
                         (const float*)s->f[s->cur], s->plane, s->plist, np, s->vy, lat_view(s), s->form);
    else if (s->prec == ABX_F32_F64)
      hipLaunchKernelGGL((k_probe<float, double>), dim3(nblk(np, 256)), dim3(256), 0, s->stream,
                         (const float*)s->f[s->cur], s->plane, s->plist, np, s->vy, lat_view(s), s->form);
    else
      hipLaunchKernelGGL((k_probe<double, double>), dim3(nblk(np, 256)), dim3(256), 0, s->stream,
                         (const double*)s->f[s->cur], s->plane, s->plist, np, s->vy, lat_view(s), s->form);
    hipLaunchKernelGGL(k_gamma, dim3(1), dim3(1024), 0, s->stream, s->vy, s->nM, s->nT, (double)s->Ain * s->u0,
                       s->red);
    ABX_HIP(hipGetLastError());
    ABX_HIP(hipMemcpyAsync(s->red_host, s->red, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    ABX_HIP(hipStreamSynchronize(s->stream));
  } else {
    // empty layers: the reference's sums are 0 and 0/0 gives NaN
    s->red_host[0] = NAN;
    s->red_host[1] = 0.0 / ((double)s->Ain * s->u0);
    s->red_host[2] = 0.0;
  }
  out[0] = s->red_host[0];
  out[1] = s->red_host[1];
  out[2] = s->red_host[2];
  if (!std::isfinite(s->red_host[2])) {
    set_errorf("Rechnung instabil bei Schritt %d. Zellgröße vergrößern oder Viskosität (tau) erhöhen.", step);
    return -1;
  }
  return 0;
}

template <typename T>
static int read_acc_t(abx_sim* s, int b, double divisor, T* out) {
  if (!s || !out || b < 0 || b >= s->B) {
    set_error("abx_sim_read_acc: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  const int64_t n = 4 * s->NF;
  if (n == 0) return 0;
  T* dout = nullptr;
  ABX_HIP(hipMalloc((void**)&dout, (size_t)n * sizeof(T)));
  if (s->prec == ABX_F64)
    hipLaunchKernelGGL((k_gather_div<double, T>), dim3(nblk(n, 256)), dim3(256), 0, s->stream,
                       (const double*)s->acc + (int64_t)b * 4 * s->plane, s->plane, s->cells, s->NF, divisor, dout);
  else
    hipLaunchKernelGGL((k_gather_div<float, T>), dim3(nblk(n, 256)), dim3(256), 0, s->stream,
                       (const float*)s->acc + (int64_t)b * 4 * s->plane, s->plane, s->cells, s->NF, divisor, dout);
  ABX_HIP(hipGetLastError());
  ABX_HIP(hipMemcpyAsync(out, dout, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, s->stream));
  ABX_HIP(hipStreamSynchronize(s->stream));
  ABX_HIP(hipFree(dout));
  return 0;
}

extern "C" int abx_sim_read_acc(abx_sim* s, int b, double divisor, float* out) {
  return read_acc_t<float>(s, b, divisor, out);
}
extern "C" int abx_sim_read_acc_f64(abx_sim* s, int b, double divisor, double* out) {
  return read_acc_t<double>(s, b, divisor, out);
}

extern "C" int abx_sim_read_f(abx_sim* s, int b, void* out) {
  if (!s || !out || b < 0 || b >= s->B) {
    set_error("abx_sim_read_f: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  const int64_t n = 19 * s->NF;
  if (n == 0) return 0;
  void* dout = nullptr;
  ABX_HIP(hipMalloc(&dout, (size_t)n * s->ssz));
  if (s->prec == ABX_F64)
    hipLaunchKernelGGL(k_gather_f<double>, dim3(nblk(s->NF, 256)), dim3(256), 0, s->stream,
                       (const double*)s->f[s->cur] + (int64_t)b * 19 * s->plane, s->plane, s->cells, s->NF,
                       (double*)dout, lat_view(s), s->form);
  else
    hipLaunchKernelGGL(k_gather_f<float>, dim3(nblk(s->NF, 256)), dim3(256), 0, s->stream,
                       (const float*)s->f[s->cur] + (int64_t)b * 19 * s->plane, s->plane, s->cells, s->NF,
                       (float*)dout, lat_view(s), s->form);
  ABX_HIP(hipGetLastError());
  ABX_HIP(hipMemcpyAsync(out, dout, (size_t)n * s->ssz, hipMemcpyDeviceToHost, s->stream));
  ABX_HIP(hipStreamSynchronize(s->stream));
  ABX_HIP(hipFree(dout));
  return 0;
}

extern "C" int abx_sim_write_f(abx_sim* s, int b, const void* in) {
  if (!s || !in || b >= s->B) {
    set_error("abx_sim_write_f: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  const int64_t n = 19 * s->NF;
  if (n == 0) return 0;
  void* din = nullptr;
  ABX_HIP(hipMalloc(&din, (size_t)n * s->ssz));
  ABX_HIP(hipMemcpyAsync(din, in, (size_t)n * s->ssz, hipMemcpyHostToDevice, s->stream));
  for (int v = (b < 0 ? 0 : b); v < (b < 0 ? s->B : b + 1); ++v) {
    if (s->prec == ABX_F64)
      hipLaunchKernelGGL(k_scatter_f<double>, dim3(nblk(s->NF, 256)), dim3(256), 0, s->stream, (const double*)din,
                         s->plane, s->cells, s->NF, (double*)s->f[s->cur] + (int64_t)v * 19 * s->plane,
                         lat_view(s), s->form);
    else
      hipLaunchKernelGGL(k_scatter_f<float>, dim3(nblk(s->NF, 256)), dim3(256), 0, s->stream, (const float*)din,
                         s->plane, s->cells, s->NF, (float*)s->f[s->cur] + (int64_t)v * 19 * s->plane,
                         lat_view(s), s->form);
  }
  ABX_HIP(hipGetLastError());
  ABX_HIP(hipStreamSynchronize(s->stream));
  ABX_HIP(hipFree(din));
  return 0;
}

namespace {
// Equilibrium of (rho, u) in cid order, already on the device, into every variant.
int sim_init_eq_dev(abx_sim* s, const float* d) {
  const int64_t NF = s->NF;
  for (int v = 0; v < s->B; ++v) {
    if (s->prec == ABX_F32)
      hipLaunchKernelGGL((k_init_eq<float, float>), dim3(nblk(NF, 256)), dim3(256), 0, s->stream, d, s->cells, NF,
                         s->plane, (float*)s->f[s->cur] + (int64_t)v * 19 * s->plane, lat_view(s), s->form);
    else if (s->prec == ABX_F32_F64)
      hipLaunchKernelGGL((k_init_eq<float, double>), dim3(nblk(NF, 256)), dim3(256), 0, s->stream, d, s->cells,
                         NF, s->plane, (float*)s->f[s->cur] + (int64_t)v * 19 * s->plane, lat_view(s), s->form);
    else
      hipLaunchKernelGGL((k_init_eq<double, double>), dim3(nblk(NF, 256)), dim3(256), 0, s->stream, d, s->cells,
                         NF, s->plane, (double*)s->f[s->cur] + (int64_t)v * 19 * s->plane, lat_view(s), s->form);
  }
  ABX_HIP(hipGetLastError());
  return 0;
}
}  // namespace

extern "C" int abx_sim_init_eq(abx_sim* s, const float* rho_u) {
  if (!s || !rho_u) {
    set_error("abx_sim_init_eq: null argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  const int64_t NF = s->NF;
  if (NF == 0) return 0;
  float* d = nullptr;
  ABX_HIP(hipMalloc((void**)&d, (size_t)(4 * NF) * sizeof(float)));
  ABX_HIP(hipMemcpyAsync(d, rho_u, (size_t)(4 * NF) * sizeof(float), hipMemcpyHostToDevice, s->stream));
  const int rc = sim_init_eq_dev(s, d);
  const hipError_t e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (rc) return rc;
  ABX_HIP(e);
  return 0;
}

extern "C" int abx_sim_init_eq_resample(abx_sim* s, const uint16_t* vol, int cstride, int vd, int vh, int vw, int nxc,
                                        int nyc, int nzc, const uint8_t* coarse_fluid, const double origin_c[3],
                                        double h_c, const double origin_f[3], double h_f, const float affine[8]) {
  if (!s || !vol || !coarse_fluid || !origin_c || !origin_f || !affine || cstride < 4 || nxc < 1 || nyc < 1 ||
      nzc < 1 || nxc > vd || nyc > vh || nzc > vw || !(h_c > 0) || !(h_f > 0)) {
    set_error("abx_sim_init_eq_resample: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  const int64_t NF = s->NF;
  if (NF == 0) return 0;
  ResampleArgs r;
  r.vol = vol;
  r.cfluid = coarse_fluid;
  r.cs = cstride;
  r.VD = vd;
  r.VH = vh;
  r.VW = vw;
  r.nxc = nxc;
  r.nyc = nyc;
  r.nzc = nzc;
  for (int a = 0; a < 3; ++a) {
    r.oc[a] = origin_c[a];
    r.of[a] = origin_f[a];
  }
  r.hc = h_c;
  r.hf = h_f;
  for (int c = 0; c < 4; ++c) {
    r.a[c] = affine[c];
    r.b[c] = affine[4 + c];
  }
  float* d = nullptr;
  ABX_HIP(hipMalloc((void**)&d, (size_t)(4 * NF) * sizeof(float)));
  hipLaunchKernelGGL(k_resample_rho_u, dim3(nblk(NF, 256)), dim3(256), 0, s->stream, r, s->cells, NF, lat_view(s),
                     s->rowinfo, s->ny, d);
  const hipError_t le = hipGetLastError();
  int rc = 0;
  if (le != hipSuccess) {
    set_errorf("abx_sim_init_eq_resample: %s", hipGetErrorString(le));
    rc = -1;
  } else {
    rc = sim_init_eq_dev(s, d);
  }
  const hipError_t e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (rc) return rc;
  ABX_HIP(e);
  return 0;
}

extern "C" int abx_sim_reset(abx_sim* s) {
  if (!s) {
    set_error("abx_sim_reset: null argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  if (fill_cold(s) || zero_acc(s)) return -1;
  ABX_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

extern "C" int abx_sim_clear_acc(abx_sim* s) {
  if (!s) {
    set_error("abx_sim_clear_acc: null argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  return zero_acc(s);
}

extern "C" int abx_sim_set_ns(abx_sim* s, const float* ns) {
  if (!s) {
    set_error("abx_sim_set_ns: null argument");
    return -1;
  }
  sim_drop_graphs(s);
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  ABX_HIP(hipStreamSynchronize(s->stream));
  // cells in cid order are needed on the host to place ns
  std::vector<int32_t> cells((size_t)s->NF);
  if (s->NF) ABX_HIP(hipMemcpy(cells.data(), s->cells, (size_t)s->NF * 4, hipMemcpyDeviceToHost));
  return upload_ns(s, ns, cells);
}

extern "C" int abx_sim_sync(abx_sim* s) {
  if (!s) {
    set_error("abx_sim_sync: null argument");
    return -1;
  }
  ABX_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

extern "C" int abx_sim_set_launch(abx_sim* s, int kernel, int xcd_map, int nt) {
  if (!s || (kernel != 2 && kernel != 4) || !nt_supported(nt)) {
    set_error("abx_sim_set_launch: bad argument (kernel 2 or 4; nt: even-step flags 0, 3 or 19 plus 256 x odd-step "
              "flags 0, 1, 2, 3, 8, 11, 16, 19, 32 or 48)");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  sim_drop_graphs(s);
  if (kernel == 2 && ensure_second_lattice(s)) return -1;
  if (kernel != s->kernel) {
    // carry the state over: read it in the old form, store it in the new one
    const size_t n = (size_t)(19 * s->NF);
    std::vector<char> host(std::max<size_t>(n * s->ssz, 1) * (size_t)s->B);
    for (int b = 0; b < s->B; ++b)
      if (abx_sim_read_f(s, b, host.data() + (size_t)b * n * s->ssz)) return -1;
    s->kernel = kernel;
    s->form = kernel == 4 ? 1 : 0;
    s->cur = 0;
    for (int b = 0; b < s->B; ++b)
      if (abx_sim_write_f(s, b, host.data() + (size_t)b * n * s->ssz)) return -1;
  }
  s->xcd_map = xcd_map < 0 ? 4 : std::min(xcd_map, 4096);
  s->nt = nt;
  return 0;
}


extern "C" int abx_sim_set_occupancy(abx_sim* s, int even_waves_per_cu, int odd_waves_per_cu) {
  if (!s || even_waves_per_cu < 0 || odd_waves_per_cu < 0 || even_waves_per_cu > 64 || odd_waves_per_cu > 64) {
    set_error("abx_sim_set_occupancy: waves per CU must be 0 (no cap) to 64");
    return -1;
  }
  sim_drop_graphs(s);
  s->occ_even = even_waves_per_cu;
  s->occ_odd = odd_waves_per_cu;
  return 0;
}

extern "C" int abx_sim_pattern_time(abx_sim* s, int iters, double* ms_even, double* ms_odd) {
  if (!s || iters < 1 || !ms_even || !ms_odd) {
    set_error("abx_sim_pattern_time: bad argument");
    return -1;
  }
  if (s->kernel != 4 || s->prec == ABX_F64 || s->B != 1) {
    set_error("abx_sim_pattern_time: needs a single float-storage handle on the AA pattern");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  sim_drop_graphs(s);
  StepArgs a = step_args(s, 0);
  const dim3 grid(seg_grid_x(a)), block(256);
  hipEvent_t ev[3];
  for (auto& e : ev) ABX_HIP(hipEventCreate(&e));
  // pairs of steps in the order the stored form asks for, so the form after
  // every pair is the one before (the populations themselves are moved)
  const bool even_first = s->form == 2;
  const size_t lds_even = occupancy_lds(s->dev, s->occ_even, 4), lds_odd = occupancy_lds(s->dev, s->occ_odd, 4);
  float te = 0, to = 0;
  int rc = 0;
  for (int i = 0; i <= iters && !rc; ++i) {  // pair 0 warms up, untimed
    for (int k = 0; k < 2; ++k) {
      const bool even = (k == 0) == even_first;
      if (hipEventRecord(ev[k], s->stream) != hipSuccess) rc = -1;
      if (even)
        hipLaunchKernelGGL((k_pattern_seg<1, 19>), grid, block, lds_even, s->stream, a);
      else
        hipLaunchKernelGGL((k_pattern_seg<2, 0>), grid, block, lds_odd, s->stream, a);
    }
    if (hipEventRecord(ev[2], s->stream) != hipSuccess || hipEventSynchronize(ev[2]) != hipSuccess ||
        hipGetLastError() != hipSuccess)
      rc = -1;
    float m0 = 0, m1 = 0;
    (void)hipEventElapsedTime(&m0, ev[0], ev[1]);
    (void)hipEventElapsedTime(&m1, ev[1], ev[2]);
    if (i > 0) {
      te += even_first ? m0 : m1;
      to += even_first ? m1 : m0;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) {
    set_error("abx_sim_pattern_time: launch failed");
    return -1;
  }
  *ms_even = te / iters;
  *ms_odd = to / iters;
  return 0;
}

extern "C" void abx_sim_destroy(abx_sim* s) { sim_free(s); }

namespace abx {
StepArgs sim_step_args(abx_sim* s, int doacc) { return step_args(s, doacc); }
int sim_form(abx_sim* s) { return s->form; }
void sim_set_form(abx_sim* s, int form) { s->form = form; }
int sim_normalize(abx_sim* s) {
  if (s->kernel != 4 || s->form == 1) return 0;
  const size_t n = (size_t)(19 * s->NF) * s->ssz;
  std::vector<char> host(std::max<size_t>(n, 1) * (size_t)s->B);
  for (int b = 0; b < s->B; ++b)
    if (abx_sim_read_f(s, b, host.data() + (size_t)b * n)) return -1;
  s->form = 1;
  for (int b = 0; b < s->B; ++b)
    if (abx_sim_write_f(s, b, host.data() + (size_t)b * n)) return -1;
  return 0;
}
std::vector<int32_t> sim_segments(abx_sim* s) {
  std::vector<int32_t> v((size_t)s->nchunks64);
  if (s->nchunks64) (void)hipMemcpy(v.data(), s->chunks64, v.size() * 4, hipMemcpyDeviceToHost);
  return v;
}
SimView sim_view(abx_sim* s) {
  SimView v;
  v.ordinal = s->dev->ordinal;
  v.stream = s->stream;
  v.B = s->B;
  v.nx = s->nx;
  v.ny = s->ny;
  v.nz = s->nz;
  v.rowid = s->rowid;
  v.rowinfo = s->rowinfo;
  v.plane = s->plane;
  v.ssz = s->ssz;
  v.acc = s->acc;
  v.ns = s->ns;
  v.fluid_bits = s->fluid_bits;
  v.u0 = s->u0;
  v.Ain = s->Ain;
  v.Af = s->Af;
  return v;
}
}  // namespace abx
