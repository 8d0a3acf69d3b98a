// D3Q19 stream-collide on gfx950: the device side of the reference's
// setupSim / cpuStep / GpuSim / probe (engine.js:239-321, 378-462, 493-594).
//
// Lattice layout (HBM).  The reference stores the compacted fluid list with a
// 19-entry neighbour table (engine.js:241-251), 76 B of indices per cell that
// every step re-reads.  Here the lattice is dense: the nx*ny*nz grid with a
// one-cell halo on every side and the k extent padded to a multiple of 32,
//     D(i,j,k) = ((i+1)*nyp + (j+1))*nzp + (k+1),  nyp = ny+2, nzp = 32*ceil((nz+2)/32),
// one plane of `plane` elements per population q (structure of arrays,
// f[q*plane + D]).  Streaming pulls f_q(x - c_q) at the constant offset
// off_q = (cx*nyp + cy)*nzp + cz; a neighbour that is not fluid (a wall, the
// halo, or outside the box) turns into halfway bounce-back f_opp(q)(x) of the
// previous step, which is exactly the reference's "nb = -1" case: prepare
// keeps a solid shell between the fluid and the grid faces (engine.js:177-178,
// 190-195), so its flat-index wrap-around never reaches a fluid cell.
//
// Per cell, the only geometry data are bits: a fluid bit (bit-packed, 32
// cells per word, staged per workgroup in LDS for the 18 neighbour tests) and
// three type bits (inlet / outlet / porous, engine.js:252-254).  A workgroup
// owns 256 consecutive lattice positions; only workgroups that hold fluid are
// launched (the chunk list).  Algorithmic traffic per fluid cell update:
// 19 loads + 19 stores of the storage type (152 B in float).
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "abx_common.hpp"
#include "sim_internal.hpp"

namespace abx {

constexpr int CHUNK = 256;

struct StepArgs {
  const void* fin;
  void* fout;
  void* acc;
  const float* ns;
  const uint32_t* fluid_bits;  // word w holds positions 32w..32w+31; readable from -margin
  const uint32_t* type_bits;   // per 32 positions: inlet, outlet, porous words
  const int32_t* chunks;       // first position of each active chunk
  int nchunks, per_xcd, xcd_map;
  int nyp, nzp;
  int64_t plane;
  double u0, tau0, sig, aniso, cs2s;
};

// min(x, 0.9) that, like Math.min, returns NaN for a NaN argument.
template <typename M>
__device__ __forceinline__ M min09(M x) {
  return x > (M)0.9 ? (M)0.9 : x;
}

// One fused pull-stream + collide step.  S = storage type, M = arithmetic
// type.  The arithmetic is the reference's, term by term and in its order
// (engine.js:269-320 for M = double; the WGSL port engine.js:386-449 for
// M = float), so <float, double> reproduces cpuStep bit for bit.
template <typename S, typename M, bool ACC, bool NS>
__global__ __launch_bounds__(CHUNK) void k_stream_collide(StepArgs a) {
  __shared__ uint32_t win[9][10];
  const int w = blockIdx.x;
  const int ci = a.xcd_map ? (w & 7) * a.per_xcd + (w >> 3) : w;
  if (ci >= a.nchunks) return;  // whole workgroup leaves together
  const int b = blockIdx.y;
  const int d0 = a.chunks[ci];
  const int t = threadIdx.x;
  // Fluid bits of the nine (di, dj) neighbour rows of the chunk, each window
  // covering positions [d0 - (di*nyp + dj)*nzp - 32, ... + 320).
  if (t < 90) {
    const int wi = t / 10, wo = t - wi * 10;
    const int di = wi / 3 - 1, dj = wi - (wi / 3) * 3 - 1;
    const int64_t base = (int64_t)d0 - ((int64_t)di * a.nyp + dj) * a.nzp - 32;
    win[wi][wo] = a.fluid_bits[(base >> 5) + wo];
  }
  __syncthreads();
  if (!((win[4][(t + 32) >> 5] >> ((t + 32) & 31)) & 1u)) return;
  const int d = d0 + t;
  const uint32_t tb = (uint32_t)d & 31u;
  const uint32_t* tw = a.type_bits + 3 * (d >> 5);
  const bool inlet = (tw[0] >> tb) & 1u, outlet = (tw[1] >> tb) & 1u, porous = (tw[2] >> tb) & 1u;
  const int64_t P = a.plane;
  const S* fin = (const S*)a.fin + (int64_t)b * 19 * P;
  S* fout = (S*)a.fout + (int64_t)b * 19 * P;

  M fi[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const int p = t + 32 - CZ[q];
    const int wi = (CX[q] + 1) * 3 + (CY[q] + 1);
    const bool nbf = (win[wi][p >> 5] >> (p & 31)) & 1u;
    const int off = (CX[q] * a.nyp + CY[q]) * a.nzp + CZ[q];
    const S* src = nbf ? fin + q * P + (d - off) : fin + OPP[q] * P + d;
    fi[q] = (M)*src;
  }
  M rho = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += fi[q];
    mx += (M)CX[q] * fi[q];
    my += (M)CY[q] * fi[q];
    mz += (M)CZ[q] * fi[q];
  }
  M ux = mx / rho, uy = my / rho, uz = mz / rho;
  const M u0 = (M)a.u0;
  if (inlet || outlet) {
    if (inlet) {
      // density from the populations that arrive from the interior (engine.js:282-284)
      M s0 = 0, sm = 0;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        if (CZ[q] == 0)
          s0 += fi[q];
        else if (CZ[q] < 0)
          sm += fi[q];
      }
      rho = (s0 + (M)2 * sm) / ((M)1 - u0);
      ux = 0;
      uy = 0;
      uz = u0;
    } else {
      rho = 1;
    }
    const M usq = (M)1.5 * (ux * ux + uy * uy + uz * uz);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const M cu = (M)3 * ((M)CX[q] * ux + (M)CY[q] * uy + (M)CZ[q] * uz);
      fout[q * P + d] = (S)((M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq));
    }
  } else {
    M usq = (M)1.5 * (ux * ux + uy * uy + uz * uz);
    M pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const M cu = (M)3 * ((M)CX[q] * ux + (M)CY[q] * uy + (M)CZ[q] * uz);
      const M e = (M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq);
      const M ne = fi[q] - e;
      pxx += (M)(CX[q] * CX[q]) * ne;
      pyy += (M)(CY[q] * CY[q]) * ne;
      pzz += (M)(CZ[q] * CZ[q]) * ne;
      pxy += (M)(CX[q] * CY[q]) * ne;
      pxz += (M)(CX[q] * CZ[q]) * ne;
      pyz += (M)(CY[q] * CZ[q]) * ne;
    }
    const M Qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + (M)2 * (pxy * pxy + pxz * pxz + pyz * pyz));
    const M tau0 = (M)a.tau0;
    const M tau = (M)0.5 * (tau0 + sqrt(tau0 * tau0 + (M)a.cs2s * Qn / rho));
    // velocity of the equilibrium the cell relaxes to (engine.js:304-312)
    M ex = ux, ey = uy, ez = uz;
    if (porous) {
      const M sg = (M)a.sig;
      const M sl = (M)1 - min09(tau * sg * (M)a.aniso), sv = (M)1 - min09(tau * sg);
      ex = ux * sl;
      ey = uy * sv;
      ez = uz * sl;
      usq = (M)1.5 * (ex * ex + ey * ey + ez * ez);
      ux = (M)0.5 * (ux + ex);
      uy = (M)0.5 * (uy + ey);
      uz = (M)0.5 * (uz + ez);
    }
    const M it = (M)1 / tau;
    M s = 0;
    if (NS) s = (M)a.ns[(int64_t)b * P + d];
    if (NS && s > (M)0) {
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const M cu = (M)3 * ((M)CX[q] * ex + (M)CY[q] * ey + (M)CZ[q] * ez);
        const M fe = (M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq);
        fout[q * P + d] = (S)(((M)1 - s) * (fi[q] - (fi[q] - fe) * it) + s * fi[OPP[q]]);
      }
      ux *= (M)1 - s;
      uy *= (M)1 - s;
      uz *= (M)1 - s;
    } else {
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const M cu = (M)3 * ((M)CX[q] * ex + (M)CY[q] * ey + (M)CZ[q] * ez);
        const M fe = (M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq);
        fout[q * P + d] = (S)(fi[q] - (fi[q] - fe) * it);
      }
    }
  }
  if (ACC) {
    S* ac = (S*)a.acc + (int64_t)b * 4 * P;
    ac[d] = (S)((M)ac[d] + ux);
    ac[P + d] = (S)((M)ac[P + d] + uy);
    ac[2 * P + d] = (S)((M)ac[2 * P + d] + uz);
    ac[3 * P + d] = (S)((M)ac[3 * P + d] + rho);
  }
}

// u_y = sum c_y f / sum f on the probe cells (engine.js:324-327, WGSL probe
// engine.js:455-461); the value is rounded to float as the reference's
// Float32Array holds it, except for double storage.
template <typename S, typename M>
__global__ void k_probe(const S* f, int64_t P, const int32_t* plist, int np, double* vy) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const int d = plist[n];
  M r = 0, my = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const M v = (M)f[q * P + d];
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

// out[c*NF + a] = src[c*P + cells[a]] (the reference's gather kernel,
// engine.js:468-476, over whole components).
template <typename S>
__global__ void k_gather(const S* src, int64_t P, int ncomp, const int32_t* cells, int64_t NF, S* out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NF * ncomp) return;
  const int64_t c = n / NF, a = n - c * NF;
  out[n] = src[c * P + cells[a]];
}

// acc / divisor in double, stored in T (engine.js:628: Float32Array /= nacc).
template <typename S, typename T>
__global__ void k_gather_div(const S* src, int64_t P, const int32_t* cells, int64_t NF, double div, T* out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NF * 4) return;
  const int64_t c = n / NF, a = n - c * NF;
  out[n] = (T)((double)src[c * P + cells[a]] / div);
}

template <typename S>
__global__ void k_scatter(const S* in, int64_t P, int ncomp, const int32_t* cells, int64_t NF, S* dst) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NF * ncomp) return;
  const int64_t c = n / NF, a = n - c * NF;
  dst[c * P + cells[a]] = in[n];
}

// f_q = feq_q(rho, u) at the fluid cells (feq as engine.js:296).
template <typename S, typename M>
__global__ void k_init_eq(const float* rho_u, const int32_t* cells, int64_t NF, int64_t P, S* f) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= NF) return;
  const M rho = (M)rho_u[a], ux = (M)rho_u[NF + a], uy = (M)rho_u[2 * NF + a], uz = (M)rho_u[3 * NF + a];
  const M usq = (M)1.5 * (ux * ux + uy * uy + uz * uz);
  const int d = cells[a];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const M cu = (M)3 * ((M)CX[q] * ux + (M)CY[q] * uy + (M)CZ[q] * uz);
    f[q * P + d] = (S)((M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq));
  }
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
  int prec = ABX_F32, B = 1, xcd_map = 1;
  int nx = 0, ny = 0, nz = 0, nxp = 0, nyp = 0, nzp = 0;
  int64_t ntot = 0, plane = 0, NF = 0;
  size_t ssz = 4;  // bytes per stored value
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  void* acc = nullptr;
  float* ns = nullptr;
  uint32_t* fluid_alloc = nullptr;
  const uint32_t* fluid_bits = nullptr;
  uint32_t* type_bits = nullptr;
  int32_t* chunks = nullptr;
  int nchunks = 0;
  int32_t* cells = nullptr;
  int32_t* plist = nullptr;
  double* vy = nullptr;
  double* red = nullptr;
  double* red_host = nullptr;
  int nM = 0, nT = 0, Ain = 0, Af = 0, Lp = 0;
  double u0 = 0, tau0 = 0, sig = 0, aniso = 0, cs2s = 0;
  int64_t device_bytes = 0;
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

StepArgs step_args(abx_sim* s, int doacc) {
  StepArgs a;
  a.fin = s->f[s->cur];
  a.fout = s->f[s->cur ^ 1];
  a.acc = doacc ? s->acc : nullptr;
  a.ns = s->ns;
  a.fluid_bits = s->fluid_bits;
  a.type_bits = s->type_bits;
  a.chunks = s->chunks;
  a.nchunks = s->nchunks;
  a.per_xcd = (s->nchunks + 7) / 8;
  a.xcd_map = s->xcd_map;
  a.nyp = s->nyp;
  a.nzp = s->nzp;
  a.plane = s->plane;
  a.u0 = s->u0;
  a.tau0 = s->tau0;
  a.sig = s->sig;
  a.aniso = s->aniso;
  a.cs2s = s->cs2s;
  return a;
}

template <typename S, typename M>
void launch_step_t(abx_sim* s, const StepArgs& a, int doacc) {
  const unsigned gx = a.xcd_map ? (unsigned)(8 * a.per_xcd) : (unsigned)a.nchunks;
  dim3 grid(gx, (unsigned)s->B), block(CHUNK);
  const bool ns = a.ns != nullptr;
  if (doacc) {
    if (ns)
      hipLaunchKernelGGL((k_stream_collide<S, M, true, true>), grid, block, 0, s->stream, a);
    else
      hipLaunchKernelGGL((k_stream_collide<S, M, true, false>), grid, block, 0, s->stream, a);
  } else {
    if (ns)
      hipLaunchKernelGGL((k_stream_collide<S, M, false, true>), grid, block, 0, s->stream, a);
    else
      hipLaunchKernelGGL((k_stream_collide<S, M, false, false>), grid, block, 0, s->stream, a);
  }
}

int launch_step(abx_sim* s, int doacc) {
  if (s->nchunks == 0) return 0;
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
  s->cur ^= 1;
  return 0;
}

int fill_cold(abx_sim* s) {
  const int64_t n = s->plane * 19 * s->B;
  for (int l = 0; l < 2; ++l) {
    if (s->prec == ABX_F64)
      hipLaunchKernelGGL(k_fill_wt<double>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (double*)s->f[l],
                         s->plane, (int)(19 * s->B));
    else
      hipLaunchKernelGGL(k_fill_wt<float>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (float*)s->f[l],
                         s->plane, (int)(19 * s->B));
  }
  ABX_HIP(hipGetLastError());
  return 0;
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
      hipFree(s->ns);
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

void sim_free(abx_sim* s) {
  if (!s) return;
  if (s->stream) hipStreamSynchronize(s->stream);
  void* ptrs[] = {s->f[0], s->f[1], s->acc, s->ns, s->fluid_alloc, s->type_bits, s->chunks, s->cells, s->plist,
                  s->vy, s->red};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (s->red_host) hipHostFree(s->red_host);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->stream) hipStreamDestroy(s->stream);
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
  s->nxp = nx + 2;
  s->nyp = ny + 2;
  s->nzp = ((nz + 2 + 31) / 32) * 32;
  s->ntot = (int64_t)s->nxp * s->nyp * s->nzp;
  s->plane = ((s->ntot + CHUNK - 1) / CHUNK) * CHUNK;
  if ((int64_t)s->plane * 19 * B > ((int64_t)1 << 40) || s->plane + (int64_t)(s->nyp + 2) * s->nzp > INT32_MAX) {
    set_error("abx_sim_create: grid too large for 32-bit lattice positions");
    delete s;
    return -1;
  }
  const int64_t N = (int64_t)nx * ny * nz;
  auto D = [&](int64_t i, int64_t j, int64_t k) { return ((i + 1) * s->nyp + (j + 1)) * s->nzp + (k + 1); };

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
  const int64_t margin = ((int64_t)(s->nyp + 1) * s->nzp) / 32 + 16;
  std::vector<uint32_t> fb((size_t)(nwords + 2 * margin), 0u);
  uint32_t* fbits = fb.data() + margin;
  int Ain = 0;
  for (int64_t c = 0; c < N; ++c) {
    if (!g->fluid[c]) continue;
    const int64_t k = c % nz, ij = c / nz, j = ij % ny, i = ij / ny;
    const int64_t d = D(i, j, k);
    if ((int64_t)cells.size() >= INT32_MAX) {
      set_error("abx_sim_create: too many fluid cells");
      delete s;
      return -1;
    }
    const int32_t a = (int32_t)cells.size();
    cells.push_back((int32_t)d);
    fbits[d >> 5] |= 1u << (d & 31);
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

  std::vector<int32_t> chunks;
  for (int64_t c0 = 0; c0 < s->plane; c0 += CHUNK) {
    uint32_t any = 0;
    for (int w = 0; w < CHUNK / 32; ++w) any |= fbits[(c0 >> 5) + w];
    if (any) chunks.push_back((int32_t)c0);
  }
  s->nchunks = (int)chunks.size();
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
    rc |= rc ? 0 : dalloc(s, (double**)&s->f[1], fcount);
    rc |= rc ? 0 : dalloc(s, (double**)&s->acc, (size_t)(s->plane * 4 * B));
  } else {
    rc |= dalloc(s, (float**)&s->f[0], fcount);
    rc |= rc ? 0 : dalloc(s, (float**)&s->f[1], fcount);
    rc |= rc ? 0 : dalloc(s, (float**)&s->acc, (size_t)(s->plane * 4 * B));
  }
  rc |= rc ? 0 : dalloc(s, &s->fluid_alloc, fb.size());
  rc |= rc ? 0 : dalloc(s, &s->type_bits, typ.size());
  rc |= rc ? 0 : dalloc(s, &s->chunks, std::max<size_t>(chunks.size(), 1));
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
  s->fluid_bits = s->fluid_alloc + margin;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream) : hipSuccess;
  };
  if (up(s->fluid_alloc, fb.data(), fb.size() * 4) != hipSuccess || up(s->type_bits, typ.data(), typ.size() * 4) ||
      up(s->chunks, chunks.data(), chunks.size() * 4) || up(s->cells, cells.data(), cells.size() * 4) ||
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
  return 0;
}

extern "C" int abx_sim_steps(abx_sim* s, int n, int doacc) {
  if (!s || n < 0) {
    set_error("abx_sim_steps: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  for (int i = 0; i < n; ++i)
    if (launch_step(s, doacc ? 1 : 0)) return -1;
  return 0;
}

extern "C" int abx_sim_time_steps(abx_sim* s, int n, int doacc, double* ms_total) {
  if (!s || n < 0 || !ms_total) {
    set_error("abx_sim_time_steps: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(s->dev->ordinal));
  ABX_HIP(hipEventRecord(s->ev0, s->stream));
  for (int i = 0; i < n; ++i)
    if (launch_step(s, doacc ? 1 : 0)) return -1;
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
                         (const float*)s->f[s->cur], s->plane, s->plist, np, s->vy);
    else if (s->prec == ABX_F32_F64)
      hipLaunchKernelGGL((k_probe<float, double>), dim3(nblk(np, 256)), dim3(256), 0, s->stream,
                         (const float*)s->f[s->cur], s->plane, s->plist, np, s->vy);
    else
      hipLaunchKernelGGL((k_probe<double, double>), dim3(nblk(np, 256)), dim3(256), 0, s->stream,
                         (const double*)s->f[s->cur], s->plane, s->plist, np, s->vy);
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
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(n, 256)), dim3(256), 0, s->stream,
                       (const double*)s->f[s->cur] + (int64_t)b * 19 * s->plane, s->plane, 19, s->cells, s->NF,
                       (double*)dout);
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(nblk(n, 256)), dim3(256), 0, s->stream,
                       (const float*)s->f[s->cur] + (int64_t)b * 19 * s->plane, s->plane, 19, s->cells, s->NF,
                       (float*)dout);
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
      hipLaunchKernelGGL(k_scatter<double>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (const double*)din,
                         s->plane, 19, s->cells, s->NF, (double*)s->f[s->cur] + (int64_t)v * 19 * s->plane);
    else
      hipLaunchKernelGGL(k_scatter<float>, dim3(nblk(n, 256)), dim3(256), 0, s->stream, (const float*)din,
                         s->plane, 19, s->cells, s->NF, (float*)s->f[s->cur] + (int64_t)v * 19 * s->plane);
  }
  ABX_HIP(hipGetLastError());
  ABX_HIP(hipStreamSynchronize(s->stream));
  ABX_HIP(hipFree(din));
  return 0;
}

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
  for (int v = 0; v < s->B; ++v) {
    if (s->prec == ABX_F32)
      hipLaunchKernelGGL((k_init_eq<float, float>), dim3(nblk(NF, 256)), dim3(256), 0, s->stream, d, s->cells, NF,
                         s->plane, (float*)s->f[s->cur] + (int64_t)v * 19 * s->plane);
    else if (s->prec == ABX_F32_F64)
      hipLaunchKernelGGL((k_init_eq<float, double>), dim3(nblk(NF, 256)), dim3(256), 0, s->stream, d, s->cells,
                         NF, s->plane, (float*)s->f[s->cur] + (int64_t)v * 19 * s->plane);
    else
      hipLaunchKernelGGL((k_init_eq<double, double>), dim3(nblk(NF, 256)), dim3(256), 0, s->stream, d, s->cells,
                         NF, s->plane, (double*)s->f[s->cur] + (int64_t)v * 19 * s->plane);
  }
  ABX_HIP(hipGetLastError());
  ABX_HIP(hipStreamSynchronize(s->stream));
  ABX_HIP(hipFree(d));
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

extern "C" int abx_sim_set_launch(abx_sim* s, int xcd_map) {
  if (!s) {
    set_error("abx_sim_set_launch: null argument");
    return -1;
  }
  s->xcd_map = xcd_map ? 1 : 0;
  return 0;
}

extern "C" void abx_sim_destroy(abx_sim* s) { sim_free(s); }

namespace abx {
SimView sim_view(abx_sim* s) {
  SimView v;
  v.ordinal = s->dev->ordinal;
  v.stream = s->stream;
  v.B = s->B;
  v.nx = s->nx;
  v.ny = s->ny;
  v.nz = s->nz;
  v.nyp = s->nyp;
  v.nzp = s->nzp;
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
