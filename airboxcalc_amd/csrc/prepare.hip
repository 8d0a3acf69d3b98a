// Geometry preparation: the reference's prepare(p) (engine.js:110-234) with
// its voxelizer (engine.js:45-93), 27-subcell vote (engine.js:127-135),
// inlet duct, floor, flood fill and leak check (engine.js:173-195) on the
// GPU, and the order-dependent opening choice (by BFS order over one 2-D
// slice, engine.js:141-171) and the lid and outlet rows (engine.js:196-214)
// on the host.
//
// Voxelizer.  The reference loops triangle by triangle over the columns of
// its bounding box, tests the column centre with f64 barycentrics, stores the
// hit depth as float32, then sorts each column and fills between pairs.  On
// the GPU the (triangle, column) pairs are enumerated flat (a prefix sum of
// per-triangle column counts, a binary search per pair), counted and
// scattered into per-column hit lists with atomics, and every column is then
// rank-sorted and filled by one wavefront.  The order in which hits land in a
// column does not matter: the rank sort orders equal values by slot and the
// fill only reads values.  Arithmetic is IEEE f64 without contraction
// (-ffp-contract=off), so every hit, every bound and every filled voxel is the
// reference's bit for bit.
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "abx_common.hpp"
#include "prepare_host.hpp"

namespace abx {

constexpr double JX = 1.37e-4, JY = 2.71e-4;  // ray jitter (engine.js:47)

struct TriCols {
  int i0, j0, ni, nj;
};

__device__ __forceinline__ double dmin3(double a, double b, double c) {
  // Math.min(a, b, c) for finite arguments
  double m = a < b ? a : b;
  return m < c ? m : c;
}
__device__ __forceinline__ double dmax3(double a, double b, double c) {
  double m = a > b ? a : b;
  return m > c ? m : c;
}

struct TriPts {
  double ax, ay, az, bx, by, bz, cx, cy, cz, den;
};

__device__ __forceinline__ TriPts tri_pts(const float* T, int64_t t, const double* o) {
  const float* p = T + 9 * t;
  TriPts r;
  r.ax = (double)p[0] - o[0] - JX;
  r.ay = (double)p[1] - o[1] - JY;
  r.az = (double)p[2] - o[2];
  r.bx = (double)p[3] - o[0] - JX;
  r.by = (double)p[4] - o[1] - JY;
  r.bz = (double)p[5] - o[2];
  r.cx = (double)p[6] - o[0] - JX;
  r.cy = (double)p[7] - o[1] - JY;
  r.cz = (double)p[8] - o[2];
  r.den = (r.by - r.cy) * (r.ax - r.cx) + (r.cx - r.bx) * (r.ay - r.cy);
  return r;
}

struct VoxArgs {
  const float* T;
  int64_t ntri;
  double o[3];
  int nx, ny, nz;
  double h;
};

// Column range of each triangle's bounding box (engine.js:56-59).
__global__ void k_tri_cols(VoxArgs a, TriCols* tc, int64_t* npairs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntri) return;
  const TriPts p = tri_pts(a.T, t, a.o);
  TriCols r = {0, 0, 0, 0};
  if (!(fabs(p.den) < 1e-12)) {
    const double i0 = fmax(0.0, floor(dmin3(p.ax, p.bx, p.cx) / a.h - 0.5));
    const double i1 = fmin((double)(a.nx - 1), ceil(dmax3(p.ax, p.bx, p.cx) / a.h - 0.5));
    const double j0 = fmax(0.0, floor(dmin3(p.ay, p.by, p.cy) / a.h - 0.5));
    const double j1 = fmin((double)(a.ny - 1), ceil(dmax3(p.ay, p.by, p.cy) / a.h - 0.5));
    if (i0 <= i1 && j0 <= j1) {
      r.i0 = (int)i0;
      r.j0 = (int)j0;
      r.ni = (int)(i1 - i0) + 1;
      r.nj = (int)(j1 - j0) + 1;
    }
  }
  tc[t] = r;
  npairs[t] = (int64_t)r.ni * r.nj;
}

// One (triangle, column) pair per thread: barycentric test at the column
// centre (engine.js:60-73).  Pass 0 counts hits per column, pass 1 stores
// the float32 hit depth into the column's slot range.
template <int PASS>
__global__ void k_tri_pairs(VoxArgs a, const TriCols* tc, const int64_t* pstart, int64_t total, int* cnt, int* pos,
                            float* hits) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= total) return;
  // last triangle whose first pair is <= n
  int64_t lo = 0, hi = a.ntri - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (pstart[mid] <= n)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int64_t t = lo;
  const TriCols c = tc[t];
  const int64_t r = n - pstart[t];
  const int i = c.i0 + (int)(r / c.nj), j = c.j0 + (int)(r % c.nj);
  const TriPts p = tri_pts(a.T, t, a.o);
  const double px = ((double)i + 0.5) * a.h, py = ((double)j + 0.5) * a.h;
  const double l1 = ((p.by - p.cy) * (px - p.cx) + (p.cx - p.bx) * (py - p.cy)) / p.den;
  if (l1 < 0) return;
  const double l2 = ((p.cy - p.ay) * (px - p.cx) + (p.ax - p.cx) * (py - p.cy)) / p.den;
  if (l2 < 0) return;
  const double l3 = 1 - l1 - l2;
  if (l3 < 0) return;
  const int col = i * a.ny + j;
  if (PASS == 0) {
    atomicAdd(&cnt[col], 1);
  } else {
    const int slot = atomicAdd(&pos[col], 1);
    hits[slot] = (float)(l1 * p.az + l2 * p.bz + l3 * p.cz);
  }
}

// One wavefront per column: rank-sort the column's hits, then set the voxels
// between consecutive pairs (engine.js:82-91).  Every byte of the column is
// written, so vox needs no clearing.
__global__ __launch_bounds__(256) void k_fill_cols(const int* off, int ncol, const float* hits, float* sorted,
                                                   int nz, double h, uint8_t* vox) {
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const bool live = col < ncol;
  const int a = live ? off[col] : 0, b = live ? off[col + 1] : 0, n = b - a;
  if (n >= 2) {
    for (int e = lane; e < n; e += 64) {
      const float x = hits[a + e];
      int rank = 0;
      for (int m = 0; m < n; ++m) {
        const float y = hits[a + m];
        rank += (y < x) || (y == x && m < e);
      }
      sorted[a + rank] = x;
    }
  }
  __threadfence_block();
  __syncthreads();
  if (!live) return;
  uint8_t* v = vox + (int64_t)col * nz;
  for (int k = lane; k < nz; k += 64) {
    uint8_t inside = 0;
    for (int m = 0; m + 1 < n; m += 2) {
      const double k0 = fmax(0.0, ceil((double)sorted[a + m] / h - 0.5));
      const double k1 = fmin((double)(nz - 1), ceil((double)sorted[a + m + 1] / h - 0.5) - 1);
      if ((double)k >= k0 && (double)k <= k1) inside = 1;
    }
    v[k] = inside;
  }
}

// solid = (sum of the 3x3x3 fine voxels) >= 7 (engine.js:127-135).
__global__ void k_vote(const uint8_t* fine, int nx, int ny, int nz, uint8_t* solid) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = (int64_t)nx * ny * nz;
  if (c >= N) return;
  const int k = (int)(c % nz);
  const int64_t ij = c / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  const int64_t nyf = 3 * (int64_t)ny, nzf = 3 * (int64_t)nz;
  int s = 0;
  for (int da = 0; da < 3; ++da)
    for (int db = 0; db < 3; ++db) {
      const int64_t base = ((int64_t)(3 * i + da) * nyf + (3 * j + db)) * nzf + 3 * k;
      s += fine[base] + fine[base + 1] + fine[base + 2];
    }
  solid[c] = s >= 7 ? 1 : 0;
}

// empty[i*nyf + j] = !fine[i][j][kf] (engine.js:142-144).
__global__ void k_slice(const uint8_t* fine, int nxf, int nyf, int nzf, int kf, uint8_t* empty) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)nxf * nyf) return;
  const bool in = kf >= 0 && kf < nzf;
  empty[c] = (in && fine[c * nzf + kf]) ? 0 : 1;
}

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---- inlet duct, floor, flood fill and leak check (engine.js:173-195) ------
// Duct: cells below kw0 are solid except under the opening; floor: rows
// whose depth lies beyond mx_d - h/2 are solid (the floor wins).
__global__ void k_duct_floor(uint8_t* solid, const uint8_t* opening, const uint8_t* floor_row, int nx, int ny, int nz,
                             int kw0) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)nx * ny * nz) return;
  const int k = (int)(c % nz);
  const int64_t ij = c / nz;
  uint8_t v = solid[c];
  if (k < kw0) v = opening[ij] ? 0 : 1;
  if (floor_row[ij % ny]) v = 1;
  solid[c] = v;
}

// Breadth-first flood of the air from the opening cells at k = 0, 6-connected,
// through cells that are not solid and lie at or below the seal (j >= top).
// The visited set is a bit per cell (set with atomicOr, so every cell enters
// the queue once); the queue holds cell indices, level by level: the level
// being expanded is Q[lo, end), new cells are appended at tail.  The final
// set is the connected part of the allowed cells reachable from the seeds,
// whatever the order in which a level is expanded.
struct FloodCtr {
  unsigned long long lo, end, tail;
  int leak;
};

__device__ __forceinline__ void flood_visit(int64_t c, int j, int top, const uint8_t* __restrict__ solid, uint32_t* fw,
                                            int32_t* Q, FloodCtr* ctr) {
  if (j < top || solid[c]) return;
  const uint32_t bit = 1u << (c & 31);
  if (fw[c >> 5] & bit) return;  // bits are only ever set: a set bit is final
  if (atomicOr(&fw[c >> 5], bit) & bit) return;
  Q[atomicAdd(&ctr->tail, 1ull)] = (int32_t)c;
}

__global__ void k_flood_seed(const uint8_t* __restrict__ opening, const uint8_t* __restrict__ solid, uint32_t* fw,
                             int32_t* Q, FloodCtr* ctr, int nx, int ny, int nz, int top) {
  const int64_t ij = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ij >= (int64_t)nx * ny || !opening[ij]) return;
  flood_visit(ij * nz, (int)(ij % ny), top, solid, fw, Q, ctr);
}

__global__ void k_flood_level(const uint8_t* __restrict__ solid, uint32_t* fw, int32_t* Q, FloodCtr* ctr, int nx,
                              int ny, int nz, int top) {
  const unsigned long long lo = ctr->lo, hi = ctr->end;
  const int64_t pl = (int64_t)ny * nz;
  for (unsigned long long q = lo + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; q < hi;
       q += (unsigned long long)gridDim.x * blockDim.x) {
    const int64_t c = Q[q];
    const int k = (int)(c % nz);
    const int64_t ij = c / nz;
    const int j = (int)(ij % ny), i = (int)(ij / ny);
    if (i > 0) flood_visit(c - pl, j, top, solid, fw, Q, ctr);
    if (i < nx - 1) flood_visit(c + pl, j, top, solid, fw, Q, ctr);
    if (j > 0) flood_visit(c - nz, j - 1, top, solid, fw, Q, ctr);
    if (j < ny - 1) flood_visit(c + nz, j + 1, top, solid, fw, Q, ctr);
    if (k > 0) flood_visit(c - 1, j, top, solid, fw, Q, ctr);
    if (k < nz - 1) flood_visit(c + 1, j, top, solid, fw, Q, ctr);
  }
}

// The next level is what the last one appended.
__global__ void k_flood_advance(FloodCtr* ctr) {
  ctr->lo = ctr->end;
  ctr->end = ctr->tail;
}

// Byte mask of the flooded cells, and the leak flag: air on the faces i = 0,
// i = nx - 1 or k = nz - 1 at or below the seal and not in the duct
// (engine.js:191-195).
__global__ void k_flood_finish(const uint32_t* fw, uint8_t* fluid, FloodCtr* ctr, int nx, int ny, int nz, int top,
                               int k0) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)nx * ny * nz) return;
  const uint8_t f = (fw[c >> 5] >> (c & 31)) & 1u;
  fluid[c] = f;
  if (!f) return;
  const int k = (int)(c % nz);
  const int64_t ij = c / nz;
  const int j = (int)(ij % ny), i = (int)(ij / ny);
  if (j >= top && k >= k0 && (i == 0 || i == nx - 1 || k == nz - 1)) atomicOr(&ctr->leak, 1);
}

// RAII device buffer for the preparation stages.
struct DBuf {
  void* p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
      set_errorf("hipMalloc of %zu bytes failed in geometry preparation", bytes);
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

// GPU voxelization of T into the device byte grid vox (nx*ny*nz).
int voxelize_device(const float* T_dev, int64_t ntri, const double origin[3], int nx, int ny, int nz, double h,
                    uint8_t* vox_dev, hipStream_t st) {
  const int64_t ncol = (int64_t)nx * ny;
  if (ncol >= INT32_MAX || (int64_t)nz * ncol > ((int64_t)1 << 40)) {
    set_error("voxelize: fine grid too large");
    return -1;
  }
  VoxArgs a;
  a.T = T_dev;
  a.ntri = ntri;
  a.o[0] = origin[0];
  a.o[1] = origin[1];
  a.o[2] = origin[2];
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.h = h;
  DBuf tc, np, ps, cnt, pos, hits, sorted;
  if (tc.alloc(sizeof(TriCols) * (size_t)std::max<int64_t>(ntri, 1)) ||
      np.alloc(8 * (size_t)std::max<int64_t>(ntri, 1)) || ps.alloc(8 * (size_t)std::max<int64_t>(ntri, 1)) ||
      cnt.alloc(4 * (size_t)(ncol + 1)) || pos.alloc(4 * (size_t)(ncol + 1)))
    return -1;
  std::vector<int64_t> pairs((size_t)std::max<int64_t>(ntri, 1), 0);
  int64_t total = 0;
  if (ntri > 0) {
    hipLaunchKernelGGL(k_tri_cols, dim3(nb(ntri, 256)), dim3(256), 0, st, a, tc.as<TriCols>(), np.as<int64_t>());
    ABX_HIP(hipGetLastError());
    ABX_HIP(hipMemcpyAsync(pairs.data(), np.p, 8 * (size_t)ntri, hipMemcpyDeviceToHost, st));
    ABX_HIP(hipStreamSynchronize(st));
    for (int64_t t = 0; t < ntri; ++t) {
      const int64_t c = pairs[(size_t)t];
      pairs[(size_t)t] = total;
      total += c;
    }
    ABX_HIP(hipMemcpyAsync(ps.p, pairs.data(), 8 * (size_t)ntri, hipMemcpyHostToDevice, st));
  }
  ABX_HIP(hipMemsetAsync(cnt.p, 0, 4 * (size_t)(ncol + 1), st));
  if (total > 0) {
    hipLaunchKernelGGL(k_tri_pairs<0>, dim3(nb(total, 256)), dim3(256), 0, st, a, tc.as<TriCols>(),
                       ps.as<int64_t>(), total, cnt.as<int>(), (int*)nullptr, (float*)nullptr);
    ABX_HIP(hipGetLastError());
  }
  std::vector<int> off((size_t)(ncol + 1));
  ABX_HIP(hipMemcpyAsync(off.data(), cnt.p, 4 * (size_t)ncol, hipMemcpyDeviceToHost, st));
  ABX_HIP(hipStreamSynchronize(st));
  int64_t run = 0;
  for (int64_t c = 0; c < ncol; ++c) {
    const int64_t v = off[(size_t)c];
    off[(size_t)c] = (int)run;
    run += v;
    if (run > INT32_MAX) {
      set_error("voxelize: too many ray hits");
      return -1;
    }
  }
  off[(size_t)ncol] = (int)run;
  ABX_HIP(hipMemcpyAsync(cnt.p, off.data(), 4 * (size_t)(ncol + 1), hipMemcpyHostToDevice, st));
  ABX_HIP(hipMemcpyAsync(pos.p, off.data(), 4 * (size_t)(ncol + 1), hipMemcpyHostToDevice, st));
  if (hits.alloc(4 * (size_t)std::max<int64_t>(run, 1)) || sorted.alloc(4 * (size_t)std::max<int64_t>(run, 1)))
    return -1;
  if (total > 0) {
    hipLaunchKernelGGL(k_tri_pairs<1>, dim3(nb(total, 256)), dim3(256), 0, st, a, tc.as<TriCols>(),
                       ps.as<int64_t>(), total, (int*)nullptr, pos.as<int>(), hits.as<float>());
    ABX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fill_cols, dim3(nb(ncol, 4)), dim3(256), 0, st, cnt.as<int>(), (int)ncol, hits.as<float>(),
                     sorted.as<float>(), nz, h, vox_dev);
  ABX_HIP(hipGetLastError());
  ABX_HIP(hipStreamSynchronize(st));
  return 0;
}

// Inlet duct, floor, flood fill and leak check on the GPU (engine.js:173-195).
// solid (N bytes) is updated in place with the duct and floor; fluid receives
// the flooded cells, *nfluid their number and *leaked whether air reaches an
// outer face.
int flood_device(std::vector<uint8_t>& solid, const std::vector<uint8_t>& opening,
                 const std::vector<uint8_t>& floor_row, int nx, int ny, int nz, int kw0, int top,
                 std::vector<uint8_t>& fluid, int64_t* nfluid, bool* leaked) {
  const int64_t N = (int64_t)nx * ny * nz;
  if (N >= (int64_t)INT32_MAX) {
    set_error("abx_prepare: grid too large for the flood fill (2^31 cells)");
    return -1;
  }
  hipStream_t st;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    set_error("abx_prepare: stream");
    return -1;
  }
  DBuf dsolid, dopen, dfloor, dfw, dQ, dfluid, dctr;
  const int64_t nw = (N + 31) / 32;
  int rc = dsolid.alloc((size_t)N) || dopen.alloc(opening.size()) || dfloor.alloc(floor_row.size()) ||
           dfw.alloc(4 * (size_t)nw) || dQ.alloc(4 * (size_t)N) || dfluid.alloc((size_t)N) ||
           dctr.alloc(sizeof(FloodCtr));
  FloodCtr h;
  memset(&h, 0, sizeof h);
  if (!rc && (hipMemcpyAsync(dsolid.p, solid.data(), (size_t)N, hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(dopen.p, opening.data(), opening.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(dfloor.p, floor_row.data(), floor_row.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemsetAsync(dfw.p, 0, 4 * (size_t)nw, st) != hipSuccess ||
              hipMemcpyAsync(dctr.p, &h, sizeof h, hipMemcpyHostToDevice, st) != hipSuccess)) {
    set_error("abx_prepare: flood fill upload failed");
    rc = -1;
  }
  uint8_t* sol = dsolid.as<uint8_t>();
  uint32_t* fw = dfw.as<uint32_t>();
  int32_t* Q = dQ.as<int32_t>();
  FloodCtr* ctr = dctr.as<FloodCtr>();
  if (!rc) {
    hipLaunchKernelGGL(k_duct_floor, dim3(nb(N, 256)), dim3(256), 0, st, sol, dopen.as<uint8_t>(),
                       dfloor.as<uint8_t>(), nx, ny, nz, kw0);
    hipLaunchKernelGGL(k_flood_seed, dim3(nb((int64_t)nx * ny, 256)), dim3(256), 0, st, dopen.as<uint8_t>(), sol, fw,
                       Q, ctr, nx, ny, nz, top);
    hipLaunchKernelGGL(k_flood_advance, dim3(1), dim3(1), 0, st, ctr);
    // levels are issued 32 at a time; an empty level (lo == end) ends the search
    for (;;) {
      for (int r = 0; r < 32; ++r) {
        hipLaunchKernelGGL(k_flood_level, dim3(1024), dim3(256), 0, st, sol, fw, Q, ctr, nx, ny, nz, top);
        hipLaunchKernelGGL(k_flood_advance, dim3(1), dim3(1), 0, st, ctr);
      }
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(&h, dctr.p, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        set_error("abx_prepare: flood fill failed on the GPU");
        rc = -1;
        break;
      }
      if (h.lo == h.end) break;
    }
  }
  if (!rc) {
    hipLaunchKernelGGL(k_flood_finish, dim3(nb(N, 256)), dim3(256), 0, st, fw, dfluid.as<uint8_t>(), ctr, nx, ny, nz,
                       top, kw0 < 0 ? 0 : kw0);
    fluid.assign((size_t)N, 0);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(fluid.data(), dfluid.p, (size_t)N, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(solid.data(), dsolid.p, (size_t)N, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&h, dctr.p, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error("abx_prepare: flood fill read-back failed");
      rc = -1;
    }
  }
  (void)hipStreamDestroy(st);
  if (rc) return -1;
  *nfluid = (int64_t)h.tail;
  *leaked = h.leak != 0;
  return 0;
}

}  // namespace abx

using namespace abx;

struct abx_geom : GeomHost {
  uint8_t* fine_dev = nullptr;
  int64_t nfine = 0;
  int ordinal = 0;
  ~abx_geom() {
    if (fine_dev) (void)hipFree(fine_dev);
  }
};

extern "C" int abx_voxelize(abx_dev* dev, const float* T, int64_t ntri, const double origin[3], int nx, int ny,
                            int nz, double h, uint8_t* vox_out) {
  if (!dev || (!T && ntri > 0) || !origin || !vox_out || nx < 1 || ny < 1 || nz < 1 || ntri < 0) {
    set_error("abx_voxelize: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(dev->ordinal));
  hipStream_t st;
  ABX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  DBuf dT, dv;
  const size_t nvox = (size_t)nx * ny * nz;
  int rc = dT.alloc(36 * (size_t)std::max<int64_t>(ntri, 1)) || dv.alloc(nvox);
  if (!rc && ntri > 0 && hipMemcpyAsync(dT.p, T, 36 * (size_t)ntri, hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error("abx_voxelize: upload failed");
    rc = -1;
  }
  if (!rc) rc = voxelize_device(dT.as<float>(), ntri, origin, nx, ny, nz, h, dv.as<uint8_t>(), st);
  if (!rc && hipMemcpyAsync(vox_out, dv.p, nvox, hipMemcpyDeviceToHost, st) != hipSuccess) {
    set_error("abx_voxelize: download failed");
    rc = -1;
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess) {
    set_error("abx_voxelize: kernel failure");
    rc = -1;
  }
  (void)hipStreamDestroy(st);
  return rc ? -1 : 0;
}

extern "C" int abx_prepare(abx_dev* dev, const float* T0, int64_t nfloat, const abx_prep_params* p,
                           abx_geom** out) {
  if (!dev || !p || !out || (!T0 && nfloat > 0) || nfloat < 0) {
    set_error("abx_prepare: bad argument");
    return -1;
  }
  *out = nullptr;
  abx_geom* g = new abx_geom();
  g->ordinal = dev->ordinal;
  auto fail = [&]() {
    delete g;
    return -1;
  };
  // frame, bounding box and grid (engine.js:111-125)
  std::vector<float> T;
  if (prep_grid(T0, nfloat, p, *g, T)) return fail();
  const abx_geom_info& I = g->info;
  const int nx = I.nx, ny = I.ny, nz = I.nz, nxf = 3 * nx, nyf = 3 * ny, nzf = 3 * nz;
  const int64_t N = I.N, ntri = nfloat / 9;
  if (hipSetDevice(dev->ordinal) != hipSuccess) {
    set_error("abx_prepare: hipSetDevice failed");
    return fail();
  }

  // voxelize and vote on the GPU (engine.js:126-135), slice kf for the openings
  hipStream_t st;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    set_error("abx_prepare: stream");
    return fail();
  }
  const int kf = prep_slice_k(*g);
  std::vector<uint8_t> empty((size_t)nxf * nyf);
  g->solid.assign((size_t)N, 0);
  {
    DBuf dT, dsolid, dempty;
    g->nfine = (int64_t)nxf * nyf * nzf;
    int rc = dT.alloc(36 * (size_t)std::max<int64_t>(ntri, 1)) || dsolid.alloc((size_t)N) ||
             dempty.alloc(empty.size());
    if (!rc && hipMalloc((void**)&g->fine_dev, (size_t)g->nfine) != hipSuccess) {
      g->fine_dev = nullptr;
      set_errorf("hipMalloc of %lld bytes (fine voxel grid) failed", (long long)g->nfine);
      rc = -1;
    }
    if (!rc && ntri > 0 && hipMemcpyAsync(dT.p, T.data(), 36 * (size_t)ntri, hipMemcpyHostToDevice, st)) {
      set_error("abx_prepare: triangle upload failed");
      rc = -1;
    }
    if (!rc) rc = voxelize_device(dT.as<float>(), ntri, I.origin, nxf, nyf, nzf, I.h / 3, g->fine_dev, st);
    if (!rc) {
      hipLaunchKernelGGL(k_vote, dim3(nb(N, 256)), dim3(256), 0, st, g->fine_dev, nx, ny, nz, dsolid.as<uint8_t>());
      hipLaunchKernelGGL(k_slice, dim3(nb((int64_t)empty.size(), 256)), dim3(256), 0, st, g->fine_dev, nxf, nyf,
                         nzf, kf, dempty.as<uint8_t>());
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(g->solid.data(), dsolid.p, (size_t)N, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(empty.data(), dempty.p, empty.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        set_error("abx_prepare: voxel vote failed on the GPU");
        rc = -1;
      }
    }
    (void)hipStreamDestroy(st);
    if (rc) return fail();
  }
  // openings (engine.js:141-171), duct depth, seal row and floor rows
  std::vector<uint8_t> floor_row;
  if (prep_openings(p, empty.data(), *g) || prep_flood_inputs(*g, floor_row)) return fail();
  // inlet duct, floor, flood fill of the air and leak check on the GPU
  // (engine.js:173-195).  The reference applies duct and floor before it
  // checks the seal row; neither changes whether that check fails.
  int64_t nfluid = 0;
  bool leaked = false;
  if (flood_device(g->solid, g->opening, floor_row, nx, ny, nz, I.kw0, I.top, g->fluid, &nfluid, &leaked))
    return fail();
  // lid plenum and outlet, filter plane, points and summary (engine.js:196-234)
  if (prep_finish(p, nfluid, leaked, *g)) return fail();
  *out = g;
  return 0;
}

extern "C" int abx_geom_info_get(const abx_geom* g, abx_geom_info* out) {
  if (!g || !out) {
    set_error("abx_geom_info_get: null argument");
    return -1;
  }
  *out = g->info;
  return 0;
}

extern "C" int abx_geom_read(const abx_geom* g, int what, void* out, int64_t cap) {
  if (!g || !out) {
    set_error("abx_geom_read: null argument");
    return -1;
  }
  auto cp = [&](const void* src, int64_t n, size_t esz) {
    if (cap < n) {
      set_errorf("abx_geom_read: capacity %lld < %lld", (long long)cap, (long long)n);
      return -1;
    }
    if (n) memcpy(out, src, (size_t)n * esz);
    return 0;
  };
  switch (what) {
    case ABX_G_SOLID:
      return cp(g->solid.data(), (int64_t)g->solid.size(), 1);
    case ABX_G_FLUID:
      return cp(g->fluid.data(), (int64_t)g->fluid.size(), 1);
    case ABX_G_OPENING:
      return cp(g->opening.data(), (int64_t)g->opening.size(), 1);
    case ABX_G_OUTLET:
      return cp(g->outlet.data(), (int64_t)g->outlet.size(), 1);
    case ABX_G_XC:
      return cp(g->xc.data(), (int64_t)g->xc.size(), 4);
    case ABX_G_DC:
      return cp(g->dc.data(), (int64_t)g->dc.size(), 4);
    case ABX_G_ZC:
      return cp(g->zc.data(), (int64_t)g->zc.size(), 4);
    case ABX_G_OPENPTS:
      return cp(g->openPts.data(), (int64_t)g->openPts.size(), 8);
    case ABX_G_FILTERPTS:
      return cp(g->filterPts.data(), (int64_t)g->filterPts.size(), 8);
    case ABX_G_OUTLETPTS:
      return cp(g->outletPts.data(), (int64_t)g->outletPts.size(), 8);
    case ABX_G_FINE: {
      if (!g->fine_dev) {
        set_error("abx_geom_read: no fine voxel grid");
        return -1;
      }
      if (cap < g->nfine) {
        set_error("abx_geom_read: capacity too small for the fine grid");
        return -1;
      }
      ABX_HIP(hipSetDevice(g->ordinal));
      ABX_HIP(hipMemcpy(out, g->fine_dev, (size_t)g->nfine, hipMemcpyDeviceToHost));
      return 0;
    }
    default:
      set_errorf("abx_geom_read: unknown array %d", what);
      return -1;
  }
}

extern "C" int abx_geom_grid(const abx_geom* g, abx_grid* out) {
  if (!g || !out) {
    set_error("abx_geom_grid: null argument");
    return -1;
  }
  out->nx = g->info.nx;
  out->ny = g->info.ny;
  out->nz = g->info.nz;
  out->fluid = g->fluid.data();
  out->opening = g->opening.data();
  out->outlet = g->outlet.data();
  out->dc = g->dc.data();
  out->top = g->info.top;
  out->jm = g->info.jm;
  out->pleat = g->info.pleat;
  return 0;
}

extern "C" int abx_geom_eval_view(const abx_geom* g, abx_eval_geom* out) {
  if (!g || !out) {
    set_error("abx_geom_eval_view: null argument");
    return -1;
  }
  const abx_geom_info& I = g->info;
  out->nx = I.nx;
  out->ny = I.ny;
  out->nz = I.nz;
  out->top = I.top;
  out->jm = I.jm;
  out->h = I.h;
  out->pleat = I.pleat;
  out->mn_z = I.mn[2];
  out->opening = g->opening.data();
  out->xc = g->xc.data();
  out->dc = g->dc.data();
  out->zc = g->zc.data();
  out->oAx = I.oAx;
  out->dAx = I.dAx;
  out->iAx = I.iAx;
  out->sd = I.sd;
  out->si = I.si;
  out->seal = I.seal;
  out->cy = I.cy;
  out->sy = I.sy;
  return 0;
}

extern "C" void abx_geom_free(abx_geom* g) { delete g; }
