// Device-side pieces of the stream-collide shared by the single-lattice and
// the batched launches: launch arguments, the collision (the reference's
// model, engine.js:277-318), the AA-pattern slot mapping and the per-lane
// segment step.  Both launches compile this same code, so a batched solve is
// bit-identical to the same solve run alone.
#pragma once

#include <type_traits>
#include <utility>

#include "abx_common.hpp"

namespace abx {

constexpr int CHUNK = 256;

// The lattice of a simulation (row-span layout).  Every row (i, j) of the
// grid that holds fluid stores the span of k from its first to its last fluid
// cell, and the rows follow one another without gaps, so cell (i, j, k) sits
// at position rowoff(i, j) + k of every population plane and nearly every
// stored position is a fluid cell.  The neighbour x + c of a cell lies in row
// (i + cx, j + cy) at position d + delta(row, cx, cy) + cz.
//   meta[d]      bit q (q < 19): the cell x - c_q is fluid (0 at non-fluid
//                positions); bits 19-25: the position's row minus the first
//                row of its 64-position segment
//   seg_row0[s]  first row overlapping positions 64 s .. 64 s + 63
//   rdelta[12 r + (di + 1) * 3 + (dj + 1)]  position delta from row r to the
//                row (i + di, j + dj) (0 where that row holds no fluid)
struct LatView {
  const uint32_t* meta;
  const int32_t* seg_row0;
  const int32_t* rdelta;
};

// Row of position d.
__device__ __forceinline__ int lat_row(const LatView& L, int d, uint32_t meta_d) {
  return L.seg_row0[d >> 6] + (int)((meta_d >> 19) & 127u);
}

struct StepArgs {
  const void* fin;
  void* fout;
  void* acc;
  const float* ns;
  const uint32_t* fluid_bits;  // word w holds positions 32w..32w+31
  const uint32_t* type_bits;   // per 32 positions: inlet, outlet, porous words
  const int32_t* chunks;       // first position of each active 64-position segment
  int nchunks, per_xcd, xcd_map;
  LatView lat;
  int64_t plane;
  double u0, tau0, sig, aniso, cs2s;
};

// min(x, 0.9) that, like Math.min, returns NaN for a NaN argument.
template <typename M>
__device__ __forceinline__ M min09(M x) {
  return x > (M)0.9 ? (M)0.9 : x;
}

// Load / store, non-temporal (streamed through the caches without keeping
// the line) when NT.
template <bool NT, typename T>
__device__ __forceinline__ T ld(const T* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// c*x for a lattice coefficient c in {-1, 0, 1}, as the reference's
// product but without the multiply: the term is dropped when c == 0 (see
// csum below), negated when c == -1.
template <typename M>
__device__ __forceinline__ M cterm(int c, M x) {
  return c > 0 ? x : -x;
}
// (cx*ux + cy*uy) + cz*uz (engine.js:288, 295) with the zero-coefficient
// products left out.  For finite operands this equals the reference's value
// except possibly in the sign of a zero result, which never reaches a stored
// value (it only enters 1 + cu and cu*cu).
template <typename M>
__device__ __forceinline__ M dot3(int cx, int cy, int cz, M ux, M uy, M uz) {
  M r = 0;
  bool have = false;
  if (cx) {
    r = cterm(cx, ux);
    have = true;
  }
  if (cy) {
    r = have ? r + cterm(cy, uy) : cterm(cy, uy);
    have = true;
  }
  if (cz) r = have ? r + cterm(cz, uz) : cterm(cz, uz);
  return r;
}

// Collision of one cell from its pulled populations fi (engine.js:277-318):
// writes the post-collision populations to fo and the moments the averaging
// accumulates (physical velocity, density) to ux, uy, uz, rho.  The
// arithmetic is the reference's in its order (cpuStep for M = double; the
// WGSL port engine.js:395-443 for M = float), so <float storage, double M>
// reproduces cpuStep bit for bit.  The only liberty taken: a product by a
// lattice coefficient 0 or +-1 is not formed (adding c*x with c = 0 to a
// running sum, or multiplying by +-1, changes no finite result), which halves
// the instruction count of the kernel.
template <typename M, bool NS>
__device__ __forceinline__ void collide(const M (&fi)[19], bool inlet, bool outlet, bool porous, M s,
                                        const StepArgs& a, M (&fo)[19], M& ux, M& uy, M& uz, M& rho) {
  rho = 0;
  M mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += fi[q];
    if (CX[q]) mx += cterm(CX[q], fi[q]);
    if (CY[q]) my += cterm(CY[q], fi[q]);
    if (CZ[q]) mz += cterm(CZ[q], fi[q]);
  }
  ux = mx / rho;
  uy = my / rho;
  uz = mz / rho;
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
      const M cu = (M)3 * dot3(CX[q], CY[q], CZ[q], ux, uy, uz);
      fo[q] = (M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq);
    }
    return;
  }
  M usq = (M)1.5 * (ux * ux + uy * uy + uz * uz);
  M pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
  M fe[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const M cu = (M)3 * dot3(CX[q], CY[q], CZ[q], ux, uy, uz);
    fe[q] = (M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq);
    const M ne = fi[q] - fe[q];
    if (CX[q]) pxx += ne;
    if (CY[q]) pyy += ne;
    if (CZ[q]) pzz += ne;
    if (CX[q] * CY[q]) pxy += cterm(CX[q] * CY[q], ne);
    if (CX[q] * CZ[q]) pxz += cterm(CX[q] * CZ[q], ne);
    if (CY[q] * CZ[q]) pyz += cterm(CY[q] * CZ[q], ne);
  }
  const M Qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + (M)2 * (pxy * pxy + pxz * pxz + pyz * pyz));
  const M tau0 = (M)a.tau0;
  const M tau = (M)0.5 * (tau0 + sqrt(tau0 * tau0 + (M)a.cs2s * Qn / rho));
  if (porous) {
    // equilibrium at the Darcy-shifted velocity (engine.js:304-312)
    const M sg = (M)a.sig;
    const M sl = (M)1 - min09(tau * sg * (M)a.aniso), sv = (M)1 - min09(tau * sg);
    const M ex = ux * sl, ey = uy * sv, ez = uz * sl;
    usq = (M)1.5 * (ex * ex + ey * ey + ez * ez);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const M cu = (M)3 * dot3(CX[q], CY[q], CZ[q], ex, ey, ez);
      fe[q] = (M)wt(q) * rho * ((M)1 + cu + (M)0.5 * cu * cu - usq);
    }
    ux = (M)0.5 * (ux + ex);
    uy = (M)0.5 * (uy + ey);
    uz = (M)0.5 * (uz + ez);
  }
  const M it = (M)1 / tau;
  if (NS && s > (M)0) {
#pragma unroll
    for (int q = 0; q < 19; ++q) fo[q] = ((M)1 - s) * (fi[q] - (fi[q] - fe[q]) * it) + s * fi[OPP[q]];
    ux *= (M)1 - s;
    uy *= (M)1 - s;
    uz *= (M)1 - s;
  } else {
#pragma unroll
    for (int q = 0; q < 19; ++q) fo[q] = fi[q] - (fi[q] - fe[q]) * it;
  }
}

// FP32 collision: the same model (BGK + Smagorinsky + anisotropic Darcy
// shift + partial bounce-back, engine.js:277-318) in a cheaper but
// mathematically identical formulation, for the float kernel:
//  - the equilibria of a pair +c/-c share their even part:
//      feq(+-c) = w rho (1 - 1.5 u.u + 4.5 (c.u)^2) +- 3 w rho (c.u);
//  - the non-equilibrium stress comes from the second moments of f, using
//    the D3Q19 identity sum_q c_a c_b feq_q = rho (u_a u_b + delta_ab / 3);
//  - relaxation is f + (feq - f) / tau as one explicit FMA (contraction is
//    otherwise off, so every launch variant rounds identically).
// Rounding differs from the reference's order at the 1e-7 level; the result
// is held to the reference's GPU/CPU gate (max |df| < 1e-3 after 12 steps).
template <bool NS>
__device__ __forceinline__ void collide_fast(const float (&fi)[19], bool inlet, bool outlet, bool porous, float s,
                                             const StepArgs& a, float (&fo)[19], float& ux, float& uy, float& uz,
                                             float& rho) {
  rho = 0;
  float mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += fi[q];
    if (CX[q]) mx += cterm(CX[q], fi[q]);
    if (CY[q]) my += cterm(CY[q], fi[q]);
    if (CZ[q]) mz += cterm(CZ[q], fi[q]);
  }
  const float ir = 1.0f / rho;
  ux = mx * ir;
  uy = my * ir;
  uz = mz * ir;
  const float u0 = (float)a.u0;
  float ex = ux, ey = uy, ez = uz, omega = 1.0f;
  if (inlet || outlet) {
    if (inlet) {
      float s0 = 0, sm = 0;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        if (CZ[q] == 0)
          s0 += fi[q];
        else if (CZ[q] < 0)
          sm += fi[q];
      }
      rho = (s0 + 2.0f * sm) / (1.0f - u0);
      ux = 0;
      uy = 0;
      uz = u0;
    } else {
      rho = 1.0f;
    }
    ex = ux;
    ey = uy;
    ez = uz;
  } else {
    // non-equilibrium second moments
    float sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      if (CX[q]) sxx += fi[q];
      if (CY[q]) syy += fi[q];
      if (CZ[q]) szz += fi[q];
      if (CX[q] * CY[q]) sxy += cterm(CX[q] * CY[q], fi[q]);
      if (CX[q] * CZ[q]) sxz += cterm(CX[q] * CZ[q], fi[q]);
      if (CY[q] * CZ[q]) syz += cterm(CY[q] * CZ[q], fi[q]);
    }
    const float r3 = rho * (1.0f / 3.0f);
    const float pxx = sxx - rho * ux * ux - r3, pyy = syy - rho * uy * uy - r3, pzz = szz - rho * uz * uz - r3;
    const float pxy = sxy - rho * ux * uy, pxz = sxz - rho * ux * uz, pyz = syz - rho * uy * uz;
    const float Qn = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
    const float tau0 = (float)a.tau0;
    const float tau = 0.5f * (tau0 + sqrtf(tau0 * tau0 + (float)a.cs2s * Qn * ir));
    if (porous) {
      const float sg = (float)a.sig;
      const float sl = 1.0f - min09(tau * sg * (float)a.aniso), sv = 1.0f - min09(tau * sg);
      ex = ux * sl;
      ey = uy * sv;
      ez = uz * sl;
      ux = 0.5f * (ux + ex);
      uy = 0.5f * (uy + ey);
      uz = 0.5f * (uz + ez);
    }
    omega = 1.0f / tau;
  }
  // equilibria at (rho, e): rest, then the nine +c/-c pairs (q, q+1)
  const float usq = 1.5f * (ex * ex + ey * ey + ez * ez);
  const bool eqonly = inlet || outlet;
  float feq[19];
  feq[0] = (float)wt(0) * rho * (1.0f - usq);
#pragma unroll
  for (int q = 1; q < 19; q += 2) {
    const float wr = (float)wt(q) * rho;
    const float cu = dot3(CX[q], CY[q], CZ[q], ex, ey, ez);
    const float even = wr * (1.0f - usq + 4.5f * cu * cu), odd = 3.0f * wr * cu;
    feq[q] = even + odd;
    feq[q + 1] = even - odd;
  }
  if (eqonly) {
#pragma unroll
    for (int q = 0; q < 19; ++q) fo[q] = feq[q];
  } else if (NS && s > 0.0f) {
#pragma unroll
    for (int q = 0; q < 19; ++q) fo[q] = (1.0f - s) * fmaf(feq[q] - fi[q], omega, fi[q]) + s * fi[OPP[q]];
    ux *= 1.0f - s;
    uy *= 1.0f - s;
    uz *= 1.0f - s;
  } else {
#pragma unroll
    for (int q = 0; q < 19; ++q) fo[q] = fmaf(feq[q] - fi[q], omega, fi[q]);
  }
}

// Dispatch: float arithmetic takes the FP32 formulation, double the
// reference's order.
template <typename M, bool NS>
__device__ __forceinline__ void collide_any(const M (&fi)[19], bool inlet, bool outlet, bool porous, M s,
                                            const StepArgs& a, M (&fo)[19], M& ux, M& uy, M& uz, M& rho) {
  if constexpr (sizeof(M) == 4)
    collide_fast<NS>(fi, inlet, outlet, porous, s, a, fo, ux, uy, uz, rho);
  else
    collide<M, NS>(fi, inlet, outlet, porous, s, a, fo, ux, uy, uz, rho);
}

// Post-collision population q of the fluid cell at d, whatever the stored
// form: two lattices / AA after an even step ("swapped": F[opp q](x)) or AA
// after an odd step ("streamed": F[q](x + c_q), or F[opp q](x) towards a
// wall; x + c_q = x - c_opp(q) is fluid iff meta bit opp(q) of x is set).
template <typename S>
__device__ __forceinline__ S post_f(const S* F, int64_t P, const LatView& L, int d, int q, int form) {
  if (form == 0) return F[q * P + d];
  if (form == 1) return F[OPP[q] * P + d];
  const uint32_t m = L.meta[d];
  if (!((m >> OPP[q]) & 1u)) return F[OPP[q] * P + d];
  const int r = lat_row(L, d, m);
  return F[q * P + d + L.rdelta[12 * r + (1 + CX[q]) * 3 + (1 + CY[q])] + CZ[q]];
}

// Slot that holds post-collision population q of the fluid cell at d in a
// given form (the inverse of post_f).
__device__ __forceinline__ int64_t post_slot(int64_t P, const LatView& L, int d, int q, int form) {
  if (form == 0) return q * P + d;
  if (form == 1) return OPP[q] * P + d;
  const uint32_t m = L.meta[d];
  if (!((m >> OPP[q]) & 1u)) return OPP[q] * P + d;
  const int r = lat_row(L, d, m);
  return q * P + d + L.rdelta[12 * r + (1 + CX[q]) * 3 + (1 + CY[q])] + CZ[q];
}

// f(integral_constant<int, q>) for q = 0 .. 18 in order, q a constant
// expression in each call.
template <typename F, int... Q>
__device__ __forceinline__ void for_each_q(F&& f, std::integer_sequence<int, Q...>) {
  (f(std::integral_constant<int, Q>{}), ...);
}
template <typename F>
__device__ __forceinline__ void for_each_q(F&& f) {
  for_each_q(f, std::make_integer_sequence<int, 19>{});
}

// One lane's cell of a 64-position segment starting at seg: the fused
// stream-collide of the scheme MODE (see k_stream_collide_seg).
// NT selects non-temporal accesses (streamed through the caches without
// keeping the lines): bit 0 / bit 1 = population loads / stores at the
// cell's own position (every population of the even AA step), bit 2 / bit 3
// = the odd step's loads / stores at neighbour positions, bit 4 = the
// averaged fields, bit 5 = the meta words.  MOVE: no collision, every
// population leaves as it came (the step's memory pattern alone, for timing).
template <typename S, typename M, bool ACC, bool NS, int MODE, int NT = 0, bool MOVE = false>
__device__ __forceinline__ void seg_cell(const StepArgs& a, int seg, int l, int b) {
  const int d = seg + l;
  // AA steps: a non-fluid position (a solid cell inside a row's span, or
  // padding) whose 32-position half of the segment holds fluid writes its own
  // 19 slots (slot opp q, as a fluid cell walled in on all sides): in the even
  // step the values it reads from its own slots, in the odd step zeros,
  // without reading.  No fluid cell ever reads or writes those slots and
  // nothing else reads them, so their values do not matter; this only makes
  // every memory line that holds fluid fully written.  (Both choices were
  // timed: the even step's unconditional aligned loads are cheaper than
  // masked ones, the odd step's masked gathers cheaper than full ones.)
  const uint32_t own = a.fluid_bits[d >> 5];
  const bool fluid = (own >> (d & 31)) & 1u;
  if (!fluid && (MODE == 0 || own == 0u)) return;
  const int64_t P = a.plane;
  const S* fin = (const S*)a.fin + (int64_t)b * 19 * P;
  S* fout = (S*)a.fout + (int64_t)b * 19 * P;
  const uint32_t tb = (uint32_t)d & 31u;
  const uint32_t* tw = a.type_bits + 3 * (d >> 5);
  const bool inlet = (tw[0] >> tb) & 1u, outlet = (tw[1] >> tb) & 1u, porous = (tw[2] >> tb) & 1u;
  M fi[19], fo[19];
  // neighbour bits and the position deltas to the nine neighbour rows (the
  // even AA step touches only the cell's own position)
  uint32_t nbm = 0;  // bit q: the cell x - c_q is fluid
  int dl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if constexpr (MODE != 1) {
    const uint32_t m = ld<(NT & 32) != 0>(a.lat.meta + d);
    nbm = m & 0x7ffffu;
    const int4* rd = (const int4*)(a.lat.rdelta + 12 * lat_row(a.lat, d, m));
    const int4 v0 = rd[0], v1 = rd[1];
    dl[0] = v0.x;
    dl[1] = v0.y;
    dl[2] = v0.z;
    dl[3] = v0.w;
    dl[4] = v1.x;
    dl[5] = v1.y;
    dl[6] = v1.z;
    dl[7] = v1.w;
    dl[8] = rd[2].x;
  }
  auto load = [&](int q, auto nt) {
    const bool nbf = (nbm >> q) & 1u;
    // x - c_q: row (i - cx, j - cy), k - cz
    const int src_nb = d + dl[(1 - CX[q]) * 3 + (1 - CY[q])] - CZ[q];
    const S* src;
    if (MODE == 0)
      src = nbf ? fin + q * P + src_nb : fin + OPP[q] * P + d;
    else if (MODE == 1)
      src = fin + q * P + d;
    else
      src = nbf ? fin + OPP[q] * P + src_nb : fin + q * P + d;
    fi[q] = (M)ld<decltype(nt)::value>(src);
  };
  if (MODE != 2 || fluid) {
    for_each_q([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr bool here = MODE == 1 || (CX[q] == 0 && CY[q] == 0 && CZ[q] == 0);
      load(q, std::integral_constant<bool, here ? (NT & 1) != 0 : (NT & 4) != 0>{});
    });
  } else {
#pragma unroll
    for (int q = 0; q < 19; ++q) fi[q] = 0;
  }
  M s = 0;
  if (NS) s = (M)a.ns[(int64_t)b * P + d];
  M ux = 0, uy = 0, uz = 0, rho = 0;
  if constexpr (MOVE) {
#pragma unroll
    for (int q = 0; q < 19; ++q) fo[q] = fi[q];
  } else {
    collide_any<M, NS>(fi, inlet, outlet, porous, s, a, fo, ux, uy, uz, rho);
  }
  auto store = [&](int q, auto nt) {
    S* dst;
    if (MODE == 0) {
      dst = fout + q * P + d;
    } else if (MODE == 1) {
      dst = fout + OPP[q] * P + d;
    } else {
      // x + c_q: row (i + cx, j + cy), k + cz; fluid iff x - c_opp(q) is
      const int dst_nb = d + dl[(1 + CX[q]) * 3 + (1 + CY[q])] + CZ[q];
      dst = ((nbm >> OPP[q]) & 1u) ? fout + q * P + dst_nb : fout + OPP[q] * P + d;
    }
    st<decltype(nt)::value>(dst, fluid ? (S)fo[q] : (S)fi[q]);
  };
  for_each_q([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr bool here = MODE != 2 || (CX[q] == 0 && CY[q] == 0 && CZ[q] == 0);
    store(q, std::integral_constant<bool, here ? (NT & 2) != 0 : (NT & 8) != 0>{});
  });
  if (ACC) {
    constexpr bool NTC = (NT & 16) != 0;
    S* ac = (S*)a.acc + (int64_t)b * 4 * P;
    // non-fluid positions write zero sums (never read) without reading
    S a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    if (fluid) {
      a0 = (S)((M)ld<NTC>(ac + d) + ux);
      a1 = (S)((M)ld<NTC>(ac + P + d) + uy);
      a2 = (S)((M)ld<NTC>(ac + 2 * P + d) + uz);
      a3 = (S)((M)ld<NTC>(ac + 3 * P + d) + rho);
    }
    st<NTC>(ac + d, a0);
    st<NTC>(ac + P + d, a1);
    st<NTC>(ac + 2 * P + d, a2);
    st<NTC>(ac + 3 * P + d, a3);
  }
}

}  // namespace abx
