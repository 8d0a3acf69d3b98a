// BF16 kernels of the 3-D U-Net flow surrogate on gfx950.
//
// The surrogate predicts the averaged steady fields (ux, uy, uz, rho) of a box
// from its voxel masks; the prediction seeds the lattice at equilibrium
// instead of the reference's rest state f = w_q (engine.js:260-261).
//
// Tensors are NDHWC bf16: voxel (n, z, y, x) of an N x D x H x W grid sits at
// ((n*D + z)*H + y)*W + x, its channels contiguous.  Every tensor argument has
// its own per-voxel stride, so a channel slice of a wider tensor (one half of
// a skip concatenation) is read and written in place.  Channel counts of the
// convolutions are multiples of 16, strides and channel offsets multiples of 8
// (16-byte vector accesses).
//
// Convolutions ('same' zero padding, 3x3x3 or 1x1x1) are implicit GEMMs on the
// matrix cores, v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   forward   Y[v][co]  = sum_{tap,ci} W[co][tap][ci] X[v+tap][ci]   rows co,  cols voxels, k (tap,ci)
//   input     dX        = forward over dZ with W'[ci][tap][co] = W[co][26-tap][ci]
//   weights   dW[co][tap][ci] = sum_v dZ[v][co] X[v+tap][ci]        rows co,  cols ci,     k voxels
// A workgroup owns a 16x4x2 tile of output voxels; its input brick with the
// one-voxel halo (18x6x4) is staged in LDS once per channel chunk and read by
// all 27 taps.  Weight gradients are summed per workgroup in registers over a
// strided range of tiles and written as partial sums that one more kernel
// reduces, so the result does not depend on scheduling.
#include <utility>

#include "abx_common.hpp"

namespace abx {
namespace {

typedef uint16_t u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef u16 u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even; NaN stays NaN.
__host__ __device__ __forceinline__ u16 f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

__device__ __forceinline__ bf16x8 as_bf8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(a), as_bf8(b), c, 0, 0, 0);
}

constexpr int TX = 16, TY = 4, TZ = 2, TV = TX * TY * TZ;  // output tile, 128 voxels

template <int TAPS>
struct Brick {
  static constexpr int HALO = TAPS == 27 ? 1 : 0;
  static constexpr int X = TX + 2 * HALO, Y = TY + 2 * HALO, Z = TZ + 2 * HALO, V = X * Y * Z;
};

struct Tile {
  int n, z0, y0, x0;
};

__device__ __forceinline__ Tile tile_of(int t, int D, int H, int W) {
  const int ntx = (W + TX - 1) / TX, nty = (H + TY - 1) / TY, ntz = (D + TZ - 1) / TZ;
  Tile r;
  r.x0 = (t % ntx) * TX;
  t /= ntx;
  r.y0 = (t % nty) * TY;
  t /= nty;
  r.z0 = (t % ntz) * TZ;
  r.n = t / ntz;
  return r;
}

__device__ __forceinline__ int64_t vox(const Tile& t, int D, int H, int W, int z, int y, int x) {
  return (((int64_t)t.n * D + z) * H + y) * W + x;
}

// Stage channels [c0, c0+ck) of the tile's brick (zero outside the grid) as
// s[brick voxel * sp + c].
template <int TAPS>
__device__ __forceinline__ void load_brick(u16* s, int sp, const u16* __restrict__ x, int xs, const Tile& tl, int D,
                                           int H, int W, int c0, int ck) {
  using Bk = Brick<TAPS>;
  const int groups = ck >> 3;
  for (int e = threadIdx.x; e < Bk::V * groups; e += blockDim.x) {
    const int vb = e / groups, g = e - vb * groups;
    const int bx = vb % Bk::X, r = vb / Bk::X, by = r % Bk::Y, bz = r / Bk::Y;
    const int gx = tl.x0 + bx - Bk::HALO, gy = tl.y0 + by - Bk::HALO, gz = tl.z0 + bz - Bk::HALO;
    u32x4 v = {0u, 0u, 0u, 0u};
    if ((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D)
      v = *(const u32x4*)(x + vox(tl, D, H, W, gz, gy, gx) * xs + c0 + 8 * g);
    *(u32x4*)(s + vb * sp + 8 * g) = v;
  }
}

// ---------------------------------------------------------------------------
// Forward convolution (also the input gradient with repacked weights).
// grid (tiles, Co / (16*NCT)), 256 threads; wave w computes voxel rows
// j = 2w, 2w+1 (16 voxels each) for NCT output-channel tiles of 16.
// wt: bf16 [Co][TAPS][Ci].  CK: channels per LDS chunk (divides Ci); the
// brick takes (CK + 8) * 2 bytes per voxel, so small chunks leave room for
// more workgroups per CU.
// ReLU backward folded into a convolution's store: each of 4 packed bf16
// values is zeroed where the gate value (the layer's ReLU output) is not
// positive, as abx_relu_backward does.
__device__ __forceinline__ u32x2 gate4(u32x2 o, u32x2 g) {
  const uint32_t m0 = bf2f((u16)(g.x & 0xffffu)) > 0.f ? 0xffffu : 0u, m1 = bf2f((u16)(g.x >> 16)) > 0.f ? 0xffffu : 0u;
  const uint32_t m2 = bf2f((u16)(g.y & 0xffffu)) > 0.f ? 0xffffu : 0u, m3 = bf2f((u16)(g.y >> 16)) > 0.f ? 0xffffu : 0u;
  return (u32x2){o.x & (m0 | m1 << 16), o.y & (m2 | m3 << 16)};
}

template <int TAPS, int NCT, int CK>
__global__ __launch_bounds__(256) void k_conv_fwd(const u16* __restrict__ x, int xs, const u16* __restrict__ wt,
                                                  const float* __restrict__ bias, u16* __restrict__ y, int ys,
                                                  const u16* __restrict__ gate, int gs, int D, int H, int W, int Ci,
                                                  int Co, int relu, int accum) {
  using Bk = Brick<TAPS>;
  constexpr int sp = CK + 8, ck = CK;
  __shared__ __attribute__((aligned(16))) u16 sx[Bk::V * sp];
  const Tile tl = tile_of(blockIdx.x, D, H, W);
  const int co0 = blockIdx.y * 16 * NCT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  f32x4 acc[NCT][2];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) acc[ct][jj] = (f32x4){0.f, 0.f, 0.f, 0.f};

  for (int c0 = 0; c0 < Ci; c0 += ck) {
    __syncthreads();
    load_brick<TAPS>(sx, sp, x, xs, tl, D, H, W, c0, ck);
    __syncthreads();
    const int K = TAPS * ck;
    for (int kk = 0; kk < K; kk += 32) {
      const int k = kk + 8 * lg;
      const int tap = k / ck, cil = k - tap * ck;
      u32x4 av[NCT], bv[2];
      if (tap < TAPS) {
        const int dz = TAPS == 27 ? tap / 9 : 0, dy = TAPS == 27 ? (tap / 3) % 3 : 0, dx = TAPS == 27 ? tap % 3 : 0;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int j = 2 * wave + jj;
          const int vb = (((j >> 2) + dz) * Bk::Y + ((j & 3) + dy)) * Bk::X + lr + dx;
          bv[jj] = *(const u32x4*)(sx + vb * sp + cil);
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
          av[ct] = *(const u32x4*)(wt + ((int64_t)(co0 + ct * 16 + lr) * TAPS + tap) * Ci + c0 + cil);
      } else {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) bv[jj] = (u32x4){0u, 0u, 0u, 0u};
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) av[ct] = (u32x4){0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[ct][jj] = mfma16(av[ct], bv[jj], acc[ct][jj]);
    }
  }

  // C element i of lane l: output channel 4*lg + i of the tile, voxel lr of row j
#pragma unroll
  for (int jj = 0; jj < 2; ++jj) {
    const int j = 2 * wave + jj;
    const int gx = tl.x0 + lr, gy = tl.y0 + (j & 3), gz = tl.z0 + (j >> 2);
    if (gx >= W || gy >= H || gz >= D) continue;
    u16* yp = y + vox(tl, D, H, W, gz, gy, gx) * ys;
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int co = co0 + ct * 16 + 4 * lg;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[ct][jj][i] + (bias ? bias[co + i] : 0.f);
      if (accum) {
        const u32x2 old = *(const u32x2*)(yp + co);
        v[0] += bf2f((u16)(old.x & 0xffffu));
        v[1] += bf2f((u16)(old.x >> 16));
        v[2] += bf2f((u16)(old.y & 0xffffu));
        v[3] += bf2f((u16)(old.y >> 16));
      }
      if (relu)
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = v[i] > 0.f ? v[i] : 0.f;
      u32x2 o;
      o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
      o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
      if (gate) o = gate4(o, *(const u32x2*)(gate + vox(tl, D, H, W, gz, gy, gx) * gs + co));
      *(u32x2*)(yp + co) = o;
    }
  }
}

// ---------------------------------------------------------------------------
// Forward convolution, 3x3x3 taps, one 16-channel output tile per workgroup
// (every 3x3x3 layer: on the 16-channel layers k_conv_fwd has one MFMA per
// weight load, and on the small deep levels its wide channel tiles leave too
// few workgroups).  A tile is 16 x 4 x 8 (x, y, z) output voxels, eight
// waves; wave w computes z-slice w, four rows of 16 voxels.  Input channels are taken 16 at a time:
// the k-steps of 32 then cover a pair of taps each (k = 32 s + 8 lg: tap A of
// pair s for lg < 2, tap B for lg >= 2, channels 8 (lg % 2) .. + 8), 14
// pairs for 27 taps (the last half empty), with one weight fragment per pair.
// The brick and the 14 weight fragments (14 KB) are loaded straight into LDS
// (global_load_lds), so no registers stage them, and each wave then reads the
// weight fragments into 56 registers once per chunk: 32 LDS fragment reads
// per wave and tile instead of 88 (the weights read from LDS for each of the
// 56 MFMAs).  124 registers leave two 8-wave workgroups per CU (4 waves per
// SIMD), whose brick loads overlap each other's MFMAs.  Measured against the
// weights read from LDS (73 registers, three workgroups per CU) and against
// brick and weights staged through registers (98, two per CU; with the
// weights in registers as well only one fitted): 57, 58 and 64 us for 16 ->
// 16 channels at 96x48x64, 153, 175 and 182 us for 48 -> 16.  Eight
// z-slices per workgroup share one brick (halo 2.1 voxels loaded per output
// voxel, 2.5 with four) and one copy of the weights.
//
// The pairs are chosen so that one input fragment read from the LDS brick
// serves several MFMAs: pairs 0-8 are the x shifts 0 and 1 of tap row
// (dz, dy) = (s / 3, s % 3), pairs 9-11 the x shift 2 of dz = 0 and 1 at
// dy = s - 9, pair 12 the x shift 2 of dz = 2 at dy = 0 and 1, pair 13 the
// last tap.  A fragment of brick row rho (pairs 0-11) is then the same for
// every output row r = rho - dy and serves up to three MFMAs: 32 fragment
// reads per wave and tile instead of one per MFMA (56).
//
// Workgroups are persistent over tiles.  With 16 input channels the weights
// are loaded once per workgroup (else once per 16-channel chunk, between the
// brick's two barriers); each thread owns fixed slots of the brick
// (decoded once).  The biases are staged in LDS once per workgroup and
// read there by each epilogue, so it issues no global load that it has to
// wait for (unless it accumulates or is gated); its addresses are a scalar
// base per tile plus one 32-bit lane offset.
constexpr int F16X = 16, F16Y = 4, F16Z = 8;
constexpr int F16BX = F16X + 2, F16BY = F16Y + 2, F16BZ = F16Z + 2, F16BV = F16BX * F16BY * F16BZ;
constexpr int F16SP = 16;  // u16 per brick voxel: 16 channels, unpadded (measured faster than 24, 32, 40)
constexpr int F16SLOTS = (2 * F16BV + 511) / 512;  // 16-byte brick pieces per thread

// Taps of pair s (27: none); tap = (dz * 3 + dy) * 3 + dx.
__host__ __device__ constexpr int f16_tapA(int s) {
  return s < 9 ? s * 3 : (s < 12 ? (s - 9) * 3 + 2 : (s == 12 ? 20 : 26));
}
__host__ __device__ constexpr int f16_tapB(int s) {
  return s < 9 ? s * 3 + 1 : (s < 12 ? 9 + (s - 9) * 3 + 2 : (s == 12 ? 23 : 27));
}

// Fragment read i (0-31) of a wave's tile: brick voxel offsets of its lanes
// with lg < 2 (lo) and lg >= 2 (hi, -1: zero), relative to the wave's own
// output voxel, and its MFMAs (pair w, output row r; w = -1: none).
struct F16Read {
  int lo, hi, w[3], r[3];
};
__host__ __device__ constexpr F16Read f16_read(int i) {
  constexpr int Z = F16BY * F16BX;
  F16Read f{0, -1, {-1, -1, -1}, {0, 0, 0}};
  if (i < 24) {
    // pairs 0-8 (dz = i / 6, x shifts 0 / 1) or 9-11 (x shift 2, dz 0 / 1):
    // brick row rho, serving output rows rho - dy
    const int rho = i < 18 ? i % 6 : i - 18;
    f.lo = i < 18 ? (i / 6) * Z + rho * F16BX : rho * F16BX + 2;
    f.hi = i < 18 ? f.lo + 1 : Z + rho * F16BX + 2;
    for (int dy = 0; dy < 3; ++dy)
      if (rho - dy >= 0 && rho - dy < F16Y) {
        f.w[dy] = i < 18 ? (i / 6) * 3 + dy : 9 + dy;
        f.r[dy] = rho - dy;
      }
  } else if (i < 28) {
    // pair 12: dz = 2, x shift 2, rows r (dy = 0) and r + 1 (dy = 1)
    const int r = i - 24;
    f.lo = 2 * Z + r * F16BX + 2;
    f.hi = f.lo + F16BX;
    f.w[0] = 12;
    f.r[0] = r;
  } else {
    // pair 13: tap 26 (dz = dy = 2, x shift 2) alone
    const int r = i - 28;
    f.lo = 2 * Z + (r + 2) * F16BX + 2;
    f.w[0] = 13;
    f.r[0] = r;
  }
  return f;
}
constexpr int F16READS = 32;

template <int... I, typename Fn>
__device__ __forceinline__ void f16_unroll(std::integer_sequence<int, I...>, Fn&& fn) {
  (fn(std::integral_constant<int, I>{}), ...);
}

struct F16Tile {
  int n, x0, y0, z0;
};

__device__ __forceinline__ F16Tile f16_tile(int t, int D, int H, int W) {
  const int ntx = (W + F16X - 1) / F16X, nty = (H + F16Y - 1) / F16Y, ntz = (D + F16Z - 1) / F16Z;
  F16Tile r;
  r.x0 = (t % ntx) * F16X;
  t /= ntx;
  r.y0 = (t % nty) * F16Y;
  t /= nty;
  r.z0 = (t % ntz) * F16Z;
  r.n = t / ntz;
  return r;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// max(v, 0) on the bits: one integer max (the float max would first quiet
// NaNs); -0 and negative NaNs give +0, positive NaNs pass.
__device__ __forceinline__ float relu_bits(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, b > 0 ? b : 0);
}

// Two floats to packed bf16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Sixteen zero bytes: the source of the brick pieces outside the grid when the
// brick is loaded straight into LDS.
__device__ u32x4 g_zero16[1];

// 16-byte load from global memory straight into LDS: lane l of the wave
// writes bytes [16 l, 16 l + 16) after lds (wave-uniform).
__device__ __forceinline__ void load_lds16(const void* src, void* lds) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                   (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}

// Channel parts: input chunks c < nch1 come from x, the others from x2
// (chunk c - nch1 there); output tiles blockIdx.y < ncot1 go to y, the others
// to y2 (tile blockIdx.y - ncot1), so a concatenation can be two tensors.
// DMA: the brick and the weights go from global memory straight into LDS
// (global_load_lds_dwordx4; brick piece e of the layout below is LDS byte
// 16 e, so the 64 pieces of a wave's slot are one contiguous kilobyte), with
// no staging registers; else through registers, the next chunk's loads issued
// before the barrier that ends this chunk's MFMAs.  F16DMA_WAVES: waves per
// SIMD the register allocation must allow.
// WREG: the 14 weight fragments are held in registers (56 of them), not read
// from LDS for every MFMA.  GST (with DMA, gated calls): the tile's gate
// values arrive in LDS with the brick of its last input chunk, so the
// epilogue does not wait on a global load per row.
template <bool DMA, int F16DMA_WAVES = 6, bool WREG = false, bool GST = false>
__global__ __launch_bounds__(512, DMA ? F16DMA_WAVES : 1) void k_conv_fwd16(
    const u16* __restrict__ x, int xs, const u16* x2, int xs2, int nch1, const u16* __restrict__ wt,
    const float* __restrict__ bias, u16* y, int ys, u16* y2, int ys2, int ncot1, const u16* __restrict__ gate, int gs,
    int D, int H, int W, int Ci, int ntiles, int relu, int accum) {
  // whole slots of 64 pieces: the last wave of the last slot writes past the brick
  __shared__ __attribute__((aligned(16))) u16 sx[(2 * F16BV + 63) / 64 * 64 * 8];
  __shared__ u32x4 sw[14 * 64];
  // GST: the gate values of the output tile, voxel v = (z 4 + y) 16 + x at 32 v bytes
  __shared__ __attribute__((aligned(16))) u16 sgate[GST ? F16X * F16Y * F16Z * 16 : 8];
  static_assert(!GST || DMA, "gate staging rides on the direct loads");
  const int co0 = blockIdx.y * 16, nch = Ci / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4, hi = lg >> 1;
  const int lbase = (wave * F16BY * F16BX + lr) * F16SP + 8 * (lg & 1);
  const int bx1 = lbase + (hi ? F16SP : 0), bxy = lbase + (hi ? F16BX * F16SP : 0),
            bxz = lbase + (hi ? F16BY * F16BX * F16SP : 0);

  // brick slots: piece e = thread + 512 m is voxel e / 2 (x fastest),
  // channels 8 (e % 2) .. + 8; two slot codes per register, bx | by << 5 |
  // bz << 8 | g << 12 in bits 16 (m % 2) .., bit 13 set past the brick
  uint32_t sb2[(F16SLOTS + 1) / 2] = {};
#pragma unroll
  for (int m = 0; m < F16SLOTS; ++m) {
    const int e = threadIdx.x + 512 * m, vb = e >> 1;
    const int bx = vb % F16BX, r = vb / F16BX, by = r % F16BY, bz = r / F16BY;
    const uint32_t code = e < 2 * F16BV ? (uint32_t)(bx | by << 5 | bz << 8 | (e & 1) << 12) : 0x2000u;
    sb2[m >> 1] |= code << (16 * (m & 1));
  }
  // slot m's code, or -1 past the brick
  auto slot_code = [](uint32_t w, int m) -> int {
    const uint32_t c = (w >> (16 * (m & 1))) & 0x3fffu;
    return (c & 0x2000u) ? -1 : (int)c;
  };
  int sb[F16SLOTS];  // the staged path's codes, decoded once
#pragma unroll
  for (int m = 0; m < F16SLOTS; ++m) sb[m] = DMA ? 0 : slot_code(sb2[m >> 1], m);
  // a buffer resource based at the brick's origin voxel (z0 - 1, y0 - 1,
  // x0 - 1; it may lie before the tensor, only in-grid pieces are read):
  // each piece is a 32-bit byte offset from it (below 2^31, checked by the
  // host), and pieces outside the grid take offset 2^31, beyond the
  // resource's range, which the hardware reads as zeros
  auto fetch_lds = [&](int t, int c) {
    const F16Tile tl = f16_tile(t, D, H, W);
    const bool p2 = c >= nch1;
    const u16* xb = p2 ? x2 : x;
    const int xsb = p2 ? xs2 : xs, cb = p2 ? c - nch1 : c;
    const char* org =
        (const char*)xb + ((((int64_t)tl.n * D + tl.z0 - 1) * H + tl.y0 - 1) * W + tl.x0 - 1) * xsb * 2 + 32 * cb;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)org, (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
    for (int m = 0; m < F16SLOTS; ++m) {
      if (512 * m + 64 * wave >= 2 * F16BV) continue;  // the wave's pieces are all past the brick
      // decode the slot here: hoisted out of the tile loop, its 15 fields
      // were spilled at six waves per SIMD
      uint32_t sw2 = sb2[m >> 1];
      asm volatile("" : "+v"(sw2));
      const int sv = slot_code(sw2, m);
      const int bx = sv & 31, by = (sv >> 5) & 7, bz = (sv >> 8) & 15;
      const bool in = sv >= 0 && (unsigned)(tl.x0 + bx - 1) < (unsigned)W &&
                      (unsigned)(tl.y0 + by - 1) < (unsigned)H && (unsigned)(tl.z0 + bz - 1) < (unsigned)D;
      const uint32_t off = in ? (uint32_t)(((bz * H + by) * W + bx) * xsb * 2 + 16 * (sv >> 12)) : 0x80000000u;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(sx + (512 * m + 64 * wave) * 8),
                                               16, off, 0, 0, 0);
    }
  };
  // the gate tile (two 16-byte pieces per voxel) through a buffer resource
  // based at the tile's first voxel, out-of-grid voxels at offset 2^31
  auto fetch_gate = [&](int t) {
    const F16Tile tl = f16_tile(t, D, H, W);
    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(gate + ((((int64_t)tl.n * D + tl.z0) * H + tl.y0) * W + tl.x0) * gs + co0), (short)0, 0x7fffffff,
        0x00020000);
    // derived here: hoisted out of the tile loop, the offsets spilled
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int e = tid + 512 * m, v = e >> 1;
      const int vx = v & 15, vy = (v >> 4) & 3, vz = v >> 6;
      const bool in = tl.x0 + vx < W && tl.y0 + vy < H && tl.z0 + vz < D;
      const uint32_t off = in ? (uint32_t)(((vz * H + vy) * W + vx) * gs * 2 + 16 * (e & 1)) : 0x80000000u;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rg, (__attribute__((address_space(3))) void*)(sgate + (512 * m + 64 * wave) * 8), 16, off, 0, 0, 0);
    }
  };
  auto load_weights_lds = [&](int c) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      if (512 * m + 64 * wave >= 14 * 64) continue;
      const int s = 8 * m + wave, l_r = lane & 15, l_g = lane >> 4;
      const int tap = (l_g >> 1) ? f16_tapB(s) : f16_tapA(s);
      const void* src = g_zero16;
      if (tap < 27) src = wt + ((int64_t)(co0 + l_r) * 27 + tap) * Ci + 16 * c + 8 * (l_g & 1);
      load_lds16(src, sw + 512 * m + 64 * wave);
    }
  };
  u32x4 pf[F16SLOTS];  // unused with DMA
  auto fetch = [&](int t, int c) {
    const F16Tile tl = f16_tile(t, D, H, W);
    const bool p2 = c >= nch1;
    const u16* xb = p2 ? x2 : x;
    const int xsb = p2 ? xs2 : xs, cb = p2 ? c - nch1 : c;
#pragma unroll
    for (int m = 0; m < F16SLOTS; ++m) {
      pf[m] = (u32x4){0u, 0u, 0u, 0u};
      if (sb[m] < 0) continue;
      const int gx = tl.x0 + (sb[m] & 31) - 1, gy = tl.y0 + ((sb[m] >> 5) & 7) - 1,
                gz = tl.z0 + ((sb[m] >> 8) & 15) - 1;
      if ((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D)
        pf[m] = *(const u32x4*)(xb + ((((int64_t)tl.n * D + gz) * H + gy) * W + gx) * xsb + 16 * cb +
                                8 * (sb[m] >> 12));
    }
  };
  auto load_weights = [&](int c) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = threadIdx.x + 512 * m;
      if (e < 14 * 64) {
        const int s = e >> 6, ln = e & 63, l_r = ln & 15, l_g = ln >> 4;
        const int tap = (l_g >> 1) ? f16_tapB(s) : f16_tapA(s);
        sw[e] = tap < 27 ? *(const u32x4*)(wt + ((int64_t)(co0 + l_r) * 27 + tap) * Ci + 16 * c + 8 * (l_g & 1))
                         : (u32x4){0u, 0u, 0u, 0u};
      }
    }
  };

  // the tile's biases, read in each epilogue (after the tile's barriers)
  __shared__ __attribute__((aligned(16))) float sbias[16];
  if (threadIdx.x < 16) sbias[threadIdx.x] = bias ? bias[co0 + threadIdx.x] : 0.f;
  // with WREG the weights still arrive in LDS and are read into registers once per chunk
  u32x4 wr[WREG ? 14 : 1];
  auto read_weights = [&]() {
#pragma unroll
    for (int s = 0; s < (WREG ? 14 : 0); ++s) wr[s] = sw[s * 64 + lane];
  };
  if (nch == 1) {
    if constexpr (DMA) {
      load_weights_lds(0);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
      load_weights(0);
    }
    __syncthreads();
    read_weights();
  }
  // C element i of lane l: output channel co0 + 4 lg + i, voxel lr of row r
  const int co = co0 + 4 * lg;
  // output: the lane's offset from a tile's first voxel (its z-slice, voxel
  // lr, channels), below 2^30 (checked by the host)
  const bool o2 = (int)blockIdx.y >= ncot1;
  u16* const yb = o2 ? y2 : y;
  const int ysb = o2 ? ys2 : ys;
  const uint32_t yoffb =
      2u * (uint32_t)((wave * H * W + lr) * ysb + (o2 ? ((int)blockIdx.y - ncot1) * 16 : co0) + 4 * lg);
  const uint32_t yrowb = 2u * (uint32_t)(W * ysb);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    f32x4 acc[F16Y];
#pragma unroll
    for (int r = 0; r < F16Y; ++r) acc[r] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < nch; ++c) {
      if constexpr (DMA) {
        __syncthreads();  // the previous chunk's MFMAs are done with the brick and the weights
        fetch_lds(t, c);
        if (nch > 1) load_weights_lds(c);
        if constexpr (GST)
          if (c == nch - 1) fetch_gate(t);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      } else {
        fetch(t, c);
        __syncthreads();  // the previous chunk's MFMAs are done with the brick
#pragma unroll
        for (int m = 0; m < F16SLOTS; ++m)
          if (sb[m] >= 0) {
            const int vb = ((sb[m] >> 8) & 15) * F16BY * F16BX + ((sb[m] >> 5) & 7) * F16BX + (sb[m] & 31);
            *(u32x4*)(sx + vb * F16SP + 8 * (sb[m] >> 12)) = pf[m];
          }
        if (nch > 1) load_weights(c);
      }
      __syncthreads();
      if (nch > 1) read_weights();
      // fragment read i + 2 is issued before the MFMAs of read i, so LDS
      // latency is not waited out at every read; scheduling barriers keep the
      // compiler from sinking the reads below those MFMAs
      // the hi lanes' voxel is the lo lanes' plus 1, a plane or a row: one
      // lane base per case, and the read's own offset is a constant
      u32x4 fr[3];
      auto read = [&](auto ic) {
        constexpr F16Read f = f16_read(decltype(ic)::value);
        constexpr int d = f.hi - f.lo;
        static_assert(f.hi < 0 || d == 1 || d == F16BY * F16BX || d == F16BX, "hi lanes: x, z or y neighbour");
        if constexpr (f.hi < 0) {
          const u32x4 v = *(const u32x4*)(sx + lbase + f.lo * F16SP);
          return hi ? (u32x4){0u, 0u, 0u, 0u} : v;
        } else {
          const int b = d == 1 ? bx1 : (d == F16BX ? bxy : bxz);
          return *(const u32x4*)(sx + b + f.lo * F16SP);
        }
      };
      fr[0] = read(std::integral_constant<int, 0>{});
      fr[1] = read(std::integral_constant<int, 1>{});
      f16_unroll(std::make_integer_sequence<int, F16READS>{}, [&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (i + 2 < F16READS) fr[(i + 2) % 3] = read(std::integral_constant<int, i + 2>{});
        __builtin_amdgcn_sched_barrier(0);
        constexpr F16Read f = f16_read(i);
        u32x4 wa[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (f.w[k] >= 0) wa[k] = WREG ? wr[WREG ? f.w[k] : 0] : sw[f.w[k] * 64 + lane];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (f.w[k] >= 0) acc[f.r[k]] = mfma16(wa[k], fr[i % 3], acc[f.r[k]]);
        __builtin_amdgcn_sched_barrier(0);
      });
    }

    // a scalar base per tile (its first voxel) and the lane's 32-bit offset
    const F16Tile tl = f16_tile(t, D, H, W);
    const int64_t v0 = (((int64_t)tl.n * D + tl.z0) * H + tl.y0) * W + tl.x0;
    char* const ytile = (char*)(yb + v0 * ysb);
    const bool lane_ok = tl.x0 + lr < W && tl.z0 + wave < D;
    // the row offsets derive from yo here: hoisted out of the tile loop they
    // held four 64-bit registers
    uint32_t yo = yoffb;
    asm volatile("" : "+v"(yo));
    const f32x4 bv = *(const f32x4*)(sbias + 4 * lg);
#pragma unroll
    for (int r = 0; r < F16Y; ++r) {
      if (tl.y0 + r >= H) break;
      if (!lane_ok) continue;
      u16* yp = (u16*)(ytile + (yo + (uint32_t)r * yrowb));
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[r][i] + bv[i];
      if (accum) {
        const u32x2 old = *(const u32x2*)yp;
        v[0] += bf2f((u16)(old.x & 0xffffu));
        v[1] += bf2f((u16)(old.x >> 16));
        v[2] += bf2f((u16)(old.y & 0xffffu));
        v[3] += bf2f((u16)(old.y >> 16));
      }
      if (relu)
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = relu_bits(v[i]);
      u32x2 o;
      o.x = pack_bf16(v[0], v[1]);
      o.y = pack_bf16(v[2], v[3]);
      // gated calls of the register-weight variant take GST, so its other
      // instantiation carries no global gate path
      if constexpr (GST)
        o = gate4(o, *(const u32x2*)(sgate + ((wave * F16Y + r) * F16X + lr) * 16 + 4 * lg));
      else if (!(DMA && WREG) && gate)
        o = gate4(o, *(const u32x2*)(gate + v0 * gs + (uint32_t)(((wave * H + r) * W + lr) * gs + co)));
      *(u32x2*)yp = o;
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: partial sums over a strided range of tiles.
// grid (S, co blocks * ci chunks * tap groups), 256 threads.  A workgroup owns
// output channels [co0, co0+16*NCT), input channels [ci0, ci0+16*NCI) and taps
// [t0, t0+tg); its tg*NCT*NCI tiles of 16x16 are dealt to the 4 waves, at most
// 8 each.  part: fp32 rows [S][Co TAPS Ci + Co]: the weight partials
// [Co][TAPS][Ci], then the bias partials [Co] (written by the workgroups of ci
// chunk 0 and the last tap group when partb, a flag, is not null).
// Bias gradient on the matrix cores: in the workgroups of input-channel chunk
// 0 and the last tap group (the group with the fewest taps), wave
// 4 - NCT + c multiplies the dZ fragment of output-channel tile c by a
// fragment of bf16 ones, so every column of its accumulator sums dZ over the
// voxels.  Waves are dealt items w, w + 4, ..., so the last waves have the
// fewest and the extra MFMA does not lengthen the longest wave.
template <int NCT>
__device__ __forceinline__ f32x4 bias_mfma(const u32x4 (&av)[NCT], int ct, f32x4 accb) {
  const u32x4 ones = {0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u};
  u32x4 a = av[0];
#pragma unroll
  for (int c = 1; c < NCT; ++c)
    if (ct == c) a = av[c];
  return mfma16(a, ones, accb);
}

// Column 0 of the bias accumulator: element i of lane group lg is output
// channel 4 lg + i of the tile.
__device__ __forceinline__ void bias_store(float* pb, f32x4 accb, int lg) {
#pragma unroll
  for (int i = 0; i < 4; ++i) pb[4 * lg + i] = accb[i];
}

template <int TAPS, int NCT, int NCI>
__global__ __launch_bounds__(256) void k_conv_wgrad(const u16* __restrict__ x, int xs, const u16* __restrict__ dz,
                                                    int dzs, float* __restrict__ part, float* __restrict__ partb,
                                                    int D, int H, int W, int Ci, int Co, int ntiles, int tg) {
  using Bk = Brick<TAPS>;
  constexpr int CO = 16 * NCT, CI = 16 * NCI, SZP = CO + 8, SXP = CI + 8, MAXW = 8;
  __shared__ __attribute__((aligned(16))) u16 sz[TV * SZP];
  __shared__ __attribute__((aligned(16))) u16 sx[Bk::V * SXP];
  const int nci = Ci / CI, ngr = (TAPS + tg - 1) / tg;
  int r = blockIdx.y;
  const int g = r % ngr;
  r /= ngr;
  const int cc = r % nci, cb = r / nci;
  const int co0 = cb * CO, ci0 = cc * CI, t0 = g * tg;
  const int ntap = min(tg, TAPS - t0);
  const int nq = ntap * NCT * NCI;
  const bool do_bias = partb && cc == 0 && g == ngr - 1;
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int bwt = wave - (4 - NCT);  // bias: output-channel tile of this wave
  const bool bias_wave = do_bias && bwt >= 0;
  f32x4 acc[MAXW];
#pragma unroll
  for (int u = 0; u < MAXW; ++u) acc[u] = (f32x4){0.f, 0.f, 0.f, 0.f};
  f32x4 accb = {0.f, 0.f, 0.f, 0.f};

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tile_of(t, D, H, W);
    __syncthreads();
    // dZ tile (zero outside the grid): sz[v * SZP + co]
    for (int e = threadIdx.x; e < TV * NCT * 2; e += blockDim.x) {
      const int v = e / (NCT * 2), q = e - v * (NCT * 2);
      const int gx = tl.x0 + (v & 15), gy = tl.y0 + ((v >> 4) & 3), gz = tl.z0 + (v >> 6);
      u32x4 val = {0u, 0u, 0u, 0u};
      if (gx < W && gy < H && gz < D) val = *(const u32x4*)(dz + vox(tl, D, H, W, gz, gy, gx) * dzs + co0 + 8 * q);
      *(u32x4*)(sz + v * SZP + 8 * q) = val;
    }
    load_brick<TAPS>(sx, SXP, x, xs, tl, D, H, W, ci0, CI);
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < TV / 32; ++s) {
      // k = voxels 32s + 8lg + e (e = 0..7), consecutive in x
      const int vb0 = 32 * s + 8 * lg;
      const int tx0 = vb0 & 15, ty = (vb0 >> 4) & 3, tz = vb0 >> 6;
      u32x4 av[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        u16x8 a;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = sz[(vb0 + e) * SZP + ct * 16 + lr];
        av[ct] = __builtin_bit_cast(u32x4, a);
      }
#pragma unroll
      for (int u = 0; u < MAXW; ++u) {
        const int q = wave + 4 * u;
        if (q >= nq) break;
        const int tap = t0 + q / (NCT * NCI), ct = (q / NCI) % NCT, it = q % NCI;
        const int dzb = TAPS == 27 ? tap / 9 : 0, dyb = TAPS == 27 ? (tap / 3) % 3 : 0,
                  dxb = TAPS == 27 ? tap % 3 : 0;
        const int base = ((tz + dzb) * Bk::Y + (ty + dyb)) * Bk::X + tx0 + dxb;
        u16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = sx[(base + e) * SXP + it * 16 + lr];
        u32x4 avc = av[0];
#pragma unroll
        for (int c = 1; c < NCT; ++c)
          if (ct == c) avc = av[c];
        acc[u] = mfma16(avc, __builtin_bit_cast(u32x4, b), acc[u]);
      }
      // after the items, where the fragments have arrived (no wait at the branch)
      if (bias_wave) accb = bias_mfma<NCT>(av, bwt, accb);
    }
  }

  // C element i of lane l: output channel 4*lg + i, input channel lr of the tile
  const int64_t nw = (int64_t)Co * TAPS * Ci;
  float* pp = part + (int64_t)blockIdx.x * (nw + Co);
#pragma unroll
  for (int u = 0; u < MAXW; ++u) {
    const int q = wave + 4 * u;
    if (q >= nq) break;
    const int tap = t0 + q / (NCT * NCI), ct = (q / NCI) % NCT, it = q % NCI;
    const int ci = ci0 + it * 16 + lr;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = co0 + ct * 16 + 4 * lg + i;
      pp[((int64_t)co * TAPS + tap) * Ci + ci] = acc[u][i];
    }
  }
  if (bias_wave && lr == 0) bias_store(pp + nw + co0 + bwt * 16, accb, lg);
}

// Weight gradient on gfx950's transposing LDS reads: the partial sums of
// k_conv_wgrad from the same voxel-major (NDHWC) staging, filled with 16-byte
// copies.  An MFMA fragment holds 8 consecutive voxels (k) of one channel.
// ds_read_b64_tr_b16 builds it from rows of the voxel-major tile: in each
// group of 16 lanes, lane s points at 4 consecutive channels (columns
// 4 (s & 3) .. +3) of row s / 4, and lane i receives column i of those 4 rows
// (4 voxels of channel i).  Two reads give a lane its 8 voxels; the x shift of
// a tap is a row offset.  Rows are not padded: padding them by 16 or 32 bytes
// measured no faster.
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

// hi_rows: the rows of the second read (elements 4-7) follow the first's by
// this many rows.
__device__ __forceinline__ u32x4 frag_tr(const u16* p, int pitch, int hi_rows = 4) {
  const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)(void*)p);
  const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)(void*)(p + hi_rows * pitch));
  const u32x2 l = __builtin_bit_cast(u32x2, lo), h = __builtin_bit_cast(u32x2, hi);
  return (u32x4){l.x, l.y, h.x, h.y};
}

// Taps per workgroup of the weight gradient: at most 32 (tap, channel tile)
// items, 8 per wave.
__host__ __device__ constexpr int wgrad_tg(int taps, int nct, int nci) {
  return taps < 32 / (nct * nci) ? taps : 32 / (nct * nci);
}

template <int TAPS, int NCT, int NCI>
__global__ __launch_bounds__(256) void k_conv_wgrad_tr(const u16* __restrict__ x, int xs, const u16* __restrict__ dz,
                                                       int dzs, float* __restrict__ part, float* __restrict__ partb,
                                                       int D, int H, int W, int Ci, int Co, int ntiles, int tg) {
  using Bk = Brick<TAPS>;
  constexpr int CO = 16 * NCT, CI = 16 * NCI;
  // items per wave, a compile-time count: the loop has no branches, so the
  // LDS reads of later items are issued ahead of the MFMAs of earlier ones; a
  // wave with fewer items computes the rest on item 0's operands and drops them
  constexpr int MAXW = (wgrad_tg(TAPS, NCT, NCI) * NCT * NCI + 3) / 4;
  constexpr int PZ = CO, PX = CI;  // row pitches of the dZ tile and the brick
  constexpr int ZITEMS = TV * NCT * 2, XITEMS = Bk::V * NCI * 2;
  constexpr int ZSL = (ZITEMS + 255) / 256, XSL = (XITEMS + 255) / 256;
  __shared__ __attribute__((aligned(16))) u16 sz[TV * PZ];
  __shared__ __attribute__((aligned(16))) u16 sx[Bk::V * PX];
  const int nci = Ci / CI, ngr = (TAPS + tg - 1) / tg;
  int r = blockIdx.y;
  const int g = r % ngr;
  r /= ngr;
  const int cc = r % nci, cb = r / nci;
  const int co0 = cb * CO, ci0 = cc * CI, t0 = g * tg;
  const int ntap = min(tg, TAPS - t0);
  const int nq = ntap * NCT * NCI;
  const bool do_bias = partb && cc == 0 && g == ngr - 1;
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // this wave's (tap, channel tile) items: brick offset of the tap plus the
  // input-channel tile's column, output-channel tile (wave-uniform)
  int boff[MAXW], bct[MAXW];
#pragma unroll
  for (int u = 0; u < MAXW; ++u) {
    const int q = wave + 4 * u, qq = q < nq ? q : 0;
    const int tap = t0 + qq / (NCT * NCI), it = qq % NCI;
    const int dzb = TAPS == 27 ? tap / 9 : 0, dyb = TAPS == 27 ? (tap / 3) % 3 : 0, dxb = TAPS == 27 ? tap % 3 : 0;
    boff[u] = ((dzb * Bk::Y + dyb) * Bk::X + dxb) * PX + it * 16;
    bct[u] = (qq / NCI) % NCT;
  }
  // this thread's staging slots (items e = thread + 256 m), decoded once; a
  // slot past the last item loads the last item and is not stored
  int zs[ZSL], xsl[XSL];
#pragma unroll
  for (int m = 0; m < ZSL; ++m) zs[m] = threadIdx.x + 256 * m;  // voxel e / (2 NCT), channel group e % (2 NCT)
#pragma unroll
  for (int m = 0; m < XSL; ++m) {
    const int e = min((int)threadIdx.x + 256 * m, XITEMS - 1);
    const int vb = e / (NCI * 2), q = e - vb * (NCI * 2);
    const int bx = vb % Bk::X, row = vb / Bk::X, by = row % Bk::Y, bz = row / Bk::Y;
    xsl[m] = bx | by << 5 | bz << 8 | q << 11;
  }
  const bool xlast = (int)threadIdx.x + 256 * (XSL - 1) < XITEMS;
  // staged values of one tile and their in-grid bits (bit m)
  u32x4 zv[ZSL], xv[XSL];
  uint32_t zok = 0, xok = 0;
  // every load from an address inside the grid (clamped); out-of-grid values
  // are zeroed when they are stored, so no wait sits between the loads.
  // (Loading the next tile while computing this one measured slower: its
  // registers cost occupancy.)
  auto fetch = [&](int t) {
    const Tile tl = tile_of(t, D, H, W);
    zok = 0;
    xok = 0;
#pragma unroll
    for (int m = 0; m < ZSL; ++m) {
      const int v = zs[m] / (NCT * 2), q = zs[m] - v * (NCT * 2);
      const int gx = tl.x0 + (v & 15), gy = tl.y0 + ((v >> 4) & 3), gz = tl.z0 + (v >> 6);
      zok |= (uint32_t)(gx < W && gy < H && gz < D) << m;
      zv[m] = *(const u32x4*)(dz + vox(tl, D, H, W, min(gz, D - 1), min(gy, H - 1), min(gx, W - 1)) * dzs + co0 +
                              8 * q);
    }
#pragma unroll
    for (int m = 0; m < XSL; ++m) {
      const int bx = xsl[m] & 31, by = (xsl[m] >> 5) & 7, bz = (xsl[m] >> 8) & 7, q = xsl[m] >> 11;
      const int gx = tl.x0 + bx - Bk::HALO, gy = tl.y0 + by - Bk::HALO, gz = tl.z0 + bz - Bk::HALO;
      xok |= (uint32_t)((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D) << m;
      xv[m] = *(const u32x4*)(x + vox(tl, D, H, W, min(max(gz, 0), D - 1), min(max(gy, 0), H - 1),
                                      min(max(gx, 0), W - 1)) * xs + ci0 + 8 * q);
    }
  };
  const u32x4 zero4 = {0u, 0u, 0u, 0u};
  // this lane's place in a transposing read: row lr / 4, columns 4 (lr % 4) ..
  const int lrow = lr >> 2, lcol = 4 * (lr & 3);
  const int bwt = wave - (4 - NCT);  // bias: output-channel tile of this wave
  const bool bias_wave = do_bias && bwt >= 0;
  f32x4 acc[MAXW];
#pragma unroll
  for (int u = 0; u < MAXW; ++u) acc[u] = (f32x4){0.f, 0.f, 0.f, 0.f};
  f32x4 accb = {0.f, 0.f, 0.f, 0.f};

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    fetch(t);
    __syncthreads();
    // dZ tile sz[v * PZ + co] and input brick sx[brick voxel * PX + ci]
#pragma unroll
    for (int m = 0; m < ZSL; ++m) {
      const int v = zs[m] / (NCT * 2), q = zs[m] - v * (NCT * 2);
      *(u32x4*)(sz + v * PZ + 8 * q) = ((zok >> m) & 1u) ? zv[m] : zero4;
    }
#pragma unroll
    for (int m = 0; m < XSL; ++m) {
      if (m == XSL - 1 && !xlast) break;
      const int bx = xsl[m] & 31, by = (xsl[m] >> 5) & 7, bz = (xsl[m] >> 8) & 7, q = xsl[m] >> 11;
      *(u32x4*)(sx + ((bz * Bk::Y + by) * Bk::X + bx) * PX + 8 * q) = ((xok >> m) & 1u) ? xv[m] : zero4;
    }
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < TV / 32; ++s) {
      // k = voxels 32s + 8lg + e (e = 0..7), consecutive in x
      const int vb0 = 32 * s + 8 * lg;
      const int tx0 = vb0 & 15, ty = (vb0 >> 4) & 3, tz = vb0 >> 6;
      u32x4 av[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) av[ct] = frag_tr(sz + (vb0 + lrow) * PZ + ct * 16 + lcol, PZ);
      const u16* lbase = sx + (((tz * Bk::Y + ty) * Bk::X + tx0) + lrow) * PX + lcol;
#pragma unroll
      for (int u = 0; u < MAXW; ++u) {
        const u32x4 b = frag_tr(lbase + boff[u], PX);
        u32x4 avc = av[0];
#pragma unroll
        for (int c = 1; c < NCT; ++c)
          if (bct[u] == c) avc = av[c];
        acc[u] = mfma16(avc, b, acc[u]);
      }
      if (bias_wave) accb = bias_mfma<NCT>(av, bwt, accb);
    }
  }

  // C element i of lane l: output channel 4*lg + i, input channel lr of the tile
  const int64_t nw = (int64_t)Co * TAPS * Ci;
  float* pp = part + (int64_t)blockIdx.x * (nw + Co);
#pragma unroll
  for (int u = 0; u < MAXW; ++u) {
    const int q = wave + 4 * u;
    if (q >= nq) break;
    const int tap = t0 + q / (NCT * NCI), ct = (q / NCI) % NCT, it = q % NCI;
    const int ci = ci0 + it * 16 + lr;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = co0 + ct * 16 + 4 * lg + i;
      pp[((int64_t)co * TAPS + tap) * Ci + ci] = acc[u][i];
    }
  }
  if (bias_wave && lr == 0) bias_store(pp + nw + co0 + bwt * 16, accb, lg);
}

// Weight gradient of a 3x3x3 convolution with every tap in one workgroup.
// The output tile is 16 x 4 x TZW voxels (2 TZW k steps of 32), its input
// brick 18 x 6 x (TZW + 2): TZW = 4, or 8 on large grids, where the deeper
// brick re-reads fewer halo planes.  Workgroup (blockIdx.x = partial row, blockIdx.y =
// block) owns output channels [co0, co0 + 16 NCT) and input channels
// [ci0, ci0 + 16 NCI) for all 27 taps, so each tile is staged once per
// block and not once per tap group.  Its 27 NCI (tap, input tile) fragments
// are dealt to the 8 waves (item q = wave + 8 u); each meets the NCT dZ
// fragments of the k step, so one B fragment read feeds NCT MFMAs.  The next
// tile's global loads are issued before this tile's MFMAs and land in
// registers, which the next iteration stores to LDS; or, where the channel
// blocking then fits more workgroups per CU (w27_wgs), the tiles are loaded
// straight into LDS and the other workgroups' MFMAs cover the loads (16 ->
// 16 channels at 96x48x64: 61 against 67 us).  Fragments are built
// with transposing LDS reads as in k_conv_wgrad_tr; partial rows
// [S][Co 27 Ci + Co] as there.  Bias partials come from the ci block 0
// workgroups: their wave 7 multiplies every dZ fragment by a fragment of
// ones (27 NCI is never a multiple of 8, so wave 7 has one item less than
// wave 0, and waves skip the MFMAs of an item they do not own).
constexpr int W27X = 16, W27Y = 4, W27BX = W27X + 2, W27BY = W27Y + 2;

__device__ __forceinline__ Tile w27_tile(int t, int D, int H, int W, int TZW) {
  const int ntx = (W + W27X - 1) / W27X, nty = (H + W27Y - 1) / W27Y, ntz = (D + TZW - 1) / TZW;
  Tile r;
  r.x0 = (t % ntx) * W27X;
  t /= ntx;
  r.y0 = (t % nty) * W27Y;
  t /= nty;
  r.z0 = (t % ntz) * TZW;
  r.n = t / ntz;
  return r;
}

// Channel parts as in k_conv_fwd16: input channels c < ci1 come from x, the
// others from x2 (channel c - ci1 there), chosen per 8-channel staging piece,
// so an input block may straddle ci1 (a multiple of 16).
// WGS > 0: the dZ tile and the brick go from global memory straight into LDS
// (piece e of either is LDS byte 16 e after its base), with no staging
// registers and no prefetch of the next tile, and WGS workgroups share a CU
// so that one's loads overlap the others' MFMAs; WGS = 0: staged through
// registers, the next tile's loads issued before this tile's MFMAs.
template <int NCT, int NCI, int TZW>
constexpr int w27_lds_bytes() {
  return (W27X * W27Y * TZW * 16 * NCT + ((W27BX * W27BY * (TZW + 2) * NCI * 2 + 63) / 64 * 64) * 8) * 2;
}

template <int NCT, int NCI, int TZW, int WGS>
__global__ __launch_bounds__(512, WGS > 0 ? 2 * WGS : 1) void k_wgrad27(
    const u16* __restrict__ x, int xs, const u16* x2, int xs2, int ci1, const u16* __restrict__ dz, int dzs,
    float* __restrict__ part, float* __restrict__ partb, int D, int H, int W, int Ci, int Co, int ntiles) {
  constexpr bool DMA = WGS > 0;
  constexpr int W27V = W27X * W27Y * TZW, W27BZ = TZW + 2, W27BV = W27BX * W27BY * W27BZ;
  constexpr int CO = 16 * NCT, CI = 16 * NCI, PZ = CO, PX = CI;
  constexpr int NI = 27 * NCI, MAXI = (NI + 7) / 8;
  constexpr int ZP = W27V * NCT * 2, XP = W27BV * NCI * 2;  // 16-byte pieces of the dZ tile and the brick
  constexpr int ZSL = (ZP + 511) / 512, XSL = (XP + 511) / 512;
  static_assert(ZP % 512 == 0, "dZ pieces fill whole slots");
  // the lane groups interleaved by 4 voxels, which gives the two groups of a
  // half-wave disjoint LDS banks on voxel rows of 32 bytes (groups 128 bytes
  // apart) and 96 bytes (384 apart: rows on banks {0-7, 24-31, 48-55, 8-15}
  // and {32-39, 56-63, 16-23, 40-47}, for every tap shift).  On 64-byte rows
  // (two channel tiles) every voxel offset lands on the same banks, so there
  // the two 16-channel halves of a voxel are exchanged where bit 2 of its x
  // (in the tile or the brick row) is set (SWZ): rows 4 apart, one in each
  // group, then read opposite halves.  128-byte rows (four output tiles) keep
  // the plain layout.
  // (the swizzle relies on the interleave: elements 4-7 come from rows 8 on,
  // whose bit 2 of x is that of elements 0-3)
  constexpr bool K4 = NCT <= 2;
  constexpr bool SWZ_Z = NCT == 2, SWZ_X = NCI == 2 && K4;
  __shared__ __attribute__((aligned(16))) u16 sz[W27V * PZ];
  // whole slots of 64 pieces: with direct loads the last wave of the last slot writes past the brick
  __shared__ __attribute__((aligned(16))) u16 sx[(XP + 63) / 64 * 64 * 8];
  static_assert(sizeof(sz) + sizeof(sx) == w27_lds_bytes<NCT, NCI, TZW>(), "LDS per workgroup");
  const int ncb = Co / CO;
  const int cb = (int)blockIdx.y % ncb, cc = (int)blockIdx.y / ncb;
  const int co0 = cb * CO, ci0 = cc * CI;
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  static_assert(NI % 8 != 0, "wave 7 has a spare item slot for the bias");
  const bool bias_wave = partb && cc == 0 && wave == 7;
  const bool last_item = wave + 8 * (MAXI - 1) < NI;
  const u32x4 ones = {0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u};

  // this wave's items: brick offset of the tap plus the input tile's column
  // (with SWZ_X the input tile's column in the lane's brick row, whose x is
  // 4 (lg & 1) + lane row + the tap's x shift)
  int boff[MAXI];
#pragma unroll
  for (int u = 0; u < MAXI; ++u) {
    const int q = wave + 8 * u, qq = q < NI ? q : 0;
    const int tap = qq / NCI, it = qq % NCI;
    const int itc = SWZ_X ? it ^ (((4 * (lg & 1) + (lr >> 2) + tap % 3) >> 2) & 1) : it;
    boff[u] = (((tap / 9) * W27BY + (tap / 3) % 3) * W27BX + tap % 3) * PX + itc * 16;
  }
  // staging slots of this thread: dZ piece threadIdx + 512 m is voxel
  // e / (2 NCT), channels 8 (e % (2 NCT)) .. + 8; brick piece likewise with
  // the brick voxel packed as bx | by << 5 | bz << 8 | piece << 12
  int xsl[XSL];
#pragma unroll
  for (int m = 0; m < XSL; ++m) {
    const int e = min((int)threadIdx.x + 512 * m, XP - 1);
    const int vb = e / (NCI * 2), q = e - vb * (NCI * 2);
    const int bx = vb % W27BX, row = vb / W27BX, by = row % W27BY, bz = row / W27BY;
    // the piece the slot holds: with SWZ_X the other half's where bit 2 of bx is set
    const int qg = SWZ_X ? q ^ (((bx >> 2) & 1) << 1) : q;
    xsl[m] = bx | by << 5 | bz << 8 | qg << 12;
  }
  const bool xlast = (int)threadIdx.x + 512 * (XSL - 1) < XP;
  u32x4 zv[ZSL], xv[XSL];
  uint32_t zok = 0, xok = 0;
  // loads from clamped in-grid addresses; out-of-grid values are zeroed when stored
  auto fetch = [&](int t) {
    const Tile tl = w27_tile(t, D, H, W, TZW);
    zok = 0;
    xok = 0;
#pragma unroll
    for (int m = 0; m < ZSL; ++m) {
      const int e = (int)threadIdx.x + 512 * m;
      const int v = e / (NCT * 2), q = e - v * (NCT * 2), qg = SWZ_Z ? q ^ ((((v & 15) >> 2) & 1) << 1) : q;
      const int gx = tl.x0 + (v & 15), gy = tl.y0 + ((v >> 4) & 3), gz = tl.z0 + (v >> 6);
      zok |= (uint32_t)(gx < W && gy < H && gz < D) << m;
      zv[m] = *(const u32x4*)(dz + vox(tl, D, H, W, min(gz, D - 1), min(gy, H - 1), min(gx, W - 1)) * dzs + co0 +
                              8 * qg);
    }
#pragma unroll
    for (int m = 0; m < XSL; ++m) {
      const int bx = xsl[m] & 31, by = (xsl[m] >> 5) & 7, bz = (xsl[m] >> 8) & 15, q = xsl[m] >> 12;
      const int gx = tl.x0 + bx - 1, gy = tl.y0 + by - 1, gz = tl.z0 + bz - 1;
      xok |= (uint32_t)((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D) << m;
      const int gc = ci0 + 8 * q;  // the piece's channel, in the first part or the second
      const bool p2 = gc >= ci1;
      xv[m] = *(const u32x4*)((p2 ? x2 : x) + vox(tl, D, H, W, min(max(gz, 0), D - 1), min(max(gy, 0), H - 1),
                                                  min(max(gx, 0), W - 1)) * (p2 ? xs2 : xs) + (p2 ? gc - ci1 : gc));
    }
  };
  // dZ tile through a buffer resource based at the tile's first voxel, the
  // brick through one based at its origin voxel when the input block lies in
  // one tensor: 32-bit byte offsets (below 2^31, checked by the host), and
  // pieces outside the grid at offset 2^31, read as zeros by the range check
  const bool x_one = ci1 >= ci0 + CI || ci1 <= ci0;
  auto fetch_lds = [&](int t) {
    const Tile tl = w27_tile(t, D, H, W, TZW);
    const __amdgpu_buffer_rsrc_t rz = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(dz + vox(tl, D, H, W, tl.z0, tl.y0, tl.x0) * dzs + co0), (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
    for (int m = 0; m < ZSL; ++m) {
      const int e = (int)threadIdx.x + 512 * m;
      const int v = e / (NCT * 2), q = e - v * (NCT * 2);
      const int vx = v & 15, vy = (v >> 4) & 3, vz = v >> 6;
      const int qg = SWZ_Z ? q ^ (((vx >> 2) & 1) << 1) : q;
      const bool in = tl.x0 + vx < W && tl.y0 + vy < H && tl.z0 + vz < D;
      const uint32_t off = in ? (uint32_t)(((vz * H + vy) * W + vx) * dzs * 2 + 16 * qg) : 0x80000000u;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rz, (__attribute__((address_space(3))) void*)(sz + (512 * m + 64 * wave) * 8),
                                               16, off, 0, 0, 0);
    }
    if (x_one) {
      const bool p2 = ci0 >= ci1;
      const u16* xb = p2 ? x2 : x;
      const int xsb = p2 ? xs2 : xs, cb = p2 ? ci0 - ci1 : ci0;
      const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(xb + ((((int64_t)tl.n * D + tl.z0 - 1) * H + tl.y0 - 1) * W + tl.x0 - 1) * xsb + cb), (short)0,
          0x7fffffff, 0x00020000);
#pragma unroll
      for (int m = 0; m < XSL; ++m) {
        if (512 * m + 64 * wave >= XP) continue;  // the wave's pieces are all past the brick
        int sv = xsl[m];
        asm volatile("" : "+v"(sv));
        const int bx = sv & 31, by = (sv >> 5) & 7, bz = (sv >> 8) & 15, q = sv >> 12;
        const bool in = (int)threadIdx.x + 512 * m < XP && (unsigned)(tl.x0 + bx - 1) < (unsigned)W &&
                        (unsigned)(tl.y0 + by - 1) < (unsigned)H && (unsigned)(tl.z0 + bz - 1) < (unsigned)D;
        const uint32_t off = in ? (uint32_t)(((bz * H + by) * W + bx) * xsb * 2 + 16 * q) : 0x80000000u;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rx, (__attribute__((address_space(3))) void*)(sx + (512 * m + 64 * wave) * 8), 16, off, 0, 0, 0);
      }
      return;
    }
    // an input block across the split of a two-part input: global addresses, the tensor chosen per piece
#pragma unroll
    for (int m = 0; m < XSL; ++m) {
      if (512 * m + 64 * wave >= XP) continue;  // the wave's pieces are all past the brick
      // decoded here, not hoisted out of the tile loop into registers
      int sv = xsl[m];
      asm volatile("" : "+v"(sv));
      const int bx = sv & 31, by = (sv >> 5) & 7, bz = (sv >> 8) & 15, q = sv >> 12;
      const int gx = tl.x0 + bx - 1, gy = tl.y0 + by - 1, gz = tl.z0 + bz - 1;
      const void* src = g_zero16;
      const bool in = (unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D;
      if (in && (int)threadIdx.x + 512 * m < XP) {
        const int gc = ci0 + 8 * q;
        const bool p2 = gc >= ci1;
        src = (p2 ? x2 : x) + vox(tl, D, H, W, gz, gy, gx) * (p2 ? xs2 : xs) + (p2 ? gc - ci1 : gc);
      }
      load_lds16(src, sx + (512 * m + 64 * wave) * 8);
    }
  };
  const u32x4 zero4 = {0u, 0u, 0u, 0u};
  const int lrow = lr >> 2, lcol = 4 * (lr & 3);
  f32x4 acc[MAXI][NCT];
#pragma unroll
  for (int u = 0; u < MAXI; ++u)
#pragma unroll
    for (int c = 0; c < NCT; ++c) acc[u][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  f32x4 accb[NCT];
#pragma unroll
  for (int c = 0; c < NCT; ++c) accb[c] = (f32x4){0.f, 0.f, 0.f, 0.f};

  int t = blockIdx.x;
  if constexpr (!DMA)
    if (t < ntiles) fetch(t);
  for (; t < ntiles; t += gridDim.x) {
    __syncthreads();  // the previous tile's fragment reads are done
    if constexpr (DMA) {
      fetch_lds(t);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
#pragma unroll
      for (int m = 0; m < ZSL; ++m) {
        const int e = (int)threadIdx.x + 512 * m;
        const int v = e / (NCT * 2), q = e - v * (NCT * 2);
        *(u32x4*)(sz + v * PZ + 8 * q) = ((zok >> m) & 1u) ? zv[m] : zero4;
      }
#pragma unroll
      for (int m = 0; m < XSL; ++m) {
        if (m == XSL - 1 && !xlast) break;
        // piece e sits at LDS byte 16 e (its slot code names the global piece it holds)
        *(u32x4*)(sx + 8 * ((int)threadIdx.x + 512 * m)) = ((xok >> m) & 1u) ? xv[m] : zero4;
      }
    }
    __syncthreads();
    if constexpr (!DMA)
      if (t + (int)gridDim.x < ntiles) fetch(t + (int)gridDim.x);
#pragma unroll 1
    for (int s = 0; s < W27V / 32; ++s) {
      // k step s = the 32 voxels of x rows 2 s and 2 s + 1.  Lane group lg
      // takes 8 of them: x 8 (lg & 1) .. + 7 of row 2 s + (lg >> 1) (elements
      // 0-3 and 4-7 from rows 4 apart), or with K4 x 4 (lg & 1) .. + 3 and
      // x 8 + 4 (lg & 1) .. + 3 (rows 8 apart), so that the two groups of a
      // half-wave read 4 voxels apart instead of 8
      const int vb0 = K4 ? 32 * s + 16 * (lg >> 1) + 4 * (lg & 1) : 32 * s + 8 * lg;
      const int tx0 = vb0 & 15, ty = (vb0 >> 4) & 3, tz = vb0 >> 6;
      u32x4 av[NCT];
#pragma unroll
      for (int c = 0; c < NCT; ++c)
        av[c] = frag_tr(sz + (vb0 + lrow) * PZ + (SWZ_Z ? c ^ (lg & 1) : c) * 16 + lcol, PZ, K4 ? 8 : 4);
      const u16* lbase = sx + ((tz * W27BY + ty) * W27BX + tx0 + lrow) * PX + lcol;
#pragma unroll
      for (int u = 0; u < MAXI; ++u) {
        const u32x4 b = frag_tr(lbase + boff[u], PX, K4 ? 8 : 4);
        if (u + 1 < MAXI || last_item)
#pragma unroll
          for (int c = 0; c < NCT; ++c) acc[u][c] = mfma16(av[c], b, acc[u][c]);
      }
      if (bias_wave)
#pragma unroll
        for (int c = 0; c < NCT; ++c) accb[c] = mfma16(av[c], ones, accb[c]);
    }
  }

  // C element i of lane l: output channel 4 lg + i, input channel lr of the tiles
  const int64_t nw = (int64_t)Co * 27 * Ci;
  float* pp = part + (int64_t)blockIdx.x * (nw + Co);
#pragma unroll
  for (int u = 0; u < MAXI; ++u) {
    const int q = wave + 8 * u;
    if (q >= NI) break;
    const int tap = q / NCI, ci = ci0 + (q % NCI) * 16 + lr;
#pragma unroll
    for (int c = 0; c < NCT; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) pp[((int64_t)(co0 + c * 16 + 4 * lg + i) * 27 + tap) * Ci + ci] = acc[u][c][i];
  }
  if (bias_wave && lr == 0)
#pragma unroll
    for (int c = 0; c < NCT; ++c) bias_store(pp + nw + co0 + c * 16, accb[c], lg);
}

// First pass of the partial-sum reduction, in place: row c * chunk of part
// becomes the sum of rows c * chunk .. c * chunk + chunk - 1 (the last chunk
// may be shorter), element by element in row order.  One thread per
// (element, chunk), so every CU takes part; k_reduce_parts then sums the
// S / chunk leading rows.
__global__ __launch_bounds__(256) void k_sum_chunks(float* __restrict__ part, int S, int64_t n, int64_t stride,
                                                    int chunk) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int s0 = (int)blockIdx.y * chunk, s1 = min(S, s0 + chunk);
  float* p = part + (int64_t)s0 * stride + i;
  float acc = 0.f;
  int s = s0;
  for (; s + 8 <= s1; s += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(s - s0 + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; s < s1; ++s) acc += p[(int64_t)(s - s0) * stride];
  *p = acc;
}

// r[i] = (accum ? r[i] : 0) + sum_s part[s*stride + i] for i < n, where r is
// out for i < n1 and out2 (from index n1 on) otherwise.  A workgroup owns 64
// consecutive i; its 4 waves sum the partials s = g, g+4, g+8, ... (8
// independent loads in flight per lane), then wave 0 adds the four sums in a
// fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, int S, int64_t n,
                                                      int64_t stride, int64_t n1, float* __restrict__ out,
                                                      float* __restrict__ out2, int accum) {
  __shared__ float red[4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  float acc = 0.f;
  if (i < n) {
    int k = g;
    for (; k + 28 < S; k += 32) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part[(int64_t)(k + 4 * u) * stride + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; k < S; k += 4) acc += part[(int64_t)k * stride + i];
  }
  red[g][lane] = acc;
  __syncthreads();
  if (g == 0 && i < n) {
    const float s = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    float* o = i < n1 ? out + i : out2 + (i - n1);
    *o = accum ? *o + s : s;
  }
}

// fp32 [Co][T][Ci] -> bf16 [Co][T][Ci], or for the input gradient the flipped
// transpose bf16 [Ci][T][Co] with tap T-1-t.
__global__ void k_pack_w(const float* __restrict__ w, int Co, int T, int Ci, int flip, u16* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)Co * T * Ci;
  if (i >= n) return;
  const int ci = (int)(i % Ci), t = (int)((i / Ci) % T), co = (int)(i / ((int64_t)Ci * T));
  if (!flip)
    out[i] = f2bf(w[i]);
  else
    out[((int64_t)ci * T + (T - 1 - t)) * Co + co] = f2bf(w[i]);
}

// ---------------------------------------------------------------------------
// 2x2x2 max pooling; one thread per (coarse voxel, 8-channel group).
// D, H, W: fine grid (even).
__global__ void k_maxpool2(const u16* __restrict__ x, int xs, int N, int D, int H, int W, int C, u16* __restrict__ y,
                           int ys) {
  const int G = C >> 3, d = D >> 1, h = H >> 1, w = W >> 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * d * h * w * G) return;
  const int g = (int)(i % G);
  int64_t v = i / G;
  const int cx = (int)(v % w), cy = (int)((v / w) % h), cz = (int)((v / ((int64_t)w * h)) % d);
  const int n = (int)(v / ((int64_t)w * h * d));
  float m[8];
  for (int c = 0; c < 8; ++c) m[c] = -__builtin_inff();
  for (int o = 0; o < 8; ++o) {
    const int64_t fv = (((int64_t)n * D + 2 * cz + (o >> 2)) * H + 2 * cy + ((o >> 1) & 1)) * W + 2 * cx + (o & 1);
    const u16x8 a = *(const u16x8*)(x + fv * xs + 8 * g);
    for (int c = 0; c < 8; ++c) m[c] = fmaxf(m[c], bf2f(a[c]));
  }
  u16x8 o8;
  for (int c = 0; c < 8; ++c) o8[c] = f2bf(m[c]);
  *(u16x8*)(y + v * ys + 8 * g) = o8;
}

// Gradient of k_maxpool2: dy goes to the first maximum (fine order z, y, x).
__global__ void k_maxpool2_bwd(const u16* __restrict__ x, int xs, const u16* __restrict__ dy, int dys, int N, int D,
                               int H, int W, int C, u16* __restrict__ dx, int dxs, int accum, int gate) {
  const int G = C >> 3, d = D >> 1, h = H >> 1, w = W >> 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * d * h * w * G) return;
  const int g = (int)(i % G);
  int64_t v = i / G;
  const int cx = (int)(v % w), cy = (int)((v / w) % h), cz = (int)((v / ((int64_t)w * h)) % d);
  const int n = (int)(v / ((int64_t)w * h * d));
  int64_t fv[8];
  float m[8];
  int am[8];
  u16x8 xa[8];  // the window's inputs, kept for the gate
  for (int c = 0; c < 8; ++c) {
    m[c] = -__builtin_inff();
    am[c] = 0;
  }
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    fv[o] = (((int64_t)n * D + 2 * cz + (o >> 2)) * H + 2 * cy + ((o >> 1) & 1)) * W + 2 * cx + (o & 1);
    xa[o] = *(const u16x8*)(x + fv[o] * xs + 8 * g);
    for (int c = 0; c < 8; ++c)
      if (bf2f(xa[o][c]) > m[c]) {
        m[c] = bf2f(xa[o][c]);
        am[c] = o;
      }
  }
  const u16x8 gy = *(const u16x8*)(dy + v * dys + 8 * g);
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    u16* p = dx + fv[o] * dxs + 8 * g;
    u16x8 cur = accum ? *(const u16x8*)p : (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
    u16x8 out;
    for (int c = 0; c < 8; ++c) {
      float val = accum ? bf2f(cur[c]) : 0.f;
      if (am[c] == o) val += bf2f(gy[c]);
      // gate: x is the output of a ReLU; its backward zeroes dx where x <= 0
      out[c] = gate && !(bf2f(xa[o][c]) > 0.f) ? (u16)0 : f2bf(val);
    }
    *(u16x8*)p = out;
  }
}

// The pooling-window kernels below with x-pair lanes: thread (coarse voxel,
// half c, 8-channel group g), g fastest, then c, then the coarse voxel, so
// consecutive lanes read consecutive 16-byte pieces of a fine row (fine x =
// 2 cx + c).  The two halves of a coarse voxel lie G lanes apart in one wave
// (2G divides 64) and are combined by a lane shuffle.  Indices are 32-bit:
// the host uses these kernels when the thread count fits.
struct PoolIx {
  int n, cz, cy, cx, c, g;
};

__device__ __forceinline__ PoolIx pool_ix(uint32_t i, int G, int d, int h, int w) {
  PoolIx r;
  r.g = (int)(i % (uint32_t)G);
  i /= (uint32_t)G;
  r.c = (int)(i & 1u);
  i >>= 1;
  r.cx = (int)(i % (uint32_t)w);
  i /= (uint32_t)w;
  r.cy = (int)(i % (uint32_t)h);
  i /= (uint32_t)h;
  r.cz = (int)(i % (uint32_t)d);
  r.n = (int)(i / (uint32_t)d);
  return r;
}

// k_maxpool2 with x-pair lanes (same result: the maximum is taken in any order).
__global__ __launch_bounds__(256) void k_maxpool2_lx(const u16* __restrict__ x, int xs, int N, int D, int H, int W,
                                                     int C, u16* __restrict__ y, int ys, uint32_t total) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= total) return;
  const int G = C >> 3, d = D >> 1, h = H >> 1, w = W >> 1;
  const PoolIx p = pool_ix(i, G, d, h, w);
  const int partner = (int)(threadIdx.x & 63) + (p.c ? -G : G);
  float m[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = -__builtin_inff();
  u16x8 a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t fv = (((int64_t)p.n * D + 2 * p.cz + (j >> 1)) * H + 2 * p.cy + (j & 1)) * W + 2 * p.cx + p.c;
    a[j] = *(const u16x8*)(x + fv * xs + 8 * p.g);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int k = 0; k < 8; ++k) m[k] = fmaxf(m[k], bf2f(a[j][k]));
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = fmaxf(m[k], __shfl(m[k], partner, 64));
  if (p.c) return;
  u16x8 o8;
#pragma unroll
  for (int k = 0; k < 8; ++k) o8[k] = f2bf(m[k]);
  const int64_t v = (((int64_t)p.n * d + p.cz) * h + p.cy) * w + p.cx;
  *(u16x8*)(y + v * ys + 8 * p.g) = o8;
}

// k_maxpool2_bwd with x-pair lanes: each half finds its first maximum in the
// fine order (z, y, x), the pair keeps the larger, the earlier on a tie, and
// each lane writes its own four fine voxels: the same values as k_maxpool2_bwd.
__global__ __launch_bounds__(256) void k_maxpool2_bwd_lx(const u16* __restrict__ x, int xs, const u16* __restrict__ dy,
                                                         int dys, int N, int D, int H, int W, int C,
                                                         u16* __restrict__ dx, int dxs, int accum, int gate,
                                                         uint32_t total) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= total) return;
  const int G = C >> 3, d = D >> 1, h = H >> 1, w = W >> 1;
  const PoolIx p = pool_ix(i, G, d, h, w);
  const int partner = (int)(threadIdx.x & 63) + (p.c ? -G : G);
  float m[8];
  int am[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    m[k] = -__builtin_inff();
    am[k] = 0;
  }
  int64_t fv[4];
  u16x8 xa[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    fv[j] = (((int64_t)p.n * D + 2 * p.cz + (j >> 1)) * H + 2 * p.cy + (j & 1)) * W + 2 * p.cx + p.c;
    xa[j] = *(const u16x8*)(x + fv[j] * xs + 8 * p.g);
  }
  const int64_t v = (((int64_t)p.n * d + p.cz) * h + p.cy) * w + p.cx;
  const u16x8 gy = *(const u16x8*)(dy + v * dys + 8 * p.g);
  // the values accumulated onto, loaded with the rest: a load right before
  // its store would wait out a round trip per fine voxel
  u16x8 cur[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    cur[j] = accum ? *(const u16x8*)(dx + fv[j] * dxs + 8 * p.g) : (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (bf2f(xa[j][k]) > m[k]) {
        m[k] = bf2f(xa[j][k]);
        am[k] = 2 * j + p.c;  // window order o = (dz << 2) | (dy << 1) | dx
      }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float mo = __shfl(m[k], partner, 64);
    const int ao = __shfl(am[k], partner, 64);
    if (mo > m[k] || (mo == m[k] && ao < am[k])) am[k] = ao;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    u16* q = dx + fv[j] * dxs + 8 * p.g;
    u16x8 out;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float val = accum ? bf2f(cur[j][k]) : 0.f;
      if (am[k] == 2 * j + p.c) val += bf2f(gy[k]);
      out[k] = gate && !(bf2f(xa[j][k]) > 0.f) ? (u16)0 : f2bf(val);
    }
    *(u16x8*)q = out;
  }
}

// k_upsample2_bwd with x-pair lanes (D, H, W coarse): each half sums its four
// fine voxels in window order, the pair adds the two sums (FP32).
__global__ __launch_bounds__(256) void k_upsample2_bwd_lx(const u16* __restrict__ dy, int dys, int N, int D, int H,
                                                          int W, int C, u16* __restrict__ dx, int dxs,
                                                          const u16* __restrict__ gate, int gs, uint32_t total) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= total) return;
  const int G = C >> 3;
  const PoolIx p = pool_ix(i, G, D, H, W);
  const int partner = (int)(threadIdx.x & 63) + (p.c ? -G : G);
  u16x8 a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t fv =
        (((int64_t)p.n * 2 * D + 2 * p.cz + (j >> 1)) * (2 * H) + 2 * p.cy + (j & 1)) * (2 * W) + 2 * p.cx + p.c;
    a[j] = *(const u16x8*)(dy + fv * dys + 8 * p.g);
  }
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] += bf2f(a[j][k]);
  float t[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) t[k] = __shfl(s[k], partner, 64);
  if (p.c) return;
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] += t[k];
  const int64_t v = (((int64_t)p.n * D + p.cz) * H + p.cy) * W + p.cx;
  const u16x8 gv = gate ? *(const u16x8*)(gate + v * gs + 8 * p.g) : (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
  u16x8 o8;
#pragma unroll
  for (int k = 0; k < 8; ++k) o8[k] = gate && !(bf2f(gv[k]) > 0.f) ? (u16)0 : f2bf(s[k]);
  *(u16x8*)(dx + v * dxs + 8 * p.g) = o8;
}

// Nearest 2x upsampling; one thread per (fine voxel, 8-channel group).
// D, H, W: coarse grid.
__global__ void k_upsample2(const u16* __restrict__ x, int xs, int N, int D, int H, int W, int C, u16* __restrict__ y,
                            int ys) {
  const int G = C >> 3, Df = 2 * D, Hf = 2 * H, Wf = 2 * W;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * Df * Hf * Wf * G) return;
  const int g = (int)(i % G);
  const int64_t v = i / G;
  const int fx = (int)(v % Wf), fy = (int)((v / Wf) % Hf), fz = (int)((v / ((int64_t)Wf * Hf)) % Df);
  const int n = (int)(v / ((int64_t)Wf * Hf * Df));
  const int64_t cv = (((int64_t)n * D + (fz >> 1)) * H + (fy >> 1)) * W + (fx >> 1);
  *(u16x8*)(y + v * ys + 8 * g) = *(const u16x8*)(x + cv * xs + 8 * g);
}

// k_upsample2 with 32-bit index arithmetic (the host uses it when the thread
// count fits): the same copy, without 64-bit divisions per thread.
__global__ __launch_bounds__(256) void k_upsample2_32(const u16* __restrict__ x, int xs, int N, int D, int H, int W,
                                                      int C, u16* __restrict__ y, int ys, uint32_t total) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= total) return;
  const uint32_t G = (uint32_t)C >> 3, Wf = 2u * W, Hf = 2u * H, Df = 2u * D;
  const uint32_t g = i % G, v = i / G;
  const uint32_t fx = v % Wf, r = v / Wf, fy = r % Hf, r2 = r / Hf, fz = r2 % Df, n = r2 / Df;
  const int64_t cv = (((int64_t)n * D + (fz >> 1)) * H + (fy >> 1)) * W + (fx >> 1);
  *(u16x8*)(y + (int64_t)v * ys + 8 * g) = *(const u16x8*)(x + cv * xs + 8 * g);
}

// Gradient of k_upsample2: dx[coarse] = sum of the 8 fine dy (FP32 sum).
__global__ void k_upsample2_bwd(const u16* __restrict__ dy, int dys, int N, int D, int H, int W, int C,
                                u16* __restrict__ dx, int dxs, const u16* __restrict__ gate, int gs) {
  const int G = C >> 3, Hf = 2 * H, Wf = 2 * W;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * D * H * W * G) return;
  const int g = (int)(i % G);
  const int64_t v = i / G;
  const int cx = (int)(v % W), cy = (int)((v / W) % H), cz = (int)((v / ((int64_t)W * H)) % D);
  const int n = (int)(v / ((int64_t)W * H * D));
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int o = 0; o < 8; ++o) {
    const int64_t fv = (((int64_t)n * 2 * D + 2 * cz + (o >> 2)) * Hf + 2 * cy + ((o >> 1) & 1)) * Wf + 2 * cx + (o & 1);
    const u16x8 a = *(const u16x8*)(dy + fv * dys + 8 * g);
    for (int c = 0; c < 8; ++c) s[c] += bf2f(a[c]);
  }
  // gate: the output of the ReLU before dx; its backward zeroes dx where gate <= 0
  const u16x8 gv = gate ? *(const u16x8*)(gate + v * gs + 8 * g) : (u16x8){0, 0, 0, 0, 0, 0, 0, 0};
  u16x8 o8;
  for (int c = 0; c < 8; ++c) o8[c] = gate && !(bf2f(gv[c]) > 0.f) ? (u16)0 : f2bf(s[c]);
  *(u16x8*)(dx + v * dxs + 8 * g) = o8;
}

// dy *= (y > 0), in place; one thread per (voxel, 8-channel group).
__global__ void k_relu_bwd(u16* __restrict__ dy, int dys, const u16* __restrict__ y, int ys, int64_t nvox, int C) {
  const int G = C >> 3;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox * G) return;
  const int64_t v = i / G;
  const int g = (int)(i % G);
  const u16x8 a = *(const u16x8*)(y + v * ys + 8 * g);
  u16x8 d = *(const u16x8*)(dy + v * dys + 8 * g);
  for (int c = 0; c < 8; ++c)
    if (!(bf2f(a[c]) > 0.f)) d[c] = 0;
  *(u16x8*)(dy + v * dys + 8 * g) = d;
}

// Masked, channel-weighted squared error and its gradient.  One thread per
// voxel; loss partial sums per workgroup into loss[blockIdx.x].
// k_mse for rows of whole 16-byte groups (ps, dps multiples of 8, C <= 8):
// a voxel's prediction and gradient rows move as 16-byte vectors instead of
// 2-byte accesses strided by the row; same sums in the same order.
__global__ __launch_bounds__(256) void k_mse_v(const u16* __restrict__ pred, int ps, const u16* __restrict__ tgt,
                                               int ts, const uint8_t* __restrict__ mask, int64_t nvox, int C,
                                               const float* __restrict__ cw, float scale, float* __restrict__ loss,
                                               u16* __restrict__ dpred, int dps) {
  __shared__ float red[256];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (v < nvox) {
    const bool m = mask[v] != 0;
    const u16x8 p = *(const u16x8*)(pred + v * ps);
    u16x8 gr = {0, 0, 0, 0, 0, 0, 0, 0};
    if (m)
      for (int c = 0; c < C; ++c) {
        const float e = bf2f(p[c]) - bf2f(tgt[v * ts + c]);
        l += cw[c] * e * e;
        gr[c] = f2bf(2.f * scale * cw[c] * e);
      }
    u16* dp = dpred + v * dps;
    *(u16x8*)dp = gr;
    const u16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 8; c < dps; c += 8) *(u16x8*)(dp + c) = z;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = red[0] * scale;
}

__global__ __launch_bounds__(256) void k_mse(const u16* __restrict__ pred, int ps, const u16* __restrict__ tgt, int ts,
                                             const uint8_t* __restrict__ mask, int64_t nvox, int C,
                                             const float* __restrict__ cw, float scale, float* __restrict__ loss,
                                             u16* __restrict__ dpred, int dps) {
  __shared__ float red[256];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (v < nvox) {
    const bool m = mask[v] != 0;
    for (int c = 0; c < dps; ++c) {
      float gr = 0.f;
      if (c < C && m) {
        const float e = bf2f(pred[v * ps + c]) - bf2f(tgt[v * ts + c]);
        l += cw[c] * e * e;
        gr = 2.f * scale * cw[c] * e;
      }
      dpred[v * dps + c] = f2bf(gr);
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = red[0] * scale;
}

// The surrogate's head and loss in one pass, for training: the head is a
// 1x1x1 convolution of the 16-channel decoder output x (itself a ReLU output),
// of which c <= 4 outputs enter the loss.  Per voxel: out_j = bf16(sum_k
// W[j][k] x_k + b_j) (the head's output as stored), the masked weighted
// squared error and its gradient d_j = bf16(2 scale cw_j (out_j - target_j))
// as in k_mse_v, and dx_k = bf16(sum_j W[j][k] d_j), zero where x_k <= 0 (the
// ReLU's backward).  The head's weight gradient sum_v d_j x_k is an MFMA
// product per 64 voxels: each wave stages its voxels' x rows and d rows
// (bf16, voxel-major) in its own 4 KB of LDS and builds the fragments with
// transposing reads as k_wgrad27 does (dW rows j = the A rows, only 0-3
// non-zero).  A workgroup writes its loss sum times scale to lossp[block] and
// its sums, reduced across its four waves in a fixed order, to part row
// block: [4][16] weight sums, then 4 bias sums.  Workgroups stride over the
// voxels, each lane loading its next voxel before it works on this one.
constexpr int HEAD_CI = 16, HEAD_C = 4, HEAD_NP = HEAD_C * HEAD_CI + HEAD_C;
__global__ __launch_bounds__(256) void k_head_loss(const u16* __restrict__ x, int xs, const u16* __restrict__ wp,
                                                   const float* __restrict__ bias, int c, const u16* __restrict__ tgt,
                                                   int ts, const uint8_t* __restrict__ mask, int64_t nvox,
                                                   const float* __restrict__ cw, float scale, float* __restrict__ lossp,
                                                   u16* __restrict__ dx, int dxs, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) u16 sx[4][64 * 16];
  __shared__ __attribute__((aligned(16))) u16 sd[4][64 * 16];
  __shared__ float red[4][HEAD_NP + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const int lrow = lr >> 2, lcol = 4 * (lr & 3);
  // the weights are the same for every lane: scalar registers
  auto uni = [](float f) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(f))); };
  float W[HEAD_C][HEAD_CI], b[HEAD_C], cwv[HEAD_C];
#pragma unroll
  for (int j = 0; j < HEAD_C; ++j) {
#pragma unroll
    for (int k = 0; k < HEAD_CI; ++k) W[j][k] = uni(j < c ? bf2f(wp[j * HEAD_CI + k]) : 0.f);
    b[j] = uni(j < c && bias ? bias[j] : 0.f);
    cwv[j] = uni(j < c ? cw[j] : 0.f);
  }
  u16* const mx = sx[wave];
  u16* const md = sd[wave];
  // d rows: channels 4-15 stay zero (the loop writes 0-3)
  *(u32x4*)(md + lane * 16) = (u32x4){0u, 0u, 0u, 0u};
  *(u32x4*)(md + lane * 16 + 8) = (u32x4){0u, 0u, 0u, 0u};
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float gb[HEAD_C] = {0.f, 0.f, 0.f, 0.f}, l = 0.f;
  const int64_t step = (int64_t)gridDim.x * 256;
  int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t v_end = ((nvox + 255) / 256) * 256;  // whole waves run the loop, lanes past nvox hold zeros
  auto load = [&](int64_t vv, u32x4& a0, u32x4& a1, uint32_t& m, u32x2& tg) {
    a0 = a1 = (u32x4){0u, 0u, 0u, 0u};
    m = 0;
    tg = (u32x2){0u, 0u};
    if (vv < nvox) {
      a0 = *(const u32x4*)(x + vv * xs);
      a1 = *(const u32x4*)(x + vv * xs + 8);
      m = mask[vv];
      if (m) {
        u16 t4[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < HEAD_C; ++j)
          if (j < c) t4[j] = tgt[vv * ts + j];
        tg.x = t4[0] | (uint32_t)t4[1] << 16;
        tg.y = t4[2] | (uint32_t)t4[3] << 16;
      }
    }
  };
  u32x4 a0, a1;
  uint32_t m;
  u32x2 tg;
  load(v, a0, a1, m, tg);
  for (; v < v_end; v += step) {
    u32x4 n0, n1;
    uint32_t nm;
    u32x2 ntg;
    load(v + step, n0, n1, nm, ntg);
    float xf[HEAD_CI];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      xf[2 * k] = __uint_as_float(a0[k] << 16);
      xf[2 * k + 1] = __uint_as_float(a0[k] & 0xffff0000u);
      xf[8 + 2 * k] = __uint_as_float(a1[k] << 16);
      xf[8 + 2 * k + 1] = __uint_as_float(a1[k] & 0xffff0000u);
    }
    float d[HEAD_C] = {0.f, 0.f, 0.f, 0.f};
    if (m) {
      const float t[HEAD_C] = {__uint_as_float(tg.x << 16), __uint_as_float(tg.x & 0xffff0000u),
                               __uint_as_float(tg.y << 16), __uint_as_float(tg.y & 0xffff0000u)};
#pragma unroll
      for (int j = 0; j < HEAD_C; ++j) {
        if (j >= c) continue;
        float o = 0.f;
#pragma unroll
        for (int k = 0; k < HEAD_CI; ++k) o = fmaf(W[j][k], xf[k], o);
        const float e = bf2f(f2bf(o + b[j])) - t[j];
        l += cwv[j] * e * e;
        d[j] = bf2f(f2bf(2.f * scale * cwv[j] * e));
      }
    }
    if (v < nvox) {
      u32x4 o0, o1;
#pragma unroll
      for (int k = 0; k < HEAD_CI; k += 2) {
        float g0 = 0.f, g1 = 0.f;
#pragma unroll
        for (int j = 0; j < HEAD_C; ++j) {
          g0 = fmaf(W[j][k], d[j], g0);
          g1 = fmaf(W[j][k + 1], d[j], g1);
        }
        const uint32_t lo = xf[k] > 0.f ? f2bf(g0) : 0u, hi = xf[k + 1] > 0.f ? f2bf(g1) : 0u;
        if (k < 8) o0[k / 2] = lo | hi << 16;
        else o1[(k - 8) / 2] = lo | hi << 16;
      }
      *(u32x4*)(dx + v * dxs) = o0;
      *(u32x4*)(dx + v * dxs + 8) = o1;
    }
#pragma unroll
    for (int j = 0; j < HEAD_C; ++j) gb[j] += d[j];
    // the wave's 64 voxels as rows of its LDS: x as loaded, d as bf16
    *(u32x4*)(mx + lane * 16) = a0;
    *(u32x4*)(mx + lane * 16 + 8) = a1;
    *(u32x2*)(md + lane * 16) = (u32x2){(uint32_t)f2bf(d[0]) | (uint32_t)f2bf(d[1]) << 16,
                                        (uint32_t)f2bf(d[2]) | (uint32_t)f2bf(d[3]) << 16};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int vb = 32 * ks + 8 * lg;
      acc = mfma16(frag_tr(md + (vb + lrow) * 16 + lcol, 16), frag_tr(mx + (vb + lrow) * 16 + lcol, 16), acc);
    }
    a0 = n0;
    a1 = n1;
    m = nm;
    tg = ntg;
  }
  // C element i of lane (lr, lg): dW[4 lg + i][lr]; rows 0-3 are in lanes 0-15
  auto wave_sum = [&](float t) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    return t;
  };
  if (lg == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][i * HEAD_CI + lr] = acc[i];
#pragma unroll
  for (int j = 0; j < HEAD_C; ++j) {
    const float t = wave_sum(gb[j]);
    if (lane == 0) red[wave][HEAD_C * HEAD_CI + j] = t;
  }
  {
    const float t = wave_sum(l);
    if (lane == 0) red[wave][HEAD_NP] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x <= HEAD_NP) {
    const int i = threadIdx.x;
    const float t = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    if (i < HEAD_NP) part[(int64_t)blockIdx.x * HEAD_NP + i] = t;
    else lossp[blockIdx.x] = t * scale;
  }
}

// Adam (Kingma & Ba 2015) on FP32 master weights.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// 3x3x3 convolutions use k_conv_fwd16, one 16-channel output tile per
// workgroup, at every channel count (abx_conv3d_set_kernel switches back to
// k_conv_fwd, for comparison).
int g_fwd16 = 1;
// Weight gradients use k_conv_wgrad_tr (transposing LDS reads); 0 = the
// gathering kernel k_conv_wgrad, for comparison.
int g_wgrad_tr = 1;
// 3x3x3 weight gradients use k_wgrad27 (every tap in one workgroup); 0 = the
// tap-group kernels above, for comparison.
int g_wgrad27 = 1;
// k_conv_fwd16: 0 = brick and weights loaded straight into LDS, the weights
// then held in registers (two workgroups per CU); 1 = the weights read from
// LDS for every MFMA (g_fwd16_wgs workgroups per CU); 2 = brick and weights
// staged through registers (two per CU).
int g_fwd16_mode = 0;
int g_fwd16_wgs = 2;
// k_wgrad27 loads its tiles straight into LDS where its blocking allows more
// than one workgroup per CU (0: staged through registers everywhere).
int g_wgrad27_dma = 1;

// Compute units of the current device (cached per device ordinal).
int device_cus() {
  static int cache[64] = {0};
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return 256;
  if (!cache[d]) {
    int v = 0;
    cache[d] = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, d) == hipSuccess && v > 0 ? v : 256;
  }
  return cache[d];
}

int64_t tiles_of(int n, int d, int h, int w) {
  return (int64_t)n * ((d + TZ - 1) / TZ) * ((h + TY - 1) / TY) * ((w + TX - 1) / TX);
}

int fwd_chunk(int ci) {
  for (int c : {64, 48, 32, 16})
    if (ci % c == 0) return c;
  return 0;
}

// Weight-gradient launch shape: channel tiles, taps per workgroup, splits.
struct WgShape {
  int nct, nci, tg, S, blocks_y;
};

WgShape wgrad_shape(int n, int d, int h, int w, int ci, int co, int taps) {
  WgShape s;
  s.nct = co % 64 == 0 ? 4 : (co % 32 == 0 ? 2 : 1);
  s.nci = (ci % 32 == 0) ? 2 : 1;
  s.tg = wgrad_tg(taps, s.nct, s.nci);
  const int ngr = (taps + s.tg - 1) / s.tg;
  s.blocks_y = (co / (16 * s.nct)) * (ci / (16 * s.nci)) * ngr;
  const int64_t nt = tiles_of(n, d, h, w);
  // enough workgroups to hide each tile's load latency, but at most about 8 M
  // partial sums (4 M and 16 M measured slower in a training step: fewer
  // workgroups wait on their loads, more partials cost their reduction)
  const int64_t by_size = std::max<int64_t>(1, ((int64_t)8 << 20) / ((int64_t)co * taps * ci + co));
  s.S = (int)std::max<int64_t>(1, std::min<int64_t>({nt, (int64_t)std::max(1, 8192 / s.blocks_y), by_size}));
  return s;
}

// k_wgrad27's launch shape: channel tiles per workgroup and partial rows.
struct W27Shape {
  int nct, nci, tz, S, blocks_y;
};

// Workgroups per CU of k_wgrad27's direct-load variant for a blocking: as many
// as LDS holds, at most three (80 registers), two with four output tiles (64
// accumulators: 128 registers); 0 where the blocking keeps the staged variant
// (2 x 3 and 4 x 2 tiles: their accumulators leave one workgroup per CU, which
// only the prefetch into registers overlaps, or spill at two).
constexpr int w27_wgs(int nct, int nci, int tz) {
  const int lds = (16 * 4 * tz * 16 * nct + ((18 * 6 * (tz + 2) * nci * 2 + 63) / 64 * 64) * 8) * 2;
  return nct * nci >= 6 ? 0 : std::min(nct >= 4 ? 2 : 3, 160 * 1024 / lds);
}

// Channel blocking nci (1-3 tiles of 16 input channels) with the given
// output blocking nct.
W27Shape wgrad27_shape_nci(int n, int d, int h, int w, int ci, int co, int nct, int nci, bool dma) {
  W27Shape s;
  s.nct = nct;
  s.nci = nci;
  s.blocks_y = (co / (16 * s.nct)) * (ci / (16 * s.nci));
  // 8-deep tiles where there are at least 8 per CU, for the one blocking compiled with them (its
  // staging registers fit; the wider input blocks would spill)
  const auto tiles = [&](int tz) {
    return (int64_t)n * ((d + tz - 1) / tz) * ((h + W27Y - 1) / W27Y) * ((w + W27X - 1) / W27X);
  };
  s.tz = (s.nct == 1 && s.nci == 1 && tiles(8) >= 8 * (int64_t)device_cus()) ? 8 : 4;
  const int64_t nt = tiles(s.tz);
  // one round of resident workgroups (two per CU for the blockings within
  // 128 registers, one for the wider ones; more rounds only add partial
  // sums to reduce), at most about 16 M partial sums
  const int wgs = dma ? w27_wgs(s.nct, s.nci, s.tz) : 0;
  const int res = wgs > 0 ? wgs : (s.nct * s.nci <= 2 ? 2 : 1);
  const int64_t want = (res * (int64_t)device_cus() + s.blocks_y - 1) / s.blocks_y;
  const int64_t by_size = std::max<int64_t>(1, ((int64_t)16 << 20) / ((int64_t)co * 27 * ci + co));
  s.S = (int)std::max<int64_t>(1, std::min<int64_t>({nt, want, by_size}));
  return s;
}

int w27_nct(int co) { return co % 64 == 0 ? 4 : (co % 32 == 0 ? 2 : 1); }

// The widest input block (48, 32 or 16 channels) that divides ci; blocks may
// straddle the split of a two-part input.
// dma: the direct-load variant may run (its workgroups per CU set the rows)
W27Shape wgrad27_shape(int n, int d, int h, int w, int ci, int co, bool dma) {
  const int nct = w27_nct(co);
  const int nci = (ci % 48 == 0 && nct <= 2) ? 3 : (ci % 32 == 0 ? 2 : 1);
  return wgrad27_shape_nci(n, d, h, w, ci, co, nct, nci, dma);
}

// cx: the input channels that x itself holds (ci, or the first part's)
bool conv_args_ok(const char* fn, const void* x, int xs, int cx, int n, int d, int h, int w, int ci, int co,
                  int ksize) {
  if (!x || n < 1 || d < 1 || h < 1 || w < 1 || ci < 16 || co < 16 || ci % 16 || co % 16 || (ksize != 1 && ksize != 3) ||
      xs < cx || xs % 8 || !aligned16(x) || (int64_t)n * d * h * w > ((int64_t)1 << 31) / 2) {
    set_errorf("%s: bad argument (channels must be multiples of 16, strides of 8, pointers 16-byte aligned)", fn);
    return false;
  }
  return true;
}

}  // namespace
}  // namespace abx

using namespace abx;

namespace abx {
namespace {
// Every layer's two packed layouts in one launch: thread i finds its layer by
// the running element offsets (at most ABX_PACK_MAX layers) and writes the
// forward and the input-gradient copy of its weight, as k_pack_w does.
struct PackTable {
  abx_pack_desc d[ABX_PACK_MAX];
  int64_t start[ABX_PACK_MAX + 1];
  int n;
};

__global__ void k_pack_all(PackTable tab) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tab.start[tab.n]) return;
  int l = 0;
  while (l + 1 < tab.n && i >= tab.start[l + 1]) ++l;
  const abx_pack_desc& d = tab.d[l];
  const int64_t j = i - tab.start[l];
  const int T = d.taps, Ci = d.ci, Co = d.co;
  const int ci = (int)(j % Ci), t = (int)((j / Ci) % T), co = (int)(j / ((int64_t)Ci * T));
  const u16 v = f2bf(d.w[j]);
  d.fwd[j] = v;
  d.bwd[((int64_t)ci * T + (T - 1 - t)) * Co + co] = v;
}
}  // namespace
}  // namespace abx

extern "C" int abx_conv3d_pack_all(void* stream, int n, const abx_pack_desc* desc) {
  if (n < 1 || n > ABX_PACK_MAX || !desc) {
    set_errorf("abx_conv3d_pack_all: between 1 and %d layers", ABX_PACK_MAX);
    return -1;
  }
  PackTable tab{};
  tab.n = n;
  tab.start[0] = 0;
  for (int l = 0; l < n; ++l) {
    const abx_pack_desc& d = desc[l];
    if (!d.w || !d.fwd || !d.bwd || d.co < 1 || d.ci < 1 || (d.taps != 1 && d.taps != 27)) {
      set_errorf("abx_conv3d_pack_all: bad layer %d", l);
      return -1;
    }
    tab.d[l] = d;
    tab.start[l + 1] = tab.start[l] + (int64_t)d.co * d.taps * d.ci;
  }
  hipLaunchKernelGGL(k_pack_all, dim3(blocks_for(tab.start[n])), dim3(256), 0, (hipStream_t)stream, tab);
  ABX_HIP(hipGetLastError());
  return 0;
}

extern "C" int abx_conv3d_pack(void* stream, const float* w, int co, int ksize, int ci, int for_input_grad,
                               uint16_t* out) {
  if (!w || !out || co < 1 || ci < 1 || (ksize != 1 && ksize != 3)) {
    set_error("abx_conv3d_pack: bad argument");
    return -1;
  }
  const int T = ksize == 3 ? 27 : 1;
  const int64_t n = (int64_t)co * T * ci;
  hipLaunchKernelGGL(k_pack_w, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, co, T, ci, for_input_grad,
                     out);
  ABX_HIP(hipGetLastError());
  return 0;
}

namespace {
// Two-part input / output bounds: channels [0, c1) in the first tensor, [c1, c) in the second.
bool parts_ok(const char* fn, const void* p2, int s2, int c1, int c) {
  if (c1 < 16 || c1 > c || c1 % 16) {
    set_errorf("%s: the first channel part must be a multiple of 16 of at most all channels", fn);
    return false;
  }
  if (c1 < c && (!p2 || !aligned16(p2) || s2 < c - c1 || s2 % 8)) {
    set_errorf("%s: bad second channel part (NULL, unaligned, or its stride below its channels)", fn);
    return false;
  }
  return true;
}

int conv_forward(const char* fn, void* stream, const uint16_t* x, int xs, int ci1, const uint16_t* x2, int xs2, int n,
                 int d, int h, int w, int ci, const uint16_t* wpack, const float* bias, int co, int ksize, int relu,
                 int accumulate, uint16_t* y, int ys, int co1, uint16_t* y2, int ys2, const uint16_t* gate, int gs) {
  if (!conv_args_ok(fn, x, xs, ci1, n, d, h, w, ci, co, ksize) || !parts_ok(fn, x2, xs2, ci1, ci) ||
      !parts_ok(fn, y2, ys2, co1, co))
    return -1;
  if (!wpack || !y || ys < co1 || ys % 8 || !aligned16(wpack) || !aligned16(y)) {
    set_errorf("%s: bad output or weight argument", fn);
    return -1;
  }
  if (gate && (gs < co || gs % 8 || !aligned16(gate))) {
    set_errorf("%s: gate stride must be >= co and a multiple of 8, the pointer 16-byte aligned", fn);
    return -1;
  }
  const bool parts = ci1 < ci || co1 < co;
  // k_conv_fwd16 addresses its outputs by 32-bit offsets within 8 planes and
  // its brick by 32-bit byte offsets within 10
  const bool f16_fits =
      (int64_t)F16Z * h * w * std::max({ys, parts ? ys2 : ys, gate ? gs : 0}) < ((int64_t)1 << 30) &&
      (int64_t)(F16Z + 2) * h * w * std::max(xs, ci1 < ci ? xs2 : xs) * 2 < ((int64_t)1 << 31);
  if (parts && (ksize != 3 || !g_fwd16 || gate || !f16_fits)) {
    set_errorf("%s: channel parts need the 16-channel 3x3x3 kernel, no gate, 8 planes of output below 2^30 "
               "elements and 10 planes of input below 2^31 bytes", fn);
    return -1;
  }
  if (ci1 == ci) {
    x2 = x;
    xs2 = xs;
  }
  if (co1 == co) {
    y2 = y;
    ys2 = ys;
  }
  const int ck = fwd_chunk(ci);
  const int nct = co % 64 == 0 ? 4 : (co % 32 == 0 ? 2 : 1);
  const dim3 grid((unsigned)tiles_of(n, d, h, w), (unsigned)(co / (16 * nct)));
  hipStream_t s = (hipStream_t)stream;
  if (ksize == 3 && g_fwd16 && f16_fits) {
    const int64_t tiles = (int64_t)n * ((d + F16Z - 1) / F16Z) * ((h + F16Y - 1) / F16Y) * ((w + F16X - 1) / F16X);
    // persistent workgroups: g_fwd16_wgs (of eight waves) per CU across the co tiles
    const int64_t per_co = std::max<int64_t>(1, (int64_t)g_fwd16_wgs * device_cus() / (co / 16));
    const unsigned gx = (unsigned)std::min<int64_t>(tiles, per_co);
    if (g_fwd16_mode == 0 && gate)
      hipLaunchKernelGGL((k_conv_fwd16<true, 4, true, true>), dim3(gx, (unsigned)(co / 16)), dim3(512), 0, s, x, xs,
                         x2, xs2, ci1 / 16, wpack, bias, y, ys, y2, ys2, co1 / 16, gate, gs, d, h, w, ci, (int)tiles,
                         relu, accumulate);
    else if (g_fwd16_mode == 0)
      hipLaunchKernelGGL((k_conv_fwd16<true, 4, true>), dim3(gx, (unsigned)(co / 16)), dim3(512), 0, s, x, xs, x2,
                         xs2, ci1 / 16, wpack, bias, y, ys, y2, ys2, co1 / 16, gate, gs, d, h, w, ci, (int)tiles, relu,
                         accumulate);
    else if (g_fwd16_mode == 1)
      hipLaunchKernelGGL(k_conv_fwd16<true>, dim3(gx, (unsigned)(co / 16)), dim3(512), 0, s, x, xs, x2, xs2,
                         ci1 / 16, wpack, bias, y, ys, y2, ys2, co1 / 16, gate, gs, d, h, w, ci, (int)tiles, relu,
                         accumulate);
    else
      hipLaunchKernelGGL(k_conv_fwd16<false>, dim3(gx, (unsigned)(co / 16)), dim3(512), 0, s, x, xs, x2, xs2,
                         ci1 / 16, wpack, bias, y, ys, y2, ys2, co1 / 16, gate, gs, d, h, w, ci, (int)tiles, relu,
                         accumulate);
    ABX_HIP(hipGetLastError());
    return 0;
  }
#define ABX_FWD(T, NC, C)                                                                                          \
  hipLaunchKernelGGL((k_conv_fwd<T, NC, C>), grid, dim3(256), 0, s, x, xs, wpack, bias, y, ys, gate, gs, d, h, w, \
                     ci, co, relu, accumulate)
#define ABX_FWD_CK(T, NC)          \
  if (ck == 64) ABX_FWD(T, NC, 64); \
  else if (ck == 48) ABX_FWD(T, NC, 48); \
  else if (ck == 32) ABX_FWD(T, NC, 32); \
  else ABX_FWD(T, NC, 16);
  if (ksize == 3) {
    if (nct == 4) { ABX_FWD_CK(27, 4) }
    else if (nct == 2) { ABX_FWD_CK(27, 2) }
    else { ABX_FWD_CK(27, 1) }
  } else {
    if (nct == 4) { ABX_FWD_CK(1, 4) }
    else if (nct == 2) { ABX_FWD_CK(1, 2) }
    else { ABX_FWD_CK(1, 1) }
  }
#undef ABX_FWD_CK
#undef ABX_FWD
  ABX_HIP(hipGetLastError());
  return 0;
}
}  // namespace

extern "C" int abx_conv3d_forward(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int ci,
                                  const uint16_t* wpack, const float* bias, int co, int ksize, int relu, int accumulate,
                                  uint16_t* y, int ys) {
  return conv_forward("abx_conv3d_forward", stream, x, xs, ci, nullptr, 0, n, d, h, w, ci, wpack, bias, co, ksize,
                      relu, accumulate, y, ys, co, nullptr, 0, nullptr, 0);
}

extern "C" int abx_conv3d_forward_parts(void* stream, const uint16_t* x, int xs, int ci1, const uint16_t* x2, int xs2,
                                        int n, int d, int h, int w, int ci, const uint16_t* wpack, const float* bias,
                                        int co, int relu, int accumulate, uint16_t* y, int ys, int co1, uint16_t* y2,
                                        int ys2) {
  return conv_forward("abx_conv3d_forward_parts", stream, x, xs, ci1, x2, xs2, n, d, h, w, ci, wpack, bias, co, 3,
                      relu, accumulate, y, ys, co1, y2, ys2, nullptr, 0);
}

extern "C" int abx_conv3d_forward_gated(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int ci,
                                        const uint16_t* wpack, const float* bias, int co, int ksize, int relu,
                                        int accumulate, uint16_t* y, int ys, const uint16_t* gate, int gs) {
  if (!gate) {
    set_error("abx_conv3d_forward_gated: gate is NULL");
    return -1;
  }
  return conv_forward("abx_conv3d_forward_gated", stream, x, xs, ci, nullptr, 0, n, d, h, w, ci, wpack, bias, co,
                      ksize, relu, accumulate, y, ys, co, nullptr, 0, gate, gs);
}

extern "C" int abx_conv3d_set_kernel(int flags) {
  g_fwd16 = (flags & 1) ? 1 : 0;
  g_wgrad_tr = (flags & 2) ? 1 : 0;
  g_wgrad27 = (flags & 4) ? 0 : 1;
  g_fwd16_mode = (flags & 8) ? 2 : ((flags & 64) ? 1 : 0);
  g_fwd16_wgs = g_fwd16_mode == 1 && !(flags & 16) ? 3 : 2;
  g_wgrad27_dma = (flags & 32) ? 0 : 1;
  return 0;
}

extern "C" int64_t abx_conv3d_wgrad_workspace(int n, int d, int h, int w, int ci, int co, int ksize) {
  if (n < 1 || d < 1 || h < 1 || w < 1 || ci < 16 || co < 16 || ci % 16 || co % 16 || (ksize != 1 && ksize != 3))
    return -1;
  const int T = ksize == 3 ? 27 : 1;
  const WgShape s = wgrad_shape(n, d, h, w, ci, co, T);
  int64_t rows = s.S;
  // enough for either 3x3x3 kernel, whichever abx_conv3d_set_kernel selects
  // the direct-load variant's rows, at least those of the staged one
  if (T == 27) rows = std::max<int64_t>(rows, wgrad27_shape(n, d, h, w, ci, co, g_wgrad27_dma).S);
  return rows * ((int64_t)co * T * ci + co) * 4;
}

namespace {
// Sums nrows partial rows (stride floats apart) element by element in a fixed
// order: chunks of 32 rows in place, then the chunk sums; elements below n1
// go to out, the others to out2 (element i - n1), added to them if accumulate.
int reduce_rows(hipStream_t s, float* part, int nrows, int64_t nred, int64_t stride, int64_t n1, float* out,
                float* out2, int accumulate) {
  constexpr int kChunk = 32;
  int rows = nrows;
  if (nrows > kChunk) {
    rows = (nrows + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(nred), (unsigned)rows), dim3(256), 0, s, part, nrows, nred, stride,
                       kChunk);
    stride *= kChunk;
  }
  hipLaunchKernelGGL(k_reduce_parts, dim3(blocks_for(nred, 64)), dim3(256), 0, s, part, rows, nred, stride, n1, out,
                     out2, accumulate);
  ABX_HIP(hipGetLastError());
  return 0;
}

template <int A, int B, int Z>
void launch_w27(dim3 g, hipStream_t s, bool dma, const u16* x, int xs, const u16* x2, int xs2, int ci1, const u16* dz,
                int dzs, float* part, float* partb, int d, int h, int w, int ci, int co, int nt) {
  constexpr int wgs = w27_wgs(A, B, Z);
  static_assert(w27_lds_bytes<A, B, Z>() * wgs <= 160 * 1024, "LDS of the resident workgroups");
  if constexpr (wgs > 0)
    if (dma) {
      hipLaunchKernelGGL((k_wgrad27<A, B, Z, wgs>), g, dim3(512), 0, s, x, xs, x2, xs2, ci1, dz, dzs, part, partb, d,
                         h, w, ci, co, nt);
      return;
    }
  hipLaunchKernelGGL((k_wgrad27<A, B, Z, 0>), g, dim3(512), 0, s, x, xs, x2, xs2, ci1, dz, dzs, part, partb, d, h, w,
                     ci, co, nt);
}

int conv_wgrad(const char* fn, void* stream, const uint16_t* x, int xs, int ci1, const uint16_t* x2, int xs2,
               const uint16_t* dz, int dzs, int n, int d, int h, int w, int ci, int co, int ksize, float* workspace,
               int64_t ws_bytes, float* dw, float* db, int accumulate) {
  if (!conv_args_ok(fn, x, xs, ci1, n, d, h, w, ci, co, ksize) || !parts_ok(fn, x2, xs2, ci1, ci)) return -1;
  if (!dz || !dw || !workspace || dzs < co || dzs % 8 || !aligned16(dz)) {
    set_errorf("%s: bad gradient or output argument", fn);
    return -1;
  }
  if (ci1 < ci && ksize != 3) {
    set_errorf("%s: channel parts need a 3x3x3 convolution", fn);
    return -1;
  }
  if (ci1 == ci) {
    x2 = x;
    xs2 = xs;
  }
  const int T = ksize == 3 ? 27 : 1;
  const WgShape sh = wgrad_shape(n, d, h, w, ci, co, T);
  // two-part inputs always take k_wgrad27; so does every output width but an
  // odd multiple of 16 above 16, where its one-tile output blocks would stage
  // the input once per output tile (48 output channels: 272 against 223 us
  // for the tap-group kernel on the 96x48x64 batch of 8)
  const bool odd_wide = w27_nct(co) == 1 && co > 16;
  const bool use27 = T == 27 && ((g_wgrad27 && !odd_wide) || ci1 < ci);
  // the direct-load variant addresses its tiles by 32-bit byte offsets: 10
  // planes of input and 8 of dZ below 2^31 bytes
  const bool dma27 = g_wgrad27_dma && (int64_t)10 * h * w * std::max(xs, ci1 < ci ? xs2 : xs) * 2 < ((int64_t)1 << 31) &&
                     (int64_t)8 * h * w * dzs * 2 < ((int64_t)1 << 31);
  const W27Shape s27 = use27 ? wgrad27_shape(n, d, h, w, ci, co, dma27) : W27Shape{1, 1, 4, 1, 1};
  const int nrows = use27 ? s27.S : sh.S;
  const int64_t nw = (int64_t)co * T * ci;
  if (ws_bytes < (int64_t)nrows * (nw + co) * 4) {
    set_errorf("%s: workspace of %lld bytes is smaller than the %lld needed", fn, (long long)ws_bytes,
               (long long)((int64_t)nrows * (nw + co) * 4));
    return -1;
  }
  float* part = workspace;              // rows of nw weight partials, then co bias partials
  float* partb = db ? workspace : nullptr;  // a flag: the bias partials are written
  hipStream_t s = (hipStream_t)stream;
  if (use27) {
    const int nt27 = (int)((int64_t)n * ((d + s27.tz - 1) / s27.tz) * ((h + W27Y - 1) / W27Y) * ((w + W27X - 1) / W27X));
    const dim3 g27((unsigned)s27.S, (unsigned)s27.blocks_y);
#define ABX_W27(A, B, Z)                                                                                         \
  launch_w27<A, B, Z>(g27, s, dma27, x, xs, x2, xs2, ci1, dz, dzs, part, partb, d, h, w, ci, co, nt27)
    if (s27.tz == 8) ABX_W27(1, 1, 8);
    else if (s27.nct == 4 && s27.nci == 2) ABX_W27(4, 2, 4);
    else if (s27.nct == 4) ABX_W27(4, 1, 4);
    else if (s27.nct == 2 && s27.nci == 3) ABX_W27(2, 3, 4);
    else if (s27.nct == 2 && s27.nci == 2) ABX_W27(2, 2, 4);
    else if (s27.nct == 2) ABX_W27(2, 1, 4);
    else if (s27.nci == 3) ABX_W27(1, 3, 4);
    else if (s27.nci == 2) ABX_W27(1, 2, 4);
    else ABX_W27(1, 1, 4);
#undef ABX_W27
  }
  const int nt = (int)tiles_of(n, d, h, w);
  const dim3 grid((unsigned)sh.S, (unsigned)sh.blocks_y);
#define ABX_LAUNCH(KERNEL)                                                                                     \
  hipLaunchKernelGGL(KERNEL, grid, dim3(256), 0, s, x, xs, dz, dzs, part, partb, d, h, w, ci, co, nt, \
                     sh.tg)
#define ABX_WG(T_, A, B)                                  \
  do {                                                    \
    if (g_wgrad_tr) ABX_LAUNCH((k_conv_wgrad_tr<T_, A, B>)); \
    else ABX_LAUNCH((k_conv_wgrad<T_, A, B>));            \
  } while (0)
#define ABX_WG_T(T_)                                      \
  if (sh.nct == 4 && sh.nci == 2) ABX_WG(T_, 4, 2);       \
  else if (sh.nct == 4) ABX_WG(T_, 4, 1);                 \
  else if (sh.nct == 2 && sh.nci == 2) ABX_WG(T_, 2, 2);  \
  else if (sh.nct == 2) ABX_WG(T_, 2, 1);                 \
  else if (sh.nci == 2) ABX_WG(T_, 1, 2);                 \
  else ABX_WG(T_, 1, 1);
  if (T == 27) {
    if (!use27) {
      ABX_WG_T(27)
    }
  } else {
    ABX_WG_T(1)
  }
#undef ABX_WG_T
#undef ABX_WG
#undef ABX_LAUNCH
  ABX_HIP(hipGetLastError());
  // weights and bias in one reduction (rows of nw + co partials)
  return reduce_rows(s, part, nrows, nw + (db ? co : 0), nw + co, nw, dw, db, accumulate);
}
}  // namespace

extern "C" int abx_conv3d_wgrad(void* stream, const uint16_t* x, int xs, const uint16_t* dz, int dzs, int n, int d,
                                int h, int w, int ci, int co, int ksize, float* workspace, int64_t ws_bytes, float* dw,
                                float* db, int accumulate) {
  return conv_wgrad("abx_conv3d_wgrad", stream, x, xs, ci, nullptr, 0, dz, dzs, n, d, h, w, ci, co, ksize, workspace,
                    ws_bytes, dw, db, accumulate);
}

extern "C" int abx_conv3d_wgrad_parts(void* stream, const uint16_t* x, int xs, int ci1, const uint16_t* x2, int xs2,
                                      const uint16_t* dz, int dzs, int n, int d, int h, int w, int ci, int co,
                                      float* workspace, int64_t ws_bytes, float* dw, float* db, int accumulate) {
  return conv_wgrad("abx_conv3d_wgrad_parts", stream, x, xs, ci1, x2, xs2, dz, dzs, n, d, h, w, ci, co, 3, workspace,
                    ws_bytes, dw, db, accumulate);
}

namespace {
// The x-pair-lane kernels need a coarse voxel's 2G lanes inside one wave
// (2G divides 64) and a 32-bit thread index.
bool pair_lanes_ok(int c, int64_t tot) {
  const int G = c / 8;
  return 64 % (2 * G) == 0 && 2 * tot < ((int64_t)1 << 31);
}

bool ew_ok(const char* fn, std::initializer_list<const void*> ptrs, std::initializer_list<int> strides, int c) {
  for (const void* p : ptrs)
    if (!p || !aligned16(p)) {
      set_errorf("%s: null or unaligned tensor", fn);
      return false;
    }
  for (int s : strides)
    if (s < c || s % 8) {
      set_errorf("%s: stride must be >= channels and a multiple of 8", fn);
      return false;
    }
  if (c < 8 || c % 8) {
    set_errorf("%s: channels must be a multiple of 8", fn);
    return false;
  }
  return true;
}
}  // namespace

extern "C" int abx_maxpool2(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int c, uint16_t* y,
                            int ys) {
  if (!ew_ok("abx_maxpool2", {x, y}, {xs, ys}, c)) return -1;
  if (n < 1 || d < 2 || h < 2 || w < 2 || d % 2 || h % 2 || w % 2) {
    set_error("abx_maxpool2: the grid must have even, positive sizes");
    return -1;
  }
  const int64_t tot = (int64_t)n * (d / 2) * (h / 2) * (w / 2) * (c / 8);
  if (pair_lanes_ok(c, tot))
    hipLaunchKernelGGL(k_maxpool2_lx, dim3(blocks_for(2 * tot)), dim3(256), 0, (hipStream_t)stream, x, xs, n, d, h, w,
                       c, y, ys, (uint32_t)(2 * tot));
  else
    hipLaunchKernelGGL(k_maxpool2, dim3(blocks_for(tot)), dim3(256), 0, (hipStream_t)stream, x, xs, n, d, h, w, c, y,
                       ys);
  ABX_HIP(hipGetLastError());
  return 0;
}

namespace {
int maxpool2_backward(const char* fn, void* stream, const uint16_t* x, int xs, const uint16_t* dy, int dys, int n,
                      int d, int h, int w, int c, uint16_t* dx, int dxs, int accumulate, int gate) {
  if (!ew_ok(fn, {x, dy, dx}, {xs, dys, dxs}, c)) return -1;
  if (n < 1 || d < 2 || h < 2 || w < 2 || d % 2 || h % 2 || w % 2) {
    set_errorf("%s: the grid must have even, positive sizes", fn);
    return -1;
  }
  const int64_t tot = (int64_t)n * (d / 2) * (h / 2) * (w / 2) * (c / 8);
  if (pair_lanes_ok(c, tot))
    hipLaunchKernelGGL(k_maxpool2_bwd_lx, dim3(blocks_for(2 * tot)), dim3(256), 0, (hipStream_t)stream, x, xs, dy, dys,
                       n, d, h, w, c, dx, dxs, accumulate, gate, (uint32_t)(2 * tot));
  else
    hipLaunchKernelGGL(k_maxpool2_bwd, dim3(blocks_for(tot)), dim3(256), 0, (hipStream_t)stream, x, xs, dy, dys, n, d,
                       h, w, c, dx, dxs, accumulate, gate);
  ABX_HIP(hipGetLastError());
  return 0;
}

int upsample2_backward(const char* fn, void* stream, const uint16_t* dy, int dys, int n, int d, int h, int w, int c,
                       uint16_t* dx, int dxs, const uint16_t* gate, int gs) {
  if (!ew_ok(fn, {dy, dx}, {dys, dxs}, c)) return -1;
  if (n < 1 || d < 1 || h < 1 || w < 1) {
    set_errorf("%s: bad grid", fn);
    return -1;
  }
  const int64_t tot = (int64_t)n * d * h * w * (c / 8);
  if (pair_lanes_ok(c, tot))
    hipLaunchKernelGGL(k_upsample2_bwd_lx, dim3(blocks_for(2 * tot)), dim3(256), 0, (hipStream_t)stream, dy, dys, n, d,
                       h, w, c, dx, dxs, gate, gs, (uint32_t)(2 * tot));
  else
    hipLaunchKernelGGL(k_upsample2_bwd, dim3(blocks_for(tot)), dim3(256), 0, (hipStream_t)stream, dy, dys, n, d, h, w,
                       c, dx, dxs, gate, gs);
  ABX_HIP(hipGetLastError());
  return 0;
}
}  // namespace

extern "C" int abx_maxpool2_backward(void* stream, const uint16_t* x, int xs, const uint16_t* dy, int dys, int n, int d,
                                     int h, int w, int c, uint16_t* dx, int dxs, int accumulate) {
  return maxpool2_backward("abx_maxpool2_backward", stream, x, xs, dy, dys, n, d, h, w, c, dx, dxs, accumulate, 0);
}

extern "C" int abx_maxpool2_backward_gated(void* stream, const uint16_t* x, int xs, const uint16_t* dy, int dys, int n,
                                           int d, int h, int w, int c, uint16_t* dx, int dxs, int accumulate) {
  return maxpool2_backward("abx_maxpool2_backward_gated", stream, x, xs, dy, dys, n, d, h, w, c, dx, dxs, accumulate,
                           1);
}

extern "C" int abx_upsample2(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int c, uint16_t* y,
                             int ys) {
  if (!ew_ok("abx_upsample2", {x, y}, {xs, ys}, c)) return -1;
  if (n < 1 || d < 1 || h < 1 || w < 1) {
    set_error("abx_upsample2: bad grid");
    return -1;
  }
  const int64_t tot = (int64_t)n * 8 * d * h * w * (c / 8);
  if (tot < ((int64_t)1 << 31))
    hipLaunchKernelGGL(k_upsample2_32, dim3(blocks_for(tot)), dim3(256), 0, (hipStream_t)stream, x, xs, n, d, h, w, c,
                       y, ys, (uint32_t)tot);
  else
    hipLaunchKernelGGL(k_upsample2, dim3(blocks_for(tot)), dim3(256), 0, (hipStream_t)stream, x, xs, n, d, h, w, c, y,
                       ys);
  ABX_HIP(hipGetLastError());
  return 0;
}

extern "C" int abx_upsample2_backward(void* stream, const uint16_t* dy, int dys, int n, int d, int h, int w, int c,
                                      uint16_t* dx, int dxs) {
  return upsample2_backward("abx_upsample2_backward", stream, dy, dys, n, d, h, w, c, dx, dxs, nullptr, 0);
}

extern "C" int abx_upsample2_backward_gated(void* stream, const uint16_t* dy, int dys, int n, int d, int h, int w,
                                            int c, uint16_t* dx, int dxs, const uint16_t* gate, int gs) {
  if (!gate || gs < c || gs % 8 || !aligned16(gate)) {
    set_error("abx_upsample2_backward_gated: gate must be 16-byte aligned with a stride >= c, a multiple of 8");
    return -1;
  }
  return upsample2_backward("abx_upsample2_backward_gated", stream, dy, dys, n, d, h, w, c, dx, dxs, gate, gs);
}

extern "C" int abx_relu_backward(void* stream, uint16_t* dy, int dys, const uint16_t* y, int ys, int64_t nvox, int c) {
  if (!ew_ok("abx_relu_backward", {dy, y}, {dys, ys}, c) || nvox < 1) return -1;
  hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(nvox * (c / 8))), dim3(256), 0, (hipStream_t)stream, dy, dys, y, ys,
                     nvox, c);
  ABX_HIP(hipGetLastError());
  return 0;
}

extern "C" int64_t abx_mse_partials(int64_t nvox) { return nvox < 1 ? -1 : (int64_t)blocks_for(nvox); }

namespace abx {
namespace {
// k_head_loss's grid: up to 1536 workgroups striding over the voxels (six per CU)
int64_t head_blocks(int64_t nvox) { return nvox < 1 ? -1 : std::min<int64_t>(blocks_for(nvox), 1536); }
}  // namespace
}  // namespace abx

extern "C" int64_t abx_head_loss_partials(int64_t nvox) { return head_blocks(nvox); }

extern "C" int64_t abx_head_loss_workspace(int64_t nvox) {
  return nvox < 1 ? -1 : head_blocks(nvox) * HEAD_NP * (int64_t)sizeof(float);
}

extern "C" int abx_head_loss_backward(void* stream, const uint16_t* x, int xs, int ci, const uint16_t* wpack,
                                      const float* bias, int c, const uint16_t* target, int ts, const uint8_t* mask,
                                      int64_t nvox, const float* chan_w, float scale, float* loss_parts, uint16_t* dx,
                                      int dxs, float* workspace, int64_t ws_bytes, float* dw, float* db) {
  if (!x || !wpack || !target || !mask || !chan_w || !loss_parts || !dx || !workspace || !dw || !db || nvox < 1 ||
      ci != HEAD_CI || c < 1 || c > HEAD_C || xs < ci || xs % 8 || dxs < ci || dxs % 8 || ts < c ||
      !aligned16(x) || !aligned16(dx)) {
    set_error("abx_head_loss_backward: bad argument (16 input channels, 1-4 outputs, strides multiples of 8, "
              "16-byte aligned x and dx)");
    return -1;
  }
  if (ws_bytes < abx_head_loss_workspace(nvox)) {
    set_errorf("abx_head_loss_backward: workspace of %lld bytes is smaller than the %lld needed", (long long)ws_bytes,
               (long long)abx_head_loss_workspace(nvox));
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t g = head_blocks(nvox);
  hipLaunchKernelGGL(k_head_loss, dim3((unsigned)g), dim3(256), 0, s, x, xs, wpack, bias, c, target, ts, mask, nvox,
                     chan_w, scale, loss_parts, dx, dxs, workspace);
  ABX_HIP(hipGetLastError());
  return reduce_rows(s, workspace, (int)g, HEAD_NP, HEAD_NP, HEAD_C * HEAD_CI, dw, db, 0);
}

extern "C" int abx_mse_loss(void* stream, const uint16_t* pred, int ps, const uint16_t* target, int ts,
                            const uint8_t* mask, int64_t nvox, int c, const float* chan_w, float scale, float* loss_parts,
                            uint16_t* dpred, int dps) {
  if (!pred || !target || !mask || !chan_w || !loss_parts || !dpred || nvox < 1 || c < 1 || ps < c || ts < c ||
      dps < c) {
    set_error("abx_mse_loss: bad argument");
    return -1;
  }
  if (ps % 8 == 0 && dps % 8 == 0 && c <= 8 && aligned16(pred) && aligned16(dpred))
    hipLaunchKernelGGL(k_mse_v, dim3(blocks_for(nvox)), dim3(256), 0, (hipStream_t)stream, pred, ps, target, ts, mask,
                       nvox, c, chan_w, scale, loss_parts, dpred, dps);
  else
    hipLaunchKernelGGL(k_mse, dim3(blocks_for(nvox)), dim3(256), 0, (hipStream_t)stream, pred, ps, target, ts, mask,
                       nvox, c, chan_w, scale, loss_parts, dpred, dps);
  ABX_HIP(hipGetLastError());
  return 0;
}

extern "C" int abx_adam(void* stream, float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                        float b2, float eps, int step) {
  if (!w || !g || !m || !v || n < 1 || step < 1) {
    set_error("abx_adam: bad argument");
    return -1;
  }
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1, b2, eps,
                     bc1, bc2);
  ABX_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Input volume of the surrogate, built on the GPU from a box's grid arrays.
//
// A lattice of nx x ny x nz cells (the reference's order, k fastest,
// engine.js:128-134) sits at the origin of a d x h x w volume; voxels beyond
// it are solid.  Cell types follow setupSim's precedence (engine.js:250-255):
// inlet = k == 0 and opening, else outlet, else porous when the row's depth
// lies in (dc[top], pleat].  Channels (bf16, the rest zero):
//   0 solid, 1 porous, 2 inlet, 3 K/6 on porous cells,
//   4 signed Euclidean distance to the wall in cells / 8, clipped to [-4, 4]
//     (positive in the fluid: distance to the nearest solid voxel; negative in
//     the solid: minus the distance to the nearest fluid voxel),
//   5 outlet, 6 aniso/20 on porous cells.
// Each value is formed in double, rounded to float and then to bf16 (nearest
// even), the rounding of the host path that builds the same volume in float
// and converts it.
//
// The distances are exact: squared distances are separable,
//   D(i,j,k) = min_i' [(i-i')^2 + min_j' [(j-j')^2 + min_k' [(k-k')^2 + F(i',j',k')]]],
// F = 0 on the target set and "infinite" elsewhere, so three passes, one per
// axis, each take the minimum over the whole line in exact integer
// arithmetic (at most 96 terms per line at the surrogate's volume).
namespace abx {
namespace {

constexpr int32_t EDT_INF = 1 << 30;  // + (line length)^2 stays below 2^31

struct FeatArgs {
  const uint8_t* fluid;    // nx*ny*nz
  const uint8_t* opening;  // nx*ny
  const uint8_t* outlet;   // nx*ny*nz
  const float* dc;         // ny
  int nx, ny, nz, top;
  double pleat, K, aniso;
  int d, h, w;
};

__device__ __forceinline__ bool feat_fluid(const FeatArgs& a, int i, int j, int k) {
  return i < a.nx && j < a.ny && k < a.nz && a.fluid[((int64_t)i * a.ny + j) * a.nz + k] != 0;
}

// Pass along k: for every voxel the squared distance along its k line to the
// nearest non-fluid voxel (in) and to the nearest fluid voxel (out).
__global__ void k_feat_edt_k(FeatArgs a, int32_t* gin, int32_t* gout) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = (int64_t)a.d * a.h * a.w;
  if (v >= nv) return;
  const int k = (int)(v % a.w), j = (int)((v / a.w) % a.h), i = (int)(v / ((int64_t)a.w * a.h));
  int32_t bi = EDT_INF, bo = EDT_INF;
  for (int kk = 0; kk < a.w; ++kk) {
    const int32_t dd = (k - kk) * (k - kk);
    if (feat_fluid(a, i, j, kk))
      bo = min(bo, dd);
    else
      bi = min(bi, dd);
  }
  gin[v] = bi;
  gout[v] = bo;
}

// Pass along j: src holds the k-pass minima.
__global__ void k_feat_edt_j(FeatArgs a, const int32_t* sin, const int32_t* sout, int32_t* din, int32_t* dout) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = (int64_t)a.d * a.h * a.w;
  if (v >= nv) return;
  const int k = (int)(v % a.w), j = (int)((v / a.w) % a.h), i = (int)(v / ((int64_t)a.w * a.h));
  int32_t bi = EDT_INF, bo = EDT_INF;
  for (int jj = 0; jj < a.h; ++jj) {
    const int64_t u = ((int64_t)i * a.h + jj) * a.w + k;
    const int32_t dd = (j - jj) * (j - jj);
    bi = min(bi, sin[u] + dd);
    bo = min(bo, sout[u] + dd);
  }
  din[v] = min(bi, EDT_INF);
  dout[v] = min(bo, EDT_INF);
}

// Pass along i, then every channel of the voxel.
__global__ void k_feat_write(FeatArgs a, const int32_t* sin, const int32_t* sout, u16* x, int xs, int c) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nv = (int64_t)a.d * a.h * a.w;
  if (v >= nv) return;
  const int k = (int)(v % a.w), j = (int)((v / a.w) % a.h), i = (int)(v / ((int64_t)a.w * a.h));
  int32_t bi = EDT_INF, bo = EDT_INF;
  for (int ii = 0; ii < a.d; ++ii) {
    const int64_t u = ((int64_t)ii * a.h + j) * a.w + k;
    const int32_t dd = (i - ii) * (i - ii);
    bi = min(bi, sin[u] + dd);
    bo = min(bo, sout[u] + dd);
  }
  const bool fl = feat_fluid(a, i, j, k);
  // a voxel is at distance 0 of its own set: fluid has bo == 0, solid bi == 0
  const double din = fl ? sqrt((double)bi) : 0.0, dout = fl ? 0.0 : sqrt((double)bo);
  bool inlet = false, outlet = false, porous = false;
  if (fl) {
    const int64_t c3 = ((int64_t)i * a.ny + j) * a.nz + k;
    inlet = k == 0 && a.opening[(int64_t)i * a.ny + j] != 0;
    outlet = !inlet && a.outlet[c3] != 0;
    porous = !inlet && !outlet && a.dc[j] > a.dc[a.top] && (double)a.dc[j] <= a.pleat;
  }
  double sd = (din - dout) / 8.0;
  sd = sd < -4.0 ? -4.0 : (sd > 4.0 ? 4.0 : sd);
  float ch[7];
  ch[0] = fl ? 0.f : 1.f;
  ch[1] = porous ? 1.f : 0.f;
  ch[2] = inlet ? 1.f : 0.f;
  ch[3] = porous ? (float)(a.K / 6.0) : 0.f;
  ch[4] = (float)sd;
  ch[5] = outlet ? 1.f : 0.f;
  ch[6] = porous ? (float)(a.aniso / 20.0) : 0.f;
  u16* o = x + v * xs;
  for (int q = 0; q < c; ++q) o[q] = q < 7 ? f2bf(ch[q]) : (u16)0;
}

}  // namespace
}  // namespace abx

extern "C" int64_t abx_surrogate_features_workspace(int d, int h, int w) {
  if (d < 1 || h < 1 || w < 1) return -1;
  return (int64_t)4 * d * h * w * (int64_t)sizeof(int32_t);
}

extern "C" int abx_surrogate_features(void* stream, const uint8_t* fluid, const uint8_t* opening,
                                      const uint8_t* outlet, const float* dc, int nx, int ny, int nz, int top,
                                      double pleat, double K, double aniso, int d, int h, int w, int32_t* workspace,
                                      int64_t ws_bytes, uint16_t* x, int xs, int c) {
  if (!fluid || !opening || !outlet || !dc || !workspace || !x || nx < 1 || ny < 1 || nz < 1 || top < 0 ||
      top >= ny || d < nx || h < ny || w < nz || d > 1024 || h > 1024 || w > 1024 || c < 7 || xs < c ||
      ws_bytes < abx_surrogate_features_workspace(d, h, w)) {
    set_error("abx_surrogate_features: bad argument (the lattice must fit the volume, 7 <= c <= xs, "
              "workspace of abx_surrogate_features_workspace bytes)");
    return -1;
  }
  FeatArgs a{fluid, opening, outlet, dc, nx, ny, nz, top, pleat, K, aniso, d, h, w};
  const int64_t nv = (int64_t)d * h * w;
  int32_t *g0 = workspace, *g1 = workspace + nv, *g2 = workspace + 2 * nv, *g3 = workspace + 3 * nv;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_feat_edt_k, dim3(blocks_for(nv)), dim3(256), 0, s, a, g0, g1);
  hipLaunchKernelGGL(k_feat_edt_j, dim3(blocks_for(nv)), dim3(256), 0, s, a, g0, g1, g2, g3);
  hipLaunchKernelGGL(k_feat_write, dim3(blocks_for(nv)), dim3(256), 0, s, a, g2, g3, x, xs, c);
  ABX_HIP(hipGetLastError());
  return 0;
}
