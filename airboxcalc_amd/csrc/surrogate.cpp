// The flow surrogate's warm start behind the C ABI (no reference counterpart:
// the reference starts every solve from rest, f = w_q, engine.js:260-261).
//
// abx_surrogate_load reads an inference weight file, packs the 15 layers'
// weights to bf16 once and allocates the activations of one 3-D U-Net pass
// (4 levels, 16/32/64/128 channels, two 3x3x3 convolutions + ReLU per level,
// 2x2x2 max pooling, nearest 2x upsampling with the encoder output
// concatenated, a 1x1x1 head).  abx_surrogate_warm_start prepares the box at
// the network's cell size (or reuses the solve's geometry when that is the
// same), builds the input volume on the device, runs the network over the
// library's own kernels (abx_surrogate_features, abx_conv3d_forward,
// abx_maxpool2, abx_upsample2) and seeds the simulation at the equilibrium of
// the prediction (abx_sim_init_eq_resample).  The Python trainer
// (airboxcalc_amd/unet.py) issues the same kernels in the same order, so both
// paths seed a simulation bit for bit alike.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "abx_common.hpp"

namespace {

using namespace abx;

constexpr int IN_CH = 16, HEAD_CH = 16;
constexpr int W1 = 16, W2 = 32, W3 = 64, W4 = 128;

struct LayerDef {
  const char* name;
  int ci, co, ksize, relu, level;
  int64_t w_off, b_off;
  int taps() const { return ksize == 3 ? 27 : 1; }
  int64_t nw() const { return (int64_t)co * taps() * ci; }
};

// The layers in parameter order: W [co][taps][ci] then bias [co] per layer.
std::vector<LayerDef> layer_table() {
  std::vector<LayerDef> L = {
      {"enc1a", IN_CH, W1, 3, 1, 0},   {"enc1b", W1, W1, 3, 1, 0},    {"enc2a", W1, W2, 3, 1, 1},
      {"enc2b", W2, W2, 3, 1, 1},      {"enc3a", W2, W3, 3, 1, 2},    {"enc3b", W3, W3, 3, 1, 2},
      {"bota", W3, W4, 3, 1, 3},       {"botb", W4, W4, 3, 1, 3},     {"dec3a", W4 + W3, W3, 3, 1, 2},
      {"dec3b", W3, W3, 3, 1, 2},      {"dec2a", W3 + W2, W2, 3, 1, 1}, {"dec2b", W2, W2, 3, 1, 1},
      {"dec1a", W2 + W1, W1, 3, 1, 0}, {"dec1b", W1, W1, 3, 1, 0},    {"head", W1, HEAD_CH, 1, 0, 0}};
  int64_t off = 0;
  for (auto& l : L) {
    l.w_off = off;
    off += l.nw();
    l.b_off = off;
    off += l.co;
  }
  return L;
}

// Inference weight file header (surrogate.py write_weights).
#pragma pack(push, 1)
struct FileHead {
  char magic[8];
  int32_t n_params, d, h, w, in_ch, head_ch, widths[4], meta_len, zero;
  double family_dx, u0_scale, rho_scale, mean[4], std[4];
};
#pragma pack(pop)
static_assert(sizeof(FileHead) == 8 + 12 * 4 + 11 * 8, "header layout");

}  // namespace

struct abx_surrogate {
  abx_dev* dev = nullptr;
  hipStream_t stream = nullptr;
  int d = 0, h = 0, w = 0;
  double family_dx = 0, u0_scale = 0, rho_scale = 0, mean[4] = {}, stdv[4] = {};
  std::vector<LayerDef> L;
  float* wdev = nullptr;      // FP32 parameters
  uint16_t* packs = nullptr;  // bf16 forward layout of every layer, layer l at pk[l]
  std::vector<int64_t> pk;
  void* arena = nullptr;      // activations, features workspace, the coarse grid
  int64_t ws_bytes = 0;
  int32_t* ws = nullptr;
  uint16_t *x = nullptr, *a1 = nullptr, *cat1 = nullptr, *p1 = nullptr, *a2 = nullptr, *cat2 = nullptr,
           *p2 = nullptr, *a3 = nullptr, *cat3 = nullptr, *p3 = nullptr, *a4 = nullptr, *b4 = nullptr,
           *d3a = nullptr, *d3b = nullptr, *d2a = nullptr, *d2b = nullptr, *d1a = nullptr, *d1b = nullptr,
           *out = nullptr;
  uint8_t *fluid = nullptr, *opening = nullptr, *outlet = nullptr;
  float* dc = nullptr;
};

namespace {

void free_net(abx_surrogate* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  if (n->wdev) (void)hipFree(n->wdev);
  if (n->packs) (void)hipFree(n->packs);
  if (n->arena) (void)hipFree(n->arena);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

// Carves the arena: every buffer 256-byte aligned.
struct Carver {
  size_t off = 0;
  char* base = nullptr;
  template <typename T>
  T* take(size_t count) {
    const size_t at = off;
    off = (off + count * sizeof(T) + 255) & ~(size_t)255;
    return base ? (T*)(base + at) : nullptr;
  }
};

void carve(abx_surrogate* n, Carver& c) {
  const size_t v0 = (size_t)n->d * n->h * n->w;
  const size_t v1 = v0 / 8, v2 = v0 / 64, v3 = v0 / 512;
  n->x = c.take<uint16_t>(v0 * IN_CH);
  n->a1 = c.take<uint16_t>(v0 * W1);
  n->cat1 = c.take<uint16_t>(v0 * (W2 + W1));
  n->p1 = c.take<uint16_t>(v1 * W1);
  n->a2 = c.take<uint16_t>(v1 * W2);
  n->cat2 = c.take<uint16_t>(v1 * (W3 + W2));
  n->p2 = c.take<uint16_t>(v2 * W2);
  n->a3 = c.take<uint16_t>(v2 * W3);
  n->cat3 = c.take<uint16_t>(v2 * (W4 + W3));
  n->p3 = c.take<uint16_t>(v3 * W3);
  n->a4 = c.take<uint16_t>(v3 * W4);
  n->b4 = c.take<uint16_t>(v3 * W4);
  n->d3a = c.take<uint16_t>(v2 * W3);
  n->d3b = c.take<uint16_t>(v2 * W3);
  n->d2a = c.take<uint16_t>(v1 * W2);
  n->d2b = c.take<uint16_t>(v1 * W2);
  n->d1a = c.take<uint16_t>(v0 * W1);
  n->d1b = c.take<uint16_t>(v0 * W1);
  n->out = c.take<uint16_t>(v0 * HEAD_CH);
  n->ws = c.take<int32_t>((size_t)(n->ws_bytes + 3) / 4);
  n->fluid = c.take<uint8_t>(v0);
  n->outlet = c.take<uint8_t>(v0);
  n->opening = c.take<uint8_t>((size_t)n->d * n->h);
  n->dc = c.take<float>((size_t)n->h);
}

const LayerDef& layer(const abx_surrogate* n, const char* name) {
  for (const auto& l : n->L)
    if (!strcmp(l.name, name)) return l;
  return n->L.back();  // unreachable: every name below is in the table
}

// y = act(conv(x) + bias) of layer `name` at its level (x, y: channel slices
// with voxel strides xs, ys).
int conv(abx_surrogate* n, const char* name, const uint16_t* x, int xs, uint16_t* y, int ys) {
  const LayerDef& l = layer(n, name);
  const size_t li = (size_t)(&l - n->L.data());
  return abx_conv3d_forward(n->stream, x, xs, 1, n->d >> l.level, n->h >> l.level, n->w >> l.level, l.ci,
                            n->packs + n->pk[li], n->wdev + l.b_off, l.co, l.ksize, l.relu, 0, y, ys);
}

int forward(abx_surrogate* n) {
  const int lv[4][3] = {{n->d, n->h, n->w}, {n->d >> 1, n->h >> 1, n->w >> 1}, {n->d >> 2, n->h >> 2, n->w >> 2},
                        {n->d >> 3, n->h >> 3, n->w >> 3}};
  auto pool = [&](const uint16_t* x, int xs, uint16_t* y, int lvl, int c) {
    return abx_maxpool2(n->stream, x, xs, 1, lv[lvl][0], lv[lvl][1], lv[lvl][2], c, y, c);
  };
  auto up = [&](const uint16_t* x, uint16_t* y, int ys, int lvl, int c) {
    return abx_upsample2(n->stream, x, c, 1, lv[lvl][0], lv[lvl][1], lv[lvl][2], c, y, ys);
  };
  constexpr int C1 = W2 + W1, C2 = W3 + W2, C3 = W4 + W3;  // concatenated widths per level
  uint16_t *e1 = n->cat1 + W2, *e2 = n->cat2 + W3, *e3 = n->cat3 + W4;  // encoder outputs: the skip halves
  if (conv(n, "enc1a", n->x, IN_CH, n->a1, W1) || conv(n, "enc1b", n->a1, W1, e1, C1) ||
      pool(e1, C1, n->p1, 0, W1) || conv(n, "enc2a", n->p1, W1, n->a2, W2) || conv(n, "enc2b", n->a2, W2, e2, C2) ||
      pool(e2, C2, n->p2, 1, W2) || conv(n, "enc3a", n->p2, W2, n->a3, W3) || conv(n, "enc3b", n->a3, W3, e3, C3) ||
      pool(e3, C3, n->p3, 2, W3) || conv(n, "bota", n->p3, W3, n->a4, W4) || conv(n, "botb", n->a4, W4, n->b4, W4) ||
      up(n->b4, n->cat3, C3, 3, W4) || conv(n, "dec3a", n->cat3, C3, n->d3a, W3) ||
      conv(n, "dec3b", n->d3a, W3, n->d3b, W3) || up(n->d3b, n->cat2, C2, 2, W3) ||
      conv(n, "dec2a", n->cat2, C2, n->d2a, W2) || conv(n, "dec2b", n->d2a, W2, n->d2b, W2) ||
      up(n->d2b, n->cat1, C1, 1, W2) || conv(n, "dec1a", n->cat1, C1, n->d1a, W1) ||
      conv(n, "dec1b", n->d1a, W1, n->d1b, W1) || conv(n, "head", n->d1b, W1, n->out, HEAD_CH))
    return -1;
  return 0;
}

}  // namespace

extern "C" int abx_surrogate_load(abx_dev* dev, const char* path, abx_surrogate** out) {
  if (!dev || !path || !out) {
    set_error("abx_surrogate_load: null argument");
    return -1;
  }
  *out = nullptr;
  FILE* fh = fopen(path, "rb");
  if (!fh) {
    set_errorf("abx_surrogate_load: cannot open %s", path);
    return -1;
  }
  FileHead H;
  std::vector<float> w;
  std::vector<LayerDef> L = layer_table();
  const int64_t np = L.back().b_off + L.back().co;
  bool ok = fread(&H, sizeof H, 1, fh) == 1 && !memcmp(H.magic, "ABXUNET1", 8);
  if (ok && (H.n_params != np || H.in_ch != IN_CH || H.head_ch != HEAD_CH || H.widths[0] != W1 ||
             H.widths[1] != W2 || H.widths[2] != W3 || H.widths[3] != W4 || H.meta_len < 0 || H.d % 8 ||
             H.h % 8 || H.w % 8 || H.d < 8 || H.h < 8 || H.w < 8 || H.d > 1024 || H.h > 1024 || H.w > 1024)) {
    fclose(fh);
    set_errorf("abx_surrogate_load: %s holds a network of another shape", path);
    return -1;
  }
  if (ok) {
    const long at = (long)sizeof H + H.meta_len;
    ok = fseek(fh, at + (long)((16 - at % 16) % 16), SEEK_SET) == 0;
  }
  if (ok) {
    w.resize((size_t)np);
    ok = fread(w.data(), sizeof(float), w.size(), fh) == w.size();
  }
  fclose(fh);
  if (!ok) {
    set_errorf("abx_surrogate_load: %s is not a complete surrogate weight file", path);
    return -1;
  }
  ABX_HIP(hipSetDevice(dev->ordinal));
  abx_surrogate* n = new abx_surrogate();
  n->dev = dev;
  n->d = H.d;
  n->h = H.h;
  n->w = H.w;
  n->family_dx = H.family_dx;
  n->u0_scale = H.u0_scale;
  n->rho_scale = H.rho_scale;
  for (int c = 0; c < 4; ++c) {
    n->mean[c] = H.mean[c];
    n->stdv[c] = H.std[c];
  }
  n->L = L;
  int64_t pks = 0;
  for (const auto& l : L) {
    n->pk.push_back(pks);
    pks += (l.nw() + 127) / 128 * 128;  // 256-byte aligned packs
  }
  n->ws_bytes = abx_surrogate_features_workspace(n->d, n->h, n->w);
  Carver sizer;
  carve(n, sizer);
  if (n->ws_bytes < 0 || hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&n->wdev, (size_t)np * 4) != hipSuccess ||
      hipMalloc((void**)&n->packs, (size_t)pks * 2) != hipSuccess || hipMalloc(&n->arena, sizer.off) != hipSuccess) {
    free_net(n);
    set_error("abx_surrogate_load: device allocation failed");
    return -1;
  }
  Carver c;
  c.base = (char*)n->arena;
  carve(n, c);
  if (hipMemcpyAsync(n->wdev, w.data(), (size_t)np * 4, hipMemcpyHostToDevice, n->stream) != hipSuccess ||
      hipMemsetAsync(n->arena, 0, sizer.off, n->stream) != hipSuccess) {
    free_net(n);
    set_error("abx_surrogate_load: upload failed");
    return -1;
  }
  for (size_t i = 0; i < L.size(); ++i)
    if (abx_conv3d_pack(n->stream, n->wdev + L[i].w_off, L[i].co, L[i].ksize, L[i].ci, 0, n->packs + n->pk[i])) {
      free_net(n);
      return -1;
    }
  if (hipStreamSynchronize(n->stream) != hipSuccess) {
    free_net(n);
    set_error("abx_surrogate_load: weight packing failed");
    return -1;
  }
  *out = n;
  return 0;
}

extern "C" int abx_surrogate_info(const abx_surrogate* n, int dims[3], double* family_dx) {
  if (!n) {
    set_error("abx_surrogate_info: null argument");
    return -1;
  }
  if (dims) {
    dims[0] = n->d;
    dims[1] = n->h;
    dims[2] = n->w;
  }
  if (family_dx) *family_dx = n->family_dx;
  return 0;
}

extern "C" void abx_surrogate_free(abx_surrogate* n) { free_net(n); }

extern "C" int abx_surrogate_warm_start(abx_surrogate* n, abx_sim* sim, const abx_geom* g_fine, const float* tris,
                                        int64_t nfloat, const abx_prep_params* p, double K, double aniso) {
  if (!n || !sim || !g_fine) {
    set_error("abx_surrogate_warm_start: null argument");
    return -1;
  }
  abx_geom_info fi;
  if (abx_geom_info_get(g_fine, &fi)) return -1;
  // a solve at the network's own cell size already holds the same prepared
  // geometry (prepare is deterministic): reuse it
  const bool same = fabs(fi.h - n->family_dx) < 1e-9;
  abx_geom* own = nullptr;
  const abx_geom* gc = g_fine;
  if (!same) {
    if (!tris || !p) {
      set_error("abx_surrogate_warm_start: the box's triangles and prepare parameters are needed to prepare it at "
                "the network's cell size");
      return -1;
    }
    abx_prep_params pc = *p;
    pc.dx = n->family_dx;
    pc.max_cells = 0;
    if (abx_prepare(n->dev, tris, nfloat, &pc, &own)) return -1;
    gc = own;
  }
  int rc = -1;
  abx_geom_info ci;
  abx_grid grid;
  if (abx_geom_info_get(gc, &ci) || abx_geom_grid(gc, &grid)) goto done;
  if (grid.nx > n->d || grid.ny > n->h || grid.nz > n->w) {
    set_errorf("abx_surrogate_warm_start: lattice %dx%dx%d does not fit the surrogate volume %dx%dx%d", grid.nx,
               grid.ny, grid.nz, n->d, n->h, n->w);
    goto done;
  }
  if (hipSetDevice(n->dev->ordinal) != hipSuccess) {
    set_error("abx_surrogate_warm_start: hipSetDevice failed");
    goto done;
  }
  {
    const size_t N = (size_t)grid.nx * grid.ny * grid.nz;
    hipStream_t st = n->stream;
    if (hipMemcpyAsync(n->fluid, grid.fluid, N, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(n->outlet, grid.outlet, N, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(n->opening, grid.opening, (size_t)grid.nx * grid.ny, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(n->dc, grid.dc, (size_t)grid.ny * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
      set_error("abx_surrogate_warm_start: upload failed");
      goto done;
    }
    if (abx_surrogate_features(st, n->fluid, n->opening, n->outlet, n->dc, grid.nx, grid.ny, grid.nz, grid.top,
                               grid.pleat, K, aniso, n->d, n->h, n->w, n->ws, n->ws_bytes, n->x, IN_CH, IN_CH) ||
        forward(n))
      goto done;
    // prediction -> lattice units: u = u0 * (std * v + mean), rho = 1 + rho_scale * (std * v + mean)
    const double scale[4] = {n->u0_scale, n->u0_scale, n->u0_scale, n->rho_scale};
    float affine[8];
    for (int c = 0; c < 4; ++c) {
      affine[c] = (float)(scale[c] * n->stdv[c]);
      affine[4 + c] = (float)(scale[c] * n->mean[c] + (c == 3 ? 1.0 : 0.0));
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
      set_error("abx_surrogate_warm_start: the network pass failed");
      goto done;
    }
    rc = abx_sim_init_eq_resample(sim, n->out, HEAD_CH, n->d, n->h, n->w, grid.nx, grid.ny, grid.nz, n->fluid,
                                  ci.origin, ci.h, fi.origin, fi.h, affine);
  }
done:
  if (own) abx_geom_free(own);
  return rc;
}
