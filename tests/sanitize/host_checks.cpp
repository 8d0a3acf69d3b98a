// Host-code checks under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Built by tests/sanitize/Makefile with -fsanitize=address,undefined from the
// library's host sources (abx_core.cpp: error state, device query, STL
// reader; prepare_host.cpp: the host stages of geometry preparation) and the
// oracle's C restatement, and run by tests/test_host_sanitize.py.
//
//   host_checks prepare STL PARAMS   the host stages of abx_prepare, with CPU
//                                    stand-ins for its GPU stages (27-subcell
//                                    vote, slice, duct, floor, flood fill),
//                                    against the oracle's prepare: every array
//                                    and count equal, or the same error
//   host_checks oracle STL PARAMS N  the oracle's setup, N steps with
//                                    averaging, probe, evaluate, field3d and
//                                    streamlines
//   host_checks fuzz SEED N          N random inputs to abx_parse_stl (against
//                                    the oracle's reader), prep_grid and
//                                    prep_openings
//   host_checks device               abx_device_open without a GPU: an error
//
// PARAMS is a text file of the oracle's prepare parameters in field order:
// dx sealAxis sealSign seal inletAxis inletSign yaw cy sy has_click cx cy cz
// pleat has_outlet ox oy oz out_diameter out_height max_cells.
// Exit status 0 = every check passed; 1 = a mismatch (printed).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "../../airboxcalc_amd/csrc/abx_error.hpp"
#include "../../airboxcalc_amd/csrc/prepare_host.hpp"
#include "../../include/airbox_hip.h"
#include "../../oracle/engine_oracle.h"

namespace {

int g_fail = 0;

void fail(const std::string& what) {
  fprintf(stderr, "MISMATCH: %s\n", what.c_str());
  g_fail = 1;
}

std::vector<uint8_t> read_file(const char* path) {
  std::vector<uint8_t> b;
  FILE* fh = fopen(path, "rb");
  if (!fh) {
    fprintf(stderr, "cannot open %s\n", path);
    exit(2);
  }
  uint8_t buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, fh)) > 0) b.insert(b.end(), buf, buf + n);
  fclose(fh);
  return b;
}

bool read_params(const char* path, orc_prep_params* o, abx_prep_params* a) {
  FILE* fh = fopen(path, "r");
  if (!fh) return false;
  double v[21];
  for (double& x : v)
    if (fscanf(fh, "%lf", &x) != 1) {
      fclose(fh);
      return false;
    }
  fclose(fh);
  memset(o, 0, sizeof *o);
  memset(a, 0, sizeof *a);
  o->dx = a->dx = v[0];
  o->sealAxis = a->sealAxis = (int)v[1];
  o->sealSign = a->sealSign = (int)v[2];
  o->seal = a->seal = v[3];
  o->inletAxis = a->inletAxis = (int)v[4];
  o->inletSign = a->inletSign = (int)v[5];
  o->yaw_deg = a->inletYaw = v[6];
  o->cy = a->cy = v[7];
  o->sy = a->sy = v[8];
  o->has_click = a->has_click = (int)v[9];
  for (int c = 0; c < 3; ++c) o->click[c] = a->click[c] = v[10 + c];
  o->pleat = a->pleat = v[13];
  o->has_outlet = a->has_outlet = (int)v[14];
  for (int c = 0; c < 3; ++c) o->out_center[c] = a->out_center[c] = v[15 + c];
  o->out_diameter = a->out_diameter = v[18];
  o->out_height = a->out_height = v[19];
  o->max_cells = a->max_cells = v[20];
  return true;
}

// abx_parse_stl against orc_parse_stl: same status, count and float bits.
void check_parse(const uint8_t* buf, int64_t len, const char* tag) {
  int64_t n_abx = -1;
  const int ra = abx_parse_stl(buf, len, nullptr, 0, &n_abx);
  float* of = nullptr;
  int64_t n_orc = -1;
  const int ro = orc_parse_stl(buf, len, &of, &n_orc);
  if ((ra == 0) != (ro == 0)) {
    fail(std::string(tag) + ": parse status differs");
  } else if (ra == 0) {
    if (n_abx != n_orc) {
      fail(std::string(tag) + ": parse count differs");
    } else {
      std::vector<float> out((size_t)n_abx + 1, -7.0f);
      int64_t n2 = -1;
      if (abx_parse_stl(buf, len, out.data(), n_abx, &n2) || n2 != n_abx ||
          memcmp(out.data(), of, (size_t)n_abx * sizeof(float)) != 0)
        fail(std::string(tag) + ": parsed values differ");
      if (out[(size_t)n_abx] != -7.0f) fail(std::string(tag) + ": wrote past cap");
      // a short buffer takes only cap floats
      if (n_abx > 3) {
        std::vector<float> part(4, -7.0f);
        if (abx_parse_stl(buf, len, part.data(), 3, &n2) || n2 != n_abx || part[3] != -7.0f)
          fail(std::string(tag) + ": cap not honoured");
      }
    }
  } else if (strcmp(abx::last_error(), orc_last_error()) != 0) {
    fail(std::string(tag) + ": parse error text differs: '" + abx::last_error() + "' vs '" + orc_last_error() + "'");
  }
  orc_free(of);
}

template <typename T>
bool same(const std::vector<T>& a, const T* b, int64_t n) {
  return (int64_t)a.size() == n && (n == 0 || memcmp(a.data(), b, (size_t)n * sizeof(T)) == 0);
}

// The host stages of abx_prepare with CPU stand-ins of its GPU stages.
// Returns 0 or -1 (error text in abx_last_error).
int host_prepare(const float* tris, int64_t nfloat, const abx_prep_params* p, abx::GeomHost& g,
                 const uint8_t* fine_from_oracle) {
  std::vector<float> T;
  if (abx::prep_grid(tris, nfloat, p, g, T)) return -1;
  const abx_geom_info& I = g.info;
  const int nx = I.nx, ny = I.ny, nz = I.nz, nyf = 3 * ny, nzf = 3 * nz;
  const int64_t N = I.N;
  // the voxelizer itself is the GPU's; its fine grid comes from the oracle
  // (k_vote, k_slice stand-ins: 27-subcell vote, slice kf)
  g.solid.assign((size_t)N, 0);
  for (int64_t c = 0; c < N; ++c) {
    const int k = (int)(c % nz);
    const int64_t ij = c / nz;
    const int j = (int)(ij % ny), i = (int)(ij / ny);
    int s = 0;
    for (int da = 0; da < 3; ++da)
      for (int db = 0; db < 3; ++db)
        for (int dc = 0; dc < 3; ++dc)
          s += fine_from_oracle[((int64_t)(3 * i + da) * nyf + (3 * j + db)) * nzf + 3 * k + dc];
    g.solid[(size_t)c] = s >= 7 ? 1 : 0;
  }
  const int kf = abx::prep_slice_k(g);
  std::vector<uint8_t> empty((size_t)3 * nx * nyf);
  for (size_t c = 0; c < empty.size(); ++c)
    empty[c] = (kf >= 0 && kf < nzf && fine_from_oracle[(int64_t)c * nzf + kf]) ? 0 : 1;
  std::vector<uint8_t> floor_row;
  if (abx::prep_openings(p, empty.data(), g) || abx::prep_flood_inputs(g, floor_row)) return -1;
  // k_duct_floor and the flood fill's stand-in: breadth-first, 6-connected
  const int kw0 = I.kw0, top = I.top;
  for (int64_t c = 0; c < N; ++c) {
    const int k = (int)(c % nz);
    const int64_t ij = c / nz;
    if (k < kw0) g.solid[(size_t)c] = g.opening[(size_t)ij] ? 0 : 1;
    if (floor_row[(size_t)(ij % ny)]) g.solid[(size_t)c] = 1;
  }
  g.fluid.assign((size_t)N, 0);
  std::vector<int64_t> Q;
  auto visit = [&](int64_t c, int j) {
    if (!g.fluid[(size_t)c] && !g.solid[(size_t)c] && j >= top) {
      g.fluid[(size_t)c] = 1;
      Q.push_back(c);
    }
  };
  for (int64_t ij = 0; ij < (int64_t)nx * ny; ++ij)
    if (g.opening[(size_t)ij]) visit(ij * nz, (int)(ij % ny));
  for (size_t h = 0; h < Q.size(); ++h) {
    const int64_t c = Q[h];
    const int k = (int)(c % nz);
    const int64_t ij = c / nz;
    const int j = (int)(ij % ny), i = (int)(ij / ny);
    if (i > 0) visit(c - (int64_t)ny * nz, j);
    if (i < nx - 1) visit(c + (int64_t)ny * nz, j);
    if (j > 0) visit(c - nz, j - 1);
    if (j < ny - 1) visit(c + nz, j + 1);
    if (k > 0) visit(c - 1, j);
    if (k < nz - 1) visit(c + 1, j);
  }
  bool leaked = false;
  for (int i = 0; i < nx; ++i)
    for (int j = top; j < ny; ++j)
      for (int k = kw0 < 0 ? 0 : kw0; k < nz; ++k)
        if ((i == 0 || i == nx - 1 || k == nz - 1) && g.fluid[((size_t)i * ny + j) * nz + k]) leaked = true;
  return abx::prep_finish(p, (int64_t)Q.size(), leaked, g);
}

int cmd_prepare(const char* stl, const char* params) {
  std::vector<uint8_t> buf = read_file(stl);
  orc_prep_params op;
  abx_prep_params ap;
  if (!read_params(params, &op, &ap)) {
    fprintf(stderr, "bad parameter file %s\n", params);
    return 2;
  }
  check_parse(buf.data(), (int64_t)buf.size(), "stl");
  int64_t nfloat = 0;
  if (abx_parse_stl(buf.data(), (int64_t)buf.size(), nullptr, 0, &nfloat)) {
    printf("parse error: %s\n", abx::last_error());
    return g_fail;
  }
  std::vector<float> tris((size_t)nfloat);
  abx_parse_stl(buf.data(), (int64_t)buf.size(), tris.data(), nfloat, &nfloat);
  orc_geom og;
  memset(&og, 0, sizeof og);
  const int ro = orc_prepare(tris.data(), nfloat, &op, &og);
  const std::string orc_err = ro ? orc_last_error() : "";
  // the product's host stages need the fine grid of a grid the oracle built
  abx::GeomHost g;
  int ra;
  if (ro == 0) {
    ra = host_prepare(tris.data(), nfloat, &ap, g, og.fine);
  } else {
    // the oracle failed: the stages up to the failure must fail alike; the
    // fine grid is computed here with the oracle's voxelizer when the grid
    // is valid
    std::vector<float> T;
    abx::GeomHost g0;
    ra = abx::prep_grid(tris.data(), nfloat, &ap, g0, T);
    if (ra == 0) {
      const abx_geom_info& I = g0.info;
      std::vector<uint8_t> fine((size_t)27 * I.N);
      orc_voxelize(T.data(), nfloat / 9, I.origin, 3 * I.nx, 3 * I.ny, 3 * I.nz, I.h / 3, fine.data());
      ra = host_prepare(tris.data(), nfloat, &ap, g, fine.data());
    }
  }
  if ((ra == 0) != (ro == 0)) {
    fail(std::string("prepare status differs: product '") + (ra ? abx::last_error() : "ok") + "' oracle '" +
         (ro ? orc_err : "ok") + "'");
  } else if (ro) {
    if (orc_err != abx::last_error())
      fail("prepare error text differs: '" + std::string(abx::last_error()) + "' vs '" + orc_err + "'");
    printf("error: %s\n", orc_err.c_str());
  } else {
    const abx_geom_info& I = g.info;
    if (I.nx != og.nx || I.ny != og.ny || I.nz != og.nz || I.N != og.N || I.nf != og.nf || I.top != og.top ||
        I.jm != og.jm || I.kw0 != og.kw0 || I.nHoles != og.nHoles || I.nOpen != og.nOpen || I.nOut != og.nOut ||
        I.Af != og.Af || I.leak != og.leak)
      fail("prepare counts differ");
    for (int a = 0; a < 3; ++a)
      if (I.origin[a] != og.origin[a] || I.mn[a] != og.mn[a] || I.mx[a] != og.mx[a]) fail("prepare bounds differ");
    const int64_t N = og.N, NR = (int64_t)og.nx * og.ny;
    if (!same(g.solid, og.solid, N)) fail("solid differs");
    if (!same(g.fluid, og.fluid, N)) fail("fluid differs");
    if (!same(g.opening, og.opening, NR)) fail("opening differs");
    if (!same(g.outlet, og.outlet, N)) fail("outlet differs");
    if (!same(g.xc, og.xc, og.nx) || !same(g.dc, og.dc, og.ny) || !same(g.zc, og.zc, og.nz))
      fail("cell centres differ");
    if (!same(g.openPts, og.openPts, 3 * og.nOpenPts) || !same(g.filterPts, og.filterPts, 3 * og.nFilterPts) ||
        !same(g.outletPts, og.outletPts, 3 * og.nOutletPts))
      fail("preview points differ");
    printf("ok grid %dx%dx%d fluid %lld openings %d outlet %d\n", I.nx, I.ny, I.nz, (long long)I.nf, I.nOpen, I.nOut);
  }
  orc_geom_free(&og);
  return g_fail;
}

int cmd_oracle(const char* stl, const char* params, int steps) {
  std::vector<uint8_t> buf = read_file(stl);
  orc_prep_params op;
  abx_prep_params ap;
  if (!read_params(params, &op, &ap)) return 2;
  float* tris = nullptr;
  int64_t nfloat = 0;
  if (orc_parse_stl(buf.data(), (int64_t)buf.size(), &tris, &nfloat)) return 2;
  orc_geom g;
  memset(&g, 0, sizeof g);
  if (orc_prepare(tris, nfloat, &op, &g)) {
    fprintf(stderr, "oracle prepare: %s\n", orc_last_error());
    return 2;
  }
  orc_sim S;
  memset(&S, 0, sizeof S);
  if (orc_setup_sim(&g, 3.0, 10.0, 0.06, 0.51, &S)) return 2;
  const int64_t NF = S.NF;
  static const double W[3] = {1.0 / 3, 1.0 / 18, 1.0 / 36};
  std::vector<float> f((size_t)(19 * NF)), f2((size_t)(19 * NF)), acc((size_t)(4 * NF), 0.0f);
  for (int q = 0; q < 19; ++q)
    for (int64_t a = 0; a < NF; ++a) f[(size_t)(q * NF + a)] = (float)W[q == 0 ? 0 : (q < 7 ? 1 : 2)];
  for (int s = 0; s < steps; ++s) {
    orc_cpu_step(&S, f.data(), f2.data(), s >= steps / 2 ? acc.data() : nullptr, nullptr);
    f.swap(f2);
  }
  double pr[3];
  if (orc_probe(&S, f.data(), steps, pr)) return 2;
  const int navg = steps - steps / 2;
  for (float& v : acc) v /= (float)navg;
  orc_metrics m;
  const int64_t N = g.N, P = 3 * (int64_t)g.nx * g.nz, L = 3 * (int64_t)g.nz * g.ny;
  std::vector<float> fmap((size_t)(g.nx * g.nz)), pin((size_t)P), plow((size_t)P), lon((size_t)L);
  std::vector<uint8_t> feff((size_t)N);
  if (orc_evaluate(&g, &S, S.u0, acc.data(), nullptr, &m, fmap.data(), pin.data(), plow.data(), lon.data(),
                   feff.data()))
    return 2;
  std::vector<float> spd((size_t)N), pst((size_t)N), loss((size_t)N);
  int kB = 0;
  double pinf = 0;
  if (orc_field3d(&g, &S, S.u0, acc.data(), feff.data(), m.rhoRef, spd.data(), pst.data(), loss.data(), &kB, &pinf))
    return 2;
  const int64_t cap = 220 * 600;
  std::vector<float> pts((size_t)(3 * cap)), sp((size_t)cap);
  std::vector<uint32_t> off(221 + 1);
  int64_t npts = 0;
  const int64_t nl = orc_streamlines(&g, &S, S.u0, acc.data(), feff.data(), 220, pts.data(), 3 * cap, sp.data(), cap,
                                     off.data(), (int64_t)off.size(), &npts);
  if (nl < 0) return 2;
  printf("oracle NF %lld gamma %.6f gammaCore %.6f zeta %.6f lines %lld points %lld\n", (long long)NF, pr[0],
         m.gammaCore, m.zeta, (long long)nl, (long long)npts);
  orc_sim_free(&S);
  orc_geom_free(&g);
  orc_free(tris);
  return 0;
}

std::string random_ascii(std::mt19937_64& R) {
  static const char* tok[] = {"vertex", "vertex ", " ", "\n", "\t", "1.5", "-2e3", "e", ".", "+-", "1e", "facet",
                              "0", "7", "-.5", "3.", "1e+", "endloop", "vertexvertex", "1e400", "-0", "9.9e-320"};
  std::string s;
  const int n = (int)(R() % 80);
  for (int i = 0; i < n; ++i) s += tok[R() % (sizeof tok / sizeof *tok)];
  return s;
}

int cmd_fuzz(uint64_t seed, int n) {
  std::mt19937_64 R(seed);
  for (int it = 0; it < n; ++it) {
    // STL reader: raw bytes, binary records with true or false counts, ASCII
    std::vector<uint8_t> b;
    const int kind = (int)(R() % 4);
    if (kind == 0) {
      b.resize(R() % 400);
      for (auto& c : b) c = (uint8_t)R();
    } else if (kind == 1) {
      const uint32_t nt = (uint32_t)(R() % 6);
      b.assign(84 + 50 * (size_t)nt + (R() % 3 == 0 ? (R() % 5) : 0), 0);
      for (size_t i = 84; i < b.size(); ++i) b[i] = (uint8_t)R();
      const uint32_t claim = R() % 4 == 0 ? (uint32_t)R() : nt;
      memcpy(b.data() + 80, &claim, 4);
    } else {
      const std::string s = random_ascii(R);
      b.assign(s.begin(), s.end());
    }
    check_parse(b.data(), (int64_t)b.size(), "fuzz stl");

    // grid stage: degenerate sizes, non-finite coordinates, bad axes
    {
      const int ntri = (int)(R() % 4);
      std::vector<float> T((size_t)(9 * ntri));
      for (auto& v : T) {
        const int r = (int)(R() % 20);
        v = r == 0 ? NAN : (r == 1 ? INFINITY : (float)((double)(R() % 2000) / 7.0 - 100));
      }
      abx_prep_params p;
      memset(&p, 0, sizeof p);
      const double dxs[] = {0.0, -1.0, NAN, 1e-9, 0.5, 2.0, 37.0, 1e300};
      p.dx = dxs[R() % 8];
      p.sealAxis = (int)(R() % 4) - (R() % 8 == 0);
      p.inletAxis = (int)(R() % 3);
      p.sealSign = R() % 2 ? 1 : -1;
      p.inletSign = R() % 2 ? 1 : -1;
      p.inletYaw = R() % 3 ? (double)(R() % 90) : NAN;
      p.cy = NAN;
      p.sy = NAN;
      p.has_outlet = (int)(R() % 2);
      p.out_height = (double)(R() % 50);
      p.max_cells = R() % 2 ? 4.5e6 : 0;
      abx::GeomHost g;
      std::vector<float> Tl;
      (void)abx::prep_grid(T.data(), (int64_t)T.size() - (int64_t)(R() % 2), &p, g, Tl);
    }
    // opening stage on a random slice of a small grid
    {
      abx::GeomHost g;
      abx_geom_info& I = g.info;
      memset(&I, 0, sizeof I);
      I.nx = 1 + (int)(R() % 12);
      I.ny = 1 + (int)(R() % 12);
      I.nz = 1 + (int)(R() % 5);
      I.N = (int64_t)I.nx * I.ny * I.nz;
      I.h = 0.5 + (double)(R() % 8);
      for (int a = 0; a < 3; ++a) {
        I.origin[a] = -(double)(R() % 20);
        I.mn[a] = I.origin[a] + I.h;
        I.mx[a] = I.origin[a] + I.h * 4;
      }
      g.fr = {1, 2, 0, -1, -1, 0.0, 1.0, 0.0, 0.0};
      abx_prep_params p;
      memset(&p, 0, sizeof p);
      p.has_click = (int)(R() % 2);
      for (double& c : p.click) c = (double)(R() % 40) - 20;
      std::vector<uint8_t> empty((size_t)9 * I.nx * I.ny);
      const int dens = 1 + (int)(R() % 4);
      for (auto& e : empty) e = (R() % 4) < (uint64_t)dens ? 1 : 0;
      std::vector<uint8_t> floor_row;
      if (abx::prep_openings(&p, empty.data(), g) == 0) {
        int n_open = 0;
        for (uint8_t o : g.opening) n_open += o;
        if (n_open != I.nOpen || n_open < 8) fail("fuzz openings: count");
        (void)abx::prep_flood_inputs(g, floor_row);
      }
    }
  }
  printf("fuzz seed %llu: %d cases\n", (unsigned long long)seed, n);
  return g_fail;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 4 && !strcmp(argv[1], "prepare")) return cmd_prepare(argv[2], argv[3]);
  if (argc >= 5 && !strcmp(argv[1], "oracle")) return cmd_oracle(argv[2], argv[3], atoi(argv[4]));
  if (argc >= 4 && !strcmp(argv[1], "fuzz")) return cmd_fuzz(strtoull(argv[2], nullptr, 10), atoi(argv[3]));
  if (argc >= 2 && !strcmp(argv[1], "device")) {
    abx_dev* d = nullptr;
    char name[64];
    const int r = abx_device_open(0, &d, name, sizeof name);
    printf("device_open -> %d (%s)\n", r, r ? abx_last_error() : name);
    if (d) abx_device_close(d);
    return 0;
  }
  fprintf(stderr, "usage: host_checks prepare STL PARAMS | oracle STL PARAMS N | fuzz SEED N | device\n");
  return 2;
}
