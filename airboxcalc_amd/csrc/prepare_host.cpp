// Host stages of geometry preparation (engine.js:110-234); abx_prepare in
// prepare.hip runs them around the GPU voxelizer and flood fill.
#include "prepare_host.hpp"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "abx_error.hpp"

namespace abx {

std::string js_num(double v) {
  char buf[64];
  for (int prec = 1; prec <= 17; ++prec) {
    snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (strtod(buf, nullptr) == v) break;
  }
  return buf;
}

// The nearest one-decimal value, ties away from zero.
std::string js_fixed1(double x) {
  const double y = x * 10, fl = floor(y);
  const double n = (y - fl == 0.5) ? fl + 1 : nearbyint(y);
  char buf[64];
  snprintf(buf, sizeof buf, "%.1f", n / 10);
  return buf;
}

int first_ge(const std::vector<float>& a, double v, bool or_last) {
  for (size_t i = 0; i < a.size(); ++i)
    if ((double)a[i] >= v) return (int)i;
  return or_last ? (int)a.size() - 1 : -1;
}

int prep_grid(const float* T0, int64_t nfloat, const abx_prep_params* p, GeomHost& g, std::vector<float>& T) {
  Frame& fr = g.fr;
  fr.dAx = p->sealAxis;
  fr.iAx = p->inletAxis;
  if (fr.dAx < 0 || fr.dAx > 2 || fr.iAx < 0 || fr.iAx > 2) {
    set_error("abx_prepare: axis out of range");
    return -1;
  }
  if (fr.dAx == fr.iAx) {
    set_error("Einlassseite und Filterebene liegen auf derselben Achse.");
    return -1;
  }
  fr.oAx = 3 - fr.dAx - fr.iAx;
  fr.sd = p->sealSign;
  fr.si = p->inletSign;
  fr.seal = p->seal;
  fr.yaw = p->inletYaw;
  const double yr = (std::isnan(p->inletYaw) ? 0.0 : p->inletYaw) * 3.141592653589793 / 180;
  fr.cy = std::isnan(p->cy) ? cos(yr) : p->cy;
  fr.sy = std::isnan(p->sy) ? sin(yr) : p->sy;

  // local frame and bounding box (engine.js:112-118)
  const int64_t ntri = nfloat / 9;
  T.assign((size_t)std::max<int64_t>(nfloat, 1), 0.0f);
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t v = 0; v < ntri * 3; ++v) {
    const double q[3] = {T0[3 * v], T0[3 * v + 1], T0[3 * v + 2]};
    double L[3];
    fr.to_local(q, L);
    for (int a = 0; a < 3; ++a) {
      T[(size_t)(3 * v + a)] = (float)L[a];
      if (L[a] < mn[a]) mn[a] = L[a];
      if (L[a] > mx[a]) mx[a] = L[a];
    }
  }
  // grid (engine.js:119-125)
  const double h = p->dx;
  const bool has_out = p->has_outlet != 0;
  const double origin[3] = {mn[0] - h, (has_out ? std::min(mn[1], -p->out_height) : mn[1]) - h, mn[2] - 10 * h};
  const double ndx = ceil((mx[0] - origin[0] + h) / h), ndy = ceil((mx[1] - origin[1] + h) / h),
               ndz = ceil((mx[2] - origin[2] + h) / h);
  // every extent of the fine grid (3 per cell) must fit an int
  constexpr double kMaxDim = 715827882.0;  // INT32_MAX / 3
  if (!(ndx >= 1 && ndy >= 1 && ndz >= 1) || ndx > kMaxDim || ndy > kMaxDim || ndz > kMaxDim ||
      ndx * ndy * ndz > 4.0e9) {
    set_error("abx_prepare: degenerate or unbounded grid");
    return -1;
  }
  const int nx = (int)ndx, ny = (int)ndy, nz = (int)ndz;
  const int64_t N = (int64_t)nx * ny * nz;
  if (p->max_cells > 0 && (double)N > p->max_cells) {
    set_errorf("Gitter zu groß (%s Mio. Zellen) – Zellgröße erhöhen.", js_fixed1((double)N / 1e6).c_str());
    return -1;
  }
  abx_geom_info& I = g.info;
  memset(&I, 0, sizeof I);
  I.nx = nx;
  I.ny = ny;
  I.nz = nz;
  I.N = N;
  I.h = h;
  for (int a = 0; a < 3; ++a) {
    I.origin[a] = origin[a];
    I.mn[a] = mn[a];
    I.mx[a] = mx[a];
  }
  I.dAx = fr.dAx;
  I.iAx = fr.iAx;
  I.oAx = fr.oAx;
  I.sd = fr.sd;
  I.si = fr.si;
  I.seal = fr.seal;
  I.yaw = std::isnan(p->inletYaw) ? 0.0 : p->inletYaw;
  I.cy = fr.cy;
  I.sy = fr.sy;
  I.pleat = p->pleat;
  I.has_outlet = has_out ? 1 : 0;
  return 0;
}

int prep_slice_k(const GeomHost& g) {
  const abx_geom_info& I = g.info;
  return (int)floor((I.mn[2] + 1.5 - I.origin[2]) / (I.h / 3));
}

int prep_openings(const abx_prep_params* p, const uint8_t* empty, GeomHost& g) {
  abx_geom_info& I = g.info;
  const int nx = I.nx, ny = I.ny, nz = I.nz, nxf = 3 * nx, nyf = 3 * ny;
  const double h = I.h, hf = h / 3;
  const double* origin = I.origin;
  g.xc.resize((size_t)nx);
  g.dc.resize((size_t)ny);
  g.zc.resize((size_t)nz);
  for (int i = 0; i < nx; ++i) g.xc[(size_t)i] = (float)(origin[0] + (i + 0.5) * h);
  for (int j = 0; j < ny; ++j) g.dc[(size_t)j] = (float)(origin[1] + (j + 0.5) * h);
  for (int k = 0; k < nz; ++k) g.zc[(size_t)k] = (float)(origin[2] + (k + 0.5) * h);

  // openings: holes of the first material layer not connected to the border
  // (engine.js:141-154); fine cells of the slice are indexed with 32 bits
  const int64_t n2 = (int64_t)nxf * nyf;
  if (n2 >= (int64_t)INT32_MAX) {
    set_error("abx_prepare: fine slice too large (2^31 cells)");
    return -1;
  }
  std::vector<uint8_t> outside((size_t)n2, 0);
  std::vector<int32_t> q((size_t)n2);
  int64_t qh = 0, qt = 0;
  auto push = [&](int64_t c) {
    if (empty[(size_t)c] && !outside[(size_t)c]) {
      outside[(size_t)c] = 1;
      q[(size_t)qt++] = (int32_t)c;
    }
  };
  for (int i = 0; i < nxf; ++i) {
    push((int64_t)i * nyf);
    push((int64_t)i * nyf + nyf - 1);
  }
  for (int j = 0; j < nyf; ++j) {
    push(j);
    push((int64_t)(nxf - 1) * nyf + j);
  }
  while (qh < qt) {
    const int64_t c = q[(size_t)qh++];
    const int i = (int)(c / nyf), j = (int)(c % nyf);
    if (i > 0) push(c - nyf);
    if (i < nxf - 1) push(c + nyf);
    if (j > 0) push(c - 1);
    if (j < nyf - 1) push(c + 1);
  }
  std::vector<uint8_t> holes((size_t)n2);
  for (int64_t c = 0; c < n2; ++c)
    holes[(size_t)c] = (empty[(size_t)c] && !outside[(size_t)c] && (origin[1] + ((double)(c % nyf) + 0.5) * hf > 0));
  // 4-connected components in scan order, cells in BFS order (engine.js:95-108)
  std::vector<int32_t> lab((size_t)n2, -1);
  std::vector<std::vector<int32_t>> comps;
  for (int64_t s0 = 0; s0 < n2; ++s0) {
    if (!holes[(size_t)s0] || lab[(size_t)s0] >= 0) continue;
    const int32_t id = (int32_t)comps.size();
    std::vector<int32_t> cells;
    cells.push_back((int32_t)s0);
    lab[(size_t)s0] = id;
    for (size_t h0 = 0; h0 < cells.size(); ++h0) {
      const int32_t c = cells[h0];
      const int i = c / nyf, j = c % nyf;
      const int32_t nbs[4] = {i > 0 ? c - nyf : -1, i < nxf - 1 ? c + nyf : -1, j > 0 ? c - 1 : -1,
                              j < nyf - 1 ? c + 1 : -1};
      for (int32_t d : nbs)
        if (d >= 0 && holes[(size_t)d] && lab[(size_t)d] < 0) {
          lab[(size_t)d] = id;
          cells.push_back(d);
        }
    }
    if ((double)cells.size() * hf * hf > 25) comps.push_back(std::move(cells));
  }
  I.nHoles = (int)comps.size();
  if (comps.empty()) {
    set_error("Auf der gewählten Einlassseite wurde keine Öffnung gefunden.");
    return -1;
  }
  // the component nearest the click (every third cell, engine.js:157-165),
  // else the largest
  size_t best = 0;
  if (p->has_click) {
    double L[3];
    g.fr.to_local(p->click, L);
    double bd = INFINITY;
    for (size_t ci = 0; ci < comps.size(); ++ci)
      for (size_t s = 0; s < comps[ci].size(); s += 3) {
        const int32_t c = comps[ci][s];
        const double x = origin[0] + ((double)(c / nyf) + 0.5) * hf, y = origin[1] + ((double)(c % nyf) + 0.5) * hf;
        const double ex = x - L[0], ey = y - L[1];
        const double d = ex * ex + ey * ey;
        if (d < bd) {
          bd = d;
          best = ci;
        }
      }
  } else {
    for (size_t ci = 1; ci < comps.size(); ++ci)
      if (comps[ci].size() > comps[best].size()) best = ci;
  }
  // a cell opens when at least 5 of its 9 fine cells do (engine.js:166-170)
  std::vector<uint8_t> sel((size_t)nx * ny, 0);
  for (int32_t c : comps[best]) sel[(size_t)((c / nyf) / 3) * ny + (size_t)((c % nyf) / 3)]++;
  g.opening.assign((size_t)nx * ny, 0);
  int nOpen = 0;
  for (size_t c = 0; c < sel.size(); ++c)
    if (sel[c] >= 5) {
      g.opening[c] = 1;
      nOpen++;
    }
  I.nOpen = nOpen;
  if (nOpen < 8) {
    set_error("Einlassöffnung ist zu klein für die gewählte Zellgröße.");
    return -1;
  }
  return 0;
}

int prep_flood_inputs(GeomHost& g, std::vector<uint8_t>& floor_row) {
  abx_geom_info& I = g.info;
  I.kw0 = first_ge(g.zc, I.mn[2], false);
  I.top = first_ge(g.dc, 0.0, false);
  if (I.top < 0) {
    set_error("Dichtebene liegt außerhalb des Modells.");
    return -1;
  }
  floor_row.assign((size_t)I.ny, 0);
  for (int j = 0; j < I.ny; ++j) floor_row[(size_t)j] = (double)g.dc[(size_t)j] > I.mx[1] - 0.5 * I.h ? 1 : 0;
  return 0;
}

int prep_finish(const abx_prep_params* p, int64_t nfluid, bool leaked, GeomHost& g) {
  abx_geom_info& I = g.info;
  const int nx = I.nx, ny = I.ny, nz = I.nz, top = I.top;
  const double h = I.h;
  const int64_t N = I.N;
  const Frame& fr = g.fr;
  const std::vector<float>&xc = g.xc, &dc = g.dc, &zc = g.zc;
  std::vector<uint8_t>& fluid = g.fluid;
  std::vector<uint8_t>& solid = g.solid;
  if ((int64_t)fluid.size() != N || (int64_t)solid.size() != N || (int64_t)g.opening.size() != (int64_t)nx * ny) {
    set_error("abx_prepare: inconsistent flood result");
    return -1;
  }
  if (leaked) {
    set_error("Bei " + js_num(h) +
              " mm Zellgröße läuft die Luft durch eine zu dünne Wand nach außen (Wände sollten mindestens so "
              "dick wie eine Zelle sein). Bitte eine kleinere Zellgröße wählen.");
    return -1;
  }
  auto IDX = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * nz + k; };
  int64_t qt = nfluid;
  // clean-air side: lid plenum and outlet disc, or the whole top row (engine.js:197-213)
  g.outlet.assign((size_t)N, 0);
  int nOut = 0;
  if (I.has_outlet) {
    const int jpt = first_ge(dc, -p->out_height, true);
    if (jpt >= top) {
      set_error("Deckelhöhe ist kleiner als eine Zelle.");
      return -1;
    }
    double L[3];
    fr.to_local(p->out_center, L);
    const double rr = p->out_diameter / 2, r2 = rr * rr;
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (!fluid[(size_t)IDX(i, top, k)]) continue;
        for (int j = jpt; j < top; ++j) {
          const int64_t c = IDX(i, j, k);
          fluid[(size_t)c] = 1;
          solid[(size_t)c] = 0;
          qt++;
        }
        const double ex = (double)xc[(size_t)i] - L[0], ez = (double)zc[(size_t)k] - L[2];
        if (ex * ex + ez * ez <= r2) {
          g.outlet[(size_t)IDX(i, jpt, k)] = 1;
          nOut++;
          double m[3];
          fr.to_model(xc[(size_t)i], dc[(size_t)jpt], zc[(size_t)k], m);
          g.outletPts.insert(g.outletPts.end(), m, m + 3);
        }
      }
    if (nOut < 4) {
      set_error("Der Auslass liegt nicht über der Filterfläche oder ist zu klein.");
      return -1;
    }
  } else {
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const int64_t c = IDX(i, top, k);
        if (fluid[(size_t)c]) {
          g.outlet[(size_t)c] = 1;
          nOut++;
        }
      }
  }
  I.nOut = nOut;
  I.nf = qt;
  // filter plane, points and summary (engine.js:215-233)
  const int jm = first_ge(dc, p->pleat / 2, true);
  I.jm = jm;
  int Af = 0;
  for (int i = 0; i < nx; ++i)
    for (int k = 0; k < nz; ++k) Af += fluid[(size_t)IDX(i, jm, k)];
  I.Af = Af;
  for (int i = 0; i < nx; ++i)
    for (int k = 0; k < nz; ++k)
      if (fluid[(size_t)IDX(i, jm, k)]) {
        double m[3];
        fr.to_model(xc[(size_t)i], dc[(size_t)jm], zc[(size_t)k], m);
        g.filterPts.insert(g.filterPts.end(), m, m + 3);
      }
  for (int64_t c = 0; c < (int64_t)nx * ny; ++c)
    if (g.opening[(size_t)c]) {
      double m[3];
      fr.to_model(xc[(size_t)(c / ny)], dc[(size_t)(c % ny)], I.mn[2], m);
      g.openPts.insert(g.openPts.end(), m, m + 3);
    }
  int leak = 0;
  for (int i = 0; i < nx; ++i)
    for (int k = 0; k < nz; ++k) leak += fluid[(size_t)IDX(i, top, k)];
  I.leak = leak;
  I.nOpenPts = (int64_t)g.openPts.size() / 3;
  I.nFilterPts = (int64_t)g.filterPts.size() / 3;
  I.nOutletPts = (int64_t)g.outletPts.size() / 3;
  I.volume_l = (double)I.nf * pow(h, 3) / 1e6;
  I.inlet_cm2 = I.nOpen * h * h / 100;
  I.filter_cm2 = Af * h * h / 100;
  I.outlet_cm2 = leak * h * h / 100;
  I.outlet_cm2_real = I.has_outlet ? nOut * h * h / 100 : 0;
  I.pleatOk = jm > top;
  return 0;
}

}  // namespace abx
