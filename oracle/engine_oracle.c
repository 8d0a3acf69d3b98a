/*
 * ORACLE - test infrastructure only (see engine_oracle.h).
 *
 * Plain-C restatement of src/engine.js.  Each function names the lines it
 * follows.  JavaScript semantics kept on purpose:
 *   - all scalar arithmetic in double, left-to-right as written in the JS;
 *   - Float32Array / Uint8Array stores round or truncate like the typed arrays;
 *   - Math.min/Math.max propagate NaN; Math.round rounds half up;
 *   - Math.hypot is V8's normalised, compensated sum.
 * Built with -O2 -ffp-contract=off -fno-fast-math so no FMA is formed.
 */
#include "engine_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static double WT[19];
static int wt_init = 0;
static void init_wt(void) {
  /* engine.js:7-8: 1/3, 1/18 x6, 1/36 x12 as doubles */
  if (wt_init) return;
  WT[0] = 1.0 / 3.0;
  for (int i = 1; i < 7; i++) WT[i] = 1.0 / 18.0;
  for (int i = 7; i < 19; i++) WT[i] = 1.0 / 36.0;
  wt_init = 1;
}

static __thread char g_err[1024];
const char* orc_last_error(void) { return g_err; }
static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}
void orc_free(void* p) { free(p); }

/* ---- JS helpers ------------------------------------------------------- */
static double js_min(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a < b) return a;
  if (b < a) return b;
  return signbit(a) ? a : b;
}
static double js_max(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}
double orc_js_round(double x) {
  /* Math.round: nearest integer, ties towards +Infinity */
  if (!isfinite(x)) return x;
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}
double orc_js_hypot3(double a, double b, double c) {
  /* V8 Math.hypot: normalise by the largest magnitude, Kahan-compensated sum */
  double v[3] = {a, b, c}, ab[3];
  int nan = 0;
  double mx = 0;
  for (int i = 0; i < 3; i++) {
    if (isnan(v[i])) {
      nan = 1;
      ab[i] = 0;
    } else {
      ab[i] = fabs(v[i]);
      if (ab[i] > mx) mx = ab[i];
    }
  }
  if (mx == INFINITY) return INFINITY;
  if (nan) return NAN;
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  for (int i = 0; i < 3; i++) {
    double n = ab[i] / mx;
    double summand = n * n - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return sqrt(sum) * mx;
}
/* firstGE (engine.js:13) over a Float32Array */
static int firstGE_f32(const float* arr, int n, double v) {
  for (int i = 0; i < n; i++)
    if ((double)arr[i] >= v) return i;
  return n - 1;
}
static int findIndex_ge(const float* arr, int n, double v) {
  for (int i = 0; i < n; i++)
    if ((double)arr[i] >= v) return i;
  return -1;
}

/* ---- parseSTL (engine.js:15-28) --------------------------------------- */
static int is_js_space(unsigned char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v';
}
static int is_num_char(unsigned char c) {
  return (c >= '0' && c <= '9') || c == '-' || c == '+' || c == '.' || c == 'e' || c == 'E';
}
static double js_to_number(const char* s, size_t n) {
  /* ToNumber of a [-+0-9.eE]+ token: the whole token must be a decimal literal */
  char buf[128];
  if (n >= sizeof buf) return NAN;
  memcpy(buf, s, n);
  buf[n] = 0;
  /* JS StringNumericLiteral: [+-]? (digits [. digits?] | . digits) ([eE][+-]?digits)? */
  size_t i = 0;
  if (buf[i] == '+' || buf[i] == '-') i++;
  size_t d0 = i;
  while (buf[i] >= '0' && buf[i] <= '9') i++;
  size_t nint = i - d0, nfrac = 0;
  if (buf[i] == '.') {
    i++;
    size_t f0 = i;
    while (buf[i] >= '0' && buf[i] <= '9') i++;
    nfrac = i - f0;
  }
  if (nint + nfrac == 0) return NAN;
  if (buf[i] == 'e' || buf[i] == 'E') {
    i++;
    if (buf[i] == '+' || buf[i] == '-') i++;
    size_t e0 = i;
    while (buf[i] >= '0' && buf[i] <= '9') i++;
    if (i == e0) return NAN;
  }
  if (i != n) return NAN;
  return strtod(buf, NULL);
}

int orc_parse_stl(const uint8_t* buf, int64_t len, float** out, int64_t* nfloat) {
  uint32_t n = 0;
  if (len >= 84) memcpy(&n, buf + 80, 4);
  if (n > 0 && 84 + 50 * (int64_t)n == len) {
    float* t = (float*)malloc(sizeof(float) * 9 * (size_t)n);
    for (uint32_t i = 0; i < n; i++)
      for (int k = 0; k < 9; k++) memcpy(&t[i * 9 + k], buf + 84 + (size_t)i * 50 + 12 + k * 4, 4);
    *out = t;
    *nfloat = 9 * (int64_t)n;
    return 0;
  }
  /* ASCII: /vertex\s+(num)\s+(num)\s+(num)/g over the decoded text */
  size_t cap = 1024, cnt = 0;
  double* v = (double*)malloc(sizeof(double) * cap);
  int64_t p = 0;
  while (p + 6 <= len) {
    if (memcmp(buf + p, "vertex", 6) != 0) {
      p++;
      continue;
    }
    int64_t q = p + 6;
    size_t tok_s[3], tok_n[3];
    int ok = 1;
    for (int c = 0; c < 3 && ok; c++) {
      int64_t ws = q;
      while (q < len && is_js_space(buf[q])) q++;
      if (q == ws) {
        ok = 0;
        break;
      }
      int64_t ts = q;
      while (q < len && is_num_char(buf[q])) q++;
      if (q == ts) {
        ok = 0;
        break;
      }
      tok_s[c] = (size_t)ts;
      tok_n[c] = (size_t)(q - ts);
    }
    if (!ok) {
      p++;
      continue;
    }
    if (cnt + 3 > cap) {
      cap *= 2;
      v = (double*)realloc(v, sizeof(double) * cap);
    }
    for (int c = 0; c < 3; c++) v[cnt++] = js_to_number((const char*)buf + tok_s[c], tok_n[c]);
    p = q;
  }
  if (cnt < 9) {
    free(v);
    return fail("STL enthält keine Dreiecke.");
  }
  float* t = (float*)malloc(sizeof(float) * cnt);
  for (size_t i = 0; i < cnt; i++) t[i] = (float)v[i];
  free(v);
  *out = t;
  *nfloat = (int64_t)cnt;
  return 0;
}

/* ---- makeFrame (engine.js:31-43) -------------------------------------- */
typedef struct {
  int dAx, iAx, oAx, sd, si;
  double seal, cy, sy;
} frame_t;
static void to_local(const frame_t* fr, const double q[3], double L[3]) {
  double a = q[fr->oAx], b = -fr->si * q[fr->iAx];
  L[0] = fr->cy * a + fr->sy * b;
  L[1] = -fr->sd * (q[fr->dAx] - fr->seal);
  L[2] = -fr->sy * a + fr->cy * b;
}
static void to_model(const frame_t* fr, double X, double D, double Z, double q[3]) {
  q[fr->oAx] = fr->cy * X - fr->sy * Z;
  q[fr->dAx] = fr->seal - fr->sd * D;
  q[fr->iAx] = -fr->si * (fr->sy * X + fr->cy * Z);
}

/* ---- voxelize (engine.js:45-93) --------------------------------------- */
static int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

static void vox_pass(const float* T, int64_t ntri, const double origin[3], int nx, int ny, double h, int pass,
                     int32_t* cnt, float* hits, int32_t* pos) {
  const double jx = 1.37e-4, jy = 2.71e-4;
  for (int64_t t = 0; t < ntri; t++) {
    const float* o = T + t * 9;
    double ax = (double)o[0] - origin[0] - jx, ay = (double)o[1] - origin[1] - jy, az = (double)o[2] - origin[2];
    double bx = (double)o[3] - origin[0] - jx, by = (double)o[4] - origin[1] - jy, bz = (double)o[5] - origin[2];
    double cx = (double)o[6] - origin[0] - jx, cy = (double)o[7] - origin[1] - jy, cz = (double)o[8] - origin[2];
    double den = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
    if (fabs(den) < 1e-12) continue;
    double mnx = js_min(js_min(ax, bx), cx), mxx = js_max(js_max(ax, bx), cx);
    double mny = js_min(js_min(ay, by), cy), mxy = js_max(js_max(ay, by), cy);
    double i0d = js_max(0, floor(mnx / h - 0.5)), i1d = js_min(nx - 1, ceil(mxx / h - 0.5));
    double j0d = js_max(0, floor(mny / h - 0.5)), j1d = js_min(ny - 1, ceil(mxy / h - 0.5));
    if (!(i0d <= i1d) || !(j0d <= j1d)) continue;
    int64_t i0 = (int64_t)i0d, i1 = (int64_t)i1d, j0 = (int64_t)j0d, j1 = (int64_t)j1d;
    for (int64_t i = i0; i <= i1; i++) {
      double px = ((double)i + 0.5) * h;
      for (int64_t j = j0; j <= j1; j++) {
        double py = ((double)j + 0.5) * h;
        double l1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / den;
        if (l1 < 0) continue;
        double l2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / den;
        if (l2 < 0) continue;
        double l3 = 1 - l1 - l2;
        if (l3 < 0) continue;
        int64_t col = i * ny + j;
        if (pass == 0)
          cnt[col + 1]++;
        else
          hits[pos[col]++] = (float)(l1 * az + l2 * bz + l3 * cz);
      }
    }
  }
}

int orc_voxel_counts(const float* T, int64_t ntri, const double origin[3], int nx, int ny, double h, int32_t* cnt) {
  int64_t ncol = (int64_t)nx * ny;
  int32_t* c = (int32_t*)calloc((size_t)ncol + 1, sizeof(int32_t));
  vox_pass(T, ntri, origin, nx, ny, h, 0, c, NULL, NULL);
  memcpy(cnt, c + 1, sizeof(int32_t) * (size_t)ncol);
  free(c);
  return 0;
}

int orc_voxelize(const float* T, int64_t ntri, const double origin[3], int nx, int ny, int nz, double h,
                 uint8_t* vox) {
  int64_t ncol = (int64_t)nx * ny;
  int32_t* cnt = (int32_t*)calloc((size_t)ncol + 1, sizeof(int32_t));
  vox_pass(T, ntri, origin, nx, ny, h, 0, cnt, NULL, NULL);
  for (int64_t c = 0; c < ncol; c++) cnt[c + 1] += cnt[c];
  float* hits = (float*)malloc(sizeof(float) * ((size_t)cnt[ncol] + 1));
  int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ncol + 1));
  memcpy(pos, cnt, sizeof(int32_t) * (size_t)ncol);
  vox_pass(T, ntri, origin, nx, ny, h, 1, cnt, hits, pos);
  memset(vox, 0, (size_t)ncol * (size_t)nz);
  double* zs = NULL;
  size_t zcap = 0;
  for (int64_t c = 0; c < ncol; c++) {
    int a = cnt[c], b = cnt[c + 1];
    if (b - a < 2) continue;
    size_t n = (size_t)(b - a);
    if (n > zcap) {
      zcap = n * 2;
      zs = (double*)realloc(zs, sizeof(double) * zcap);
    }
    for (size_t q = 0; q < n; q++) zs[q] = (double)hits[a + q];
    qsort(zs, n, sizeof(double), cmp_double);
    size_t m = n - (n % 2);
    int64_t base = c * nz;
    for (size_t k = 0; k < m; k += 2) {
      double k0 = js_max(0, ceil(zs[k] / h - 0.5)), k1 = js_min(nz - 1, ceil(zs[k + 1] / h - 0.5) - 1);
      for (double kk = k0; kk <= k1; kk++) vox[base + (int64_t)kk] = 1;
    }
  }
  free(zs);
  free(hits);
  free(pos);
  free(cnt);
  return 0;
}

/* ---- components2d (engine.js:95-108) ---------------------------------- */
typedef struct {
  int32_t* cells;
  int64_t n;
} comp_t;

static comp_t* components2d(const uint8_t* mask, int nx, int ny, int64_t* ncomp) {
  int64_t NN = (int64_t)nx * ny;
  int32_t* lab = (int32_t*)malloc(sizeof(int32_t) * (size_t)NN);
  for (int64_t s = 0; s < NN; s++) lab[s] = -1;
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)NN);
  comp_t* comps = NULL;
  int64_t nc = 0, cap = 0;
  for (int64_t s = 0; s < NN; s++) {
    if (!mask[s] || lab[s] >= 0) continue;
    int32_t id = (int32_t)nc;
    int64_t qh = 0, qt = 0;
    q[qt++] = (int32_t)s;
    lab[s] = id;
    while (qh < qt) {
      int32_t c = q[qh++];
      int i = c / ny, j = c % ny;
      int32_t nb[4] = {i > 0 ? c - ny : -1, i < nx - 1 ? c + ny : -1, j > 0 ? c - 1 : -1, j < ny - 1 ? c + 1 : -1};
      for (int d = 0; d < 4; d++)
        if (nb[d] >= 0 && mask[nb[d]] && lab[nb[d]] < 0) {
          lab[nb[d]] = id;
          q[qt++] = nb[d];
        }
    }
    if (nc == cap) {
      cap = cap ? cap * 2 : 16;
      comps = (comp_t*)realloc(comps, sizeof(comp_t) * (size_t)cap);
    }
    comps[nc].cells = (int32_t*)malloc(sizeof(int32_t) * (size_t)qt);
    memcpy(comps[nc].cells, q, sizeof(int32_t) * (size_t)qt); /* BFS dequeue order */
    comps[nc].n = qt;
    nc++;
  }
  free(lab);
  free(q);
  *ncomp = nc;
  return comps;
}

static void fmt_js_number(char* out, size_t n, double v) {
  /* enough for the values the messages carry (dx such as 4, 2.5, 0.75) */
  char buf[64];
  for (int prec = 1; prec <= 17; prec++) {
    snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (strtod(buf, NULL) == v) break;
  }
  snprintf(out, n, "%s", buf);
}

/* ---- prepare (engine.js:110-234) -------------------------------------- */
int orc_prepare(const float* T0, int64_t nfloat, const orc_prep_params* p, orc_geom* g) {
  memset(g, 0, sizeof *g);
  frame_t fr;
  fr.dAx = p->sealAxis;
  fr.iAx = p->inletAxis;
  if (fr.dAx == fr.iAx) return fail("Einlassseite und Filterebene liegen auf derselben Achse.");
  fr.oAx = 3 - fr.dAx - fr.iAx;
  fr.sd = p->sealSign;
  fr.si = p->inletSign;
  fr.seal = p->seal;
  fr.cy = p->cy;
  fr.sy = p->sy;
  g->oAx = fr.oAx;
  g->dAx = fr.dAx;
  g->iAx = fr.iAx;
  g->sd = fr.sd;
  g->si = fr.si;
  g->seal = fr.seal;
  g->cy = fr.cy;
  g->sy = fr.sy;
  g->yaw_deg = p->yaw_deg;
  g->pleat = p->pleat;
  int64_t ntri = nfloat / 9;
  float* T = (float*)malloc(sizeof(float) * (size_t)(nfloat > 0 ? nfloat : 1));
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t v = 0; v < ntri * 3; v++) {
    double q[3] = {T0[v * 3], T0[v * 3 + 1], T0[v * 3 + 2]}, L[3];
    to_local(&fr, q, L);
    for (int a = 0; a < 3; a++) {
      T[v * 3 + a] = (float)L[a];
      if (L[a] < mn[a]) mn[a] = L[a];
      if (L[a] > mx[a]) mx[a] = L[a];
    }
  }
  g->T = T;
  double h = p->dx, hf = h / 3;
  const int sub = 3, duct = 10;
  int has_out = p->has_outlet;
  g->has_outlet = has_out;
  double origin[3] = {mn[0] - h, (has_out ? js_min(mn[1], -p->out_height) : mn[1]) - h, mn[2] - duct * h};
  int nx = (int)ceil((mx[0] - origin[0] + h) / h), ny = (int)ceil((mx[1] - origin[1] + h) / h),
      nz = (int)ceil((mx[2] - origin[2] + h) / h);
  int64_t N = (int64_t)nx * ny * nz;
  memcpy(g->origin, origin, sizeof origin);
  memcpy(g->mn, mn, sizeof mn);
  memcpy(g->mx, mx, sizeof mx);
  g->h = h;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->N = N;
  if (p->max_cells > 0 && (double)N > p->max_cells) {
    char num[64];
    snprintf(num, sizeof num, "%.1f", (double)N / 1e6);
    return fail("Gitter zu groß (%s Mio. Zellen) – Zellgröße erhöhen.", num);
  }
  int nxf = nx * sub, nyf = ny * sub, nzf = nz * sub;
  g->nxf = nxf;
  g->nyf = nyf;
  g->nzf = nzf;
  uint8_t* fine = (uint8_t*)malloc((size_t)nxf * nyf * nzf);
  orc_voxelize(T, ntri, origin, nxf, nyf, nzf, hf, fine);
  g->fine = fine;
  uint8_t* solid = (uint8_t*)calloc((size_t)N, 1);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        int s = 0;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            int64_t base = ((int64_t)(i * 3 + a) * nyf + (j * 3 + b)) * nzf + k * 3;
            s += fine[base] + fine[base + 1] + fine[base + 2];
          }
        if (s >= 7) solid[((int64_t)i * ny + j) * nz + k] = 1;
      }
  g->solid = solid;
  float* xc = (float*)malloc(sizeof(float) * nx);
  float* dc = (float*)malloc(sizeof(float) * ny);
  float* zc = (float*)malloc(sizeof(float) * nz);
  for (int i = 0; i < nx; i++) xc[i] = (float)(origin[0] + (i + 0.5) * h);
  for (int j = 0; j < ny; j++) dc[j] = (float)(origin[1] + (j + 0.5) * h);
  for (int k = 0; k < nz; k++) zc[k] = (float)(origin[2] + (k + 0.5) * h);
  g->xc = xc;
  g->dc = dc;
  g->zc = zc;

  /* openings: holes in the first material layer (engine.js:141-171) */
  int kf = (int)floor((mn[2] + 1.5 - origin[2]) / hf);
  int64_t NF2 = (int64_t)nxf * nyf;
  uint8_t* empty = (uint8_t*)malloc((size_t)NF2);
  for (int i = 0; i < nxf; i++)
    for (int j = 0; j < nyf; j++) {
      int64_t src = ((int64_t)i * nyf + j) * nzf + kf;
      int solidf = (kf >= 0 && kf < nzf) ? fine[src] : 0;
      empty[(int64_t)i * nyf + j] = solidf ? 0 : 1;
    }
  uint8_t* outside = (uint8_t*)calloc((size_t)NF2, 1);
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)NF2);
  int64_t qh = 0, qt = 0;
#define PUSH(c)                      \
  do {                               \
    int64_t cc_ = (c);               \
    if (empty[cc_] && !outside[cc_]) { \
      outside[cc_] = 1;              \
      q[qt++] = (int32_t)cc_;        \
    }                                \
  } while (0)
  for (int i = 0; i < nxf; i++) {
    PUSH((int64_t)i * nyf);
    PUSH((int64_t)i * nyf + nyf - 1);
  }
  for (int j = 0; j < nyf; j++) {
    PUSH(j);
    PUSH((int64_t)(nxf - 1) * nyf + j);
  }
  while (qh < qt) {
    int64_t c = q[qh++];
    int i = (int)(c / nyf), j = (int)(c % nyf);
    if (i > 0) PUSH(c - nyf);
    if (i < nxf - 1) PUSH(c + nyf);
    if (j > 0) PUSH(c - 1);
    if (j < nyf - 1) PUSH(c + 1);
  }
#undef PUSH
  uint8_t* holes = (uint8_t*)malloc((size_t)NF2);
  for (int64_t c = 0; c < NF2; c++)
    holes[c] = (empty[c] && !outside[c] && (origin[1] + ((double)(c % nyf) + 0.5) * hf > 0)) ? 1 : 0;
  free(empty);
  free(outside);
  free(q);
  int64_t nall = 0;
  comp_t* all = components2d(holes, nxf, nyf, &nall);
  free(holes);
  /* filter: c.length * hf * hf > 25 */
  comp_t* comps = (comp_t*)malloc(sizeof(comp_t) * (size_t)(nall > 0 ? nall : 1));
  int64_t nc = 0;
  for (int64_t c = 0; c < nall; c++) {
    if ((double)all[c].n * hf * hf > 25)
      comps[nc++] = all[c];
    else
      free(all[c].cells);
  }
  free(all);
  g->nHoles = (int)nc;
  if (!nc) {
    free(comps);
    return fail("Auf der gewählten Einlassseite wurde keine Öffnung gefunden.");
  }
  comp_t* best = &comps[0];
  if (p->has_click) {
    double L[3];
    to_local(&fr, p->click, L);
    double bd = INFINITY;
    for (int64_t ci = 0; ci < nc; ci++) {
      for (int64_t s = 0; s < comps[ci].n; s += 3) {
        int32_t cs = comps[ci].cells[s];
        double x = origin[0] + ((double)(cs / nyf) + 0.5) * hf, y = origin[1] + ((double)(cs % nyf) + 0.5) * hf;
        double dxl = x - L[0], dyl = y - L[1];
        double d = dxl * dxl + dyl * dyl;
        if (d < bd) {
          bd = d;
          best = &comps[ci];
        }
      }
    }
  } else {
    for (int64_t ci = 1; ci < nc; ci++)
      if (comps[ci].n > best->n) best = &comps[ci];
  }
  uint8_t* sel = (uint8_t*)calloc((size_t)nx * ny, 1);
  for (int64_t s = 0; s < best->n; s++) {
    int32_t c = best->cells[s];
    sel[(int64_t)((c / nyf) / 3) * ny + ((c % nyf) / 3)]++;
  }
  for (int64_t ci = 0; ci < nc; ci++) free(comps[ci].cells);
  free(comps);
  uint8_t* opening = (uint8_t*)calloc((size_t)nx * ny, 1);
  int nOpen = 0;
  for (int64_t c = 0; c < (int64_t)nx * ny; c++)
    if (sel[c] >= 5) {
      opening[c] = 1;
      nOpen++;
    }
  free(sel);
  g->opening = opening;
  g->nOpen = nOpen;
  if (nOpen < 8) return fail("Einlassöffnung ist zu klein für die gewählte Zellgröße.");

  /* duct, floor, fluid flood, leak check, outlet (engine.js:173-214) */
  int kw0 = findIndex_ge(zc, nz, mn[2]);
  g->kw0 = kw0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < kw0; k++) solid[((int64_t)i * ny + j) * nz + k] = opening[(int64_t)i * ny + j] ? 0 : 1;
  for (int j = 0; j < ny; j++)
    if ((double)dc[j] > mx[1] - 0.5 * h)
      for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) solid[((int64_t)i * ny + j) * nz + k] = 1;
  int top = findIndex_ge(dc, ny, 0.0);
  if (top < 0) return fail("Dichtebene liegt außerhalb des Modells.");
  g->top = top;
  uint8_t* fluid = (uint8_t*)calloc((size_t)N, 1);
  g->fluid = fluid;
  int32_t* Q = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  qh = 0;
  qt = 0;
#define VISIT(c, j)                                  \
  do {                                               \
    int64_t cc_ = (c);                               \
    if (!fluid[cc_] && !solid[cc_] && (j) >= top) {  \
      fluid[cc_] = 1;                                \
      Q[qt++] = (int32_t)cc_;                        \
    }                                                \
  } while (0)
  for (int64_t c = 0; c < (int64_t)nx * ny; c++)
    if (opening[c]) VISIT(c * nz, (int)(c % ny));
  while (qh < qt) {
    int64_t c = Q[qh++];
    int k = (int)(c % nz);
    int64_t ij = c / nz;
    int j = (int)(ij % ny), i = (int)(ij / ny);
    if (i > 0) VISIT(c - (int64_t)ny * nz, j);
    if (i < nx - 1) VISIT(c + (int64_t)ny * nz, j);
    if (j > 0) VISIT(c - nz, j - 1);
    if (j < ny - 1) VISIT(c + nz, j + 1);
    if (k > 0) VISIT(c - 1, j);
    if (k < nz - 1) VISIT(c + 1, j);
  }
#undef VISIT
  free(Q);
  int leaked = 0;
  for (int i = 0; i < nx && !leaked; i++)
    for (int j = top; j < ny && !leaked; j++)
      for (int k = kw0; k < nz; k++) {
        int64_t c = ((int64_t)i * ny + j) * nz + k;
        if ((i == 0 || i == nx - 1 || k == nz - 1) && c >= 0 && fluid[c]) {
          leaked = 1;
          break;
        }
      }
  if (leaked) {
    char num[64];
    fmt_js_number(num, sizeof num, h);
    return fail(
        "Bei %s mm Zellgröße läuft die Luft durch eine zu dünne Wand nach außen (Wände sollten mindestens so dick "
        "wie eine Zelle sein). Bitte eine kleinere Zellgröße wählen.",
        num);
  }
  uint8_t* outlet = (uint8_t*)calloc((size_t)N, 1);
  g->outlet = outlet;
  int nOut = 0;
  int64_t cap_out = 64;
  double* outletPts = (double*)malloc(sizeof(double) * 3 * (size_t)cap_out);
  int64_t nOutPts = 0;
  if (has_out) {
    int jpt = firstGE_f32(dc, ny, -p->out_height);
    if (jpt >= top) {
      free(outletPts);
      return fail("Deckelhöhe ist kleiner als eine Zelle.");
    }
    double L[3];
    to_local(&fr, p->out_center, L);
    double rr = p->out_diameter / 2, r2 = rr * rr;
    for (int i = 0; i < nx; i++)
      for (int k = 0; k < nz; k++) {
        if (!fluid[((int64_t)i * ny + top) * nz + k]) continue;
        for (int j = jpt; j < top; j++) {
          int64_t c = ((int64_t)i * ny + j) * nz + k;
          fluid[c] = 1;
          solid[c] = 0;
          qt++;
        }
        double ddx = (double)xc[i] - L[0], ddz = (double)zc[k] - L[2];
        if (ddx * ddx + ddz * ddz <= r2) {
          outlet[((int64_t)i * ny + jpt) * nz + k] = 1;
          nOut++;
          if (nOutPts == cap_out) {
            cap_out *= 2;
            outletPts = (double*)realloc(outletPts, sizeof(double) * 3 * (size_t)cap_out);
          }
          to_model(&fr, xc[i], dc[jpt], zc[k], outletPts + 3 * nOutPts);
          nOutPts++;
        }
      }
    if (nOut < 4) {
      free(outletPts);
      return fail("Der Auslass liegt nicht über der Filterfläche oder ist zu klein.");
    }
  } else {
    for (int i = 0; i < nx; i++)
      for (int k = 0; k < nz; k++) {
        int64_t c = ((int64_t)i * ny + top) * nz + k;
        if (fluid[c]) {
          outlet[c] = 1;
          nOut++;
        }
      }
  }
  g->outletPts = outletPts;
  g->nOutletPts = nOutPts;
  g->nOut = nOut;
  g->nf = qt;
  int jm = firstGE_f32(dc, ny, p->pleat / 2);
  g->jm = jm;
  int Af = 0;
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) Af += fluid[((int64_t)i * ny + jm) * nz + k];
  g->Af = Af;
  double* filterPts = (double*)malloc(sizeof(double) * 3 * (size_t)(Af > 0 ? Af : 1));
  int64_t nfp = 0;
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++)
      if (fluid[((int64_t)i * ny + jm) * nz + k]) to_model(&fr, xc[i], dc[jm], zc[k], filterPts + 3 * nfp++);
  g->filterPts = filterPts;
  g->nFilterPts = nfp;
  double* openPts = (double*)malloc(sizeof(double) * 3 * (size_t)(nOpen > 0 ? nOpen : 1));
  int64_t nop = 0;
  for (int64_t c = 0; c < (int64_t)nx * ny; c++)
    if (opening[c]) to_model(&fr, xc[c / ny], dc[c % ny], mn[2], openPts + 3 * nop++);
  g->openPts = openPts;
  g->nOpenPts = nop;
  int leak = 0;
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) leak += fluid[((int64_t)i * ny + top) * nz + k];
  g->leak = leak;
  return 0;
}

void orc_geom_free(orc_geom* g) {
  free(g->fine);
  free(g->solid);
  free(g->fluid);
  free(g->opening);
  free(g->outlet);
  free(g->xc);
  free(g->dc);
  free(g->zc);
  free(g->T);
  free(g->openPts);
  free(g->filterPts);
  free(g->outletPts);
  memset(g, 0, sizeof *g);
}

/* ---- setupSim (engine.js:239-264) ------------------------------------- */
int orc_setup_sim(const orc_geom* g, double K, double aniso, double u, double tau, orc_sim* S) {
  init_wt();
  memset(S, 0, sizeof *S);
  int ny = g->ny, nz = g->nz;
  int64_t N = g->N, NF = g->nf;
  int32_t* cid = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  int32_t* cells = (int32_t*)malloc(sizeof(int32_t) * (size_t)(NF > 0 ? NF : 1));
  for (int64_t c = 0; c < N; c++) cid[c] = -1;
  int64_t m = 0;
  for (int64_t c = 0; c < N; c++)
    if (g->fluid[c]) {
      cid[c] = (int32_t)m;
      cells[m++] = (int32_t)c;
    }
  if (m != NF) return fail("fluid count mismatch");
  int64_t off[19];
  for (int i = 0; i < 19; i++) off[i] = ((int64_t)CX[i] * ny + CY[i]) * nz + CZ[i];
  int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * 19 * (size_t)(NF > 0 ? NF : 1));
  uint32_t* flag = (uint32_t*)calloc((size_t)(NF > 0 ? NF : 1), sizeof(uint32_t));
  int32_t* jsOf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(NF > 0 ? NF : 1));
  int Ain = 0, Lp = 0;
  const float* dc = g->dc;
  int top = g->top;
  for (int j = 0; j < ny; j++)
    if (dc[j] > dc[top] && (double)dc[j] <= g->pleat) Lp++;
  for (int64_t a = 0; a < NF; a++) {
    int64_t c = cells[a];
    int k = (int)(c % nz);
    int64_t ij = c / nz;
    int j = (int)(ij % ny);
    jsOf[a] = j;
    for (int i = 0; i < 19; i++) {
      int64_t s = c - off[i];
      nb[a * 19 + i] = (s >= 0 && s < N) ? cid[s] : -1;
    }
    if (k == 0 && g->opening[ij]) {
      flag[a] = 1;
      Ain++;
    } else if (g->outlet[c])
      flag[a] = 2;
    else if (dc[j] > dc[top] && (double)dc[j] <= g->pleat)
      flag[a] = 3;
  }
  int64_t nM = 0, nT = 0;
  for (int64_t a = 0; a < NF; a++) {
    if (jsOf[a] == g->jm) nM++;
    if (jsOf[a] == top + 1) nT++;
  }
  int32_t* layerM = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nM > 0 ? nM : 1));
  int32_t* layerT = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nT > 0 ? nT : 1));
  nM = nT = 0;
  for (int64_t a = 0; a < NF; a++) {
    if (jsOf[a] == g->jm) layerM[nM++] = (int32_t)a;
    if (jsOf[a] == top + 1) layerT[nT++] = (int32_t)a;
  }
  S->NF = NF;
  S->cid = cid;
  S->cells = cells;
  S->nb = nb;
  S->flag = flag;
  S->jsOf = jsOf;
  S->layerM = layerM;
  S->layerT = layerT;
  S->nM = nM;
  S->nT = nT;
  S->Ain = Ain;
  S->Af = (int)nM;
  S->Lp = Lp;
  double u0 = u, uface = u0 * Ain / (double)nM;
  S->u0 = u0;
  S->sig = K * 0.5 * u0 * u0 / (uface * (double)(Lp > 1 ? Lp : 1));
  S->tau0 = tau;
  S->aniso = aniso;
  S->cs2s = 18 * M_SQRT2 * 0.17 * 0.17;
  return 0;
}

void orc_sim_free(orc_sim* S) {
  free(S->cid);
  free(S->cells);
  free(S->nb);
  free(S->flag);
  free(S->jsOf);
  free(S->layerM);
  free(S->layerT);
  memset(S, 0, sizeof *S);
}

/* ---- cpuStep (engine.js:266-321) -------------------------------------- */
/* Cells a0 <= a < a1 of one step; every cell reads f and writes only its own
 * entries of f2 and acc, so disjoint ranges may run concurrently. */
#define CPU_STEP_BODY(LOAD, STORE, ACC_ADD)                                                                  \
  const int64_t NF = S->NF;                                                                                  \
  const double u0 = S->u0, tau0 = S->tau0, sig = S->sig, aniso = S->aniso, cs2s = S->cs2s;                 \
  for (int64_t a = a0; a < a1; a++) {                                                                        \
    double fi[19], fe[19];                                                                                   \
    const int64_t base = a * 19;                                                                             \
    double rho = 0, mx = 0, my = 0, mz = 0;                                                                  \
    for (int i = 0; i < 19; i++) {                                                                           \
      int32_t s = S->nb[base + i];                                                                           \
      double v = s >= 0 ? LOAD(f, (int64_t)i * NF + s) : LOAD(f, (int64_t)OPP[i] * NF + a);                 \
      fi[i] = v;                                                                                             \
      rho += v;                                                                                              \
      mx += CX[i] * v;                                                                                       \
      my += CY[i] * v;                                                                                       \
      mz += CZ[i] * v;                                                                                       \
    }                                                                                                        \
    double ux = mx / rho, uy = my / rho, uz = mz / rho;                                                      \
    uint32_t fl = S->flag[a];                                                                                \
    if (fl == 1 || fl == 2) {                                                                                \
      if (fl == 1) {                                                                                         \
        double s0 = 0, sm = 0;                                                                               \
        for (int i = 0; i < 19; i++) {                                                                       \
          if (CZ[i] == 0)                                                                                    \
            s0 += fi[i];                                                                                     \
          else if (CZ[i] < 0)                                                                                \
            sm += fi[i];                                                                                     \
        }                                                                                                    \
        rho = (s0 + 2 * sm) / (1 - u0);                                                                      \
        ux = 0;                                                                                              \
        uy = 0;                                                                                              \
        uz = u0;                                                                                             \
      } else                                                                                                 \
        rho = 1;                                                                                             \
      double usq = 1.5 * (ux * ux + uy * uy + uz * uz);                                                      \
      for (int i = 0; i < 19; i++) {                                                                         \
        double cu = 3 * (CX[i] * ux + CY[i] * uy + CZ[i] * uz);                                              \
        STORE(f2, (int64_t)i * NF + a, WT[i] * rho * (1 + cu + 0.5 * cu * cu - usq));                        \
      }                                                                                                      \
    } else {                                                                                                 \
      double usq = 1.5 * (ux * ux + uy * uy + uz * uz);                                                      \
      double pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;                                           \
      for (int i = 0; i < 19; i++) {                                                                         \
        double cu = 3 * ((double)CX[i] * ux + (double)CY[i] * uy + (double)CZ[i] * uz);                      \
        double e = WT[i] * rho * (1 + cu + 0.5 * cu * cu - usq);                                             \
        fe[i] = e;                                                                                           \
        double ne = fi[i] - e;                                                                               \
        pxx += (double)(CX[i] * CX[i]) * ne;                                                                 \
        pyy += (double)(CY[i] * CY[i]) * ne;                                                                 \
        pzz += (double)(CZ[i] * CZ[i]) * ne;                                                                 \
        pxy += (double)(CX[i] * CY[i]) * ne;                                                                 \
        pxz += (double)(CX[i] * CZ[i]) * ne;                                                                 \
        pyz += (double)(CY[i] * CZ[i]) * ne;                                                                 \
      }                                                                                                      \
      double Qn = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + 2 * (pxy * pxy + pxz * pxz + pyz * pyz));         \
      double tau = 0.5 * (tau0 + sqrt(tau0 * tau0 + cs2s * Qn / rho));                                       \
      if (fl == 3) {                                                                                         \
        double sl = 1 - js_min(tau * sig * aniso, 0.9), sv = 1 - js_min(tau * sig, 0.9);                      \
        double ex = ux * sl, ey = uy * sv, ez = uz * sl;                                                     \
        usq = 1.5 * (ex * ex + ey * ey + ez * ez);                                                           \
        for (int i = 0; i < 19; i++) {                                                                       \
          double cu = 3 * ((double)CX[i] * ex + (double)CY[i] * ey + (double)CZ[i] * ez);                    \
          fe[i] = WT[i] * rho * (1 + cu + 0.5 * cu * cu - usq);                                              \
        }                                                                                                    \
        ux = 0.5 * (ux + ex);                                                                                \
        uy = 0.5 * (uy + ey);                                                                                \
        uz = 0.5 * (uz + ez);                                                                                \
      }                                                                                                      \
      double it = 1 / tau, sv = ns ? (double)ns[a] : 0;                                                      \
      if (sv > 0) {                                                                                          \
        for (int i = 0; i < 19; i++)                                                                         \
          STORE(f2, (int64_t)i * NF + a, (1 - sv) * (fi[i] - (fi[i] - fe[i]) * it) + sv * fi[OPP[i]]);      \
        ux *= 1 - sv;                                                                                        \
        uy *= 1 - sv;                                                                                        \
        uz *= 1 - sv;                                                                                        \
      } else                                                                                                 \
        for (int i = 0; i < 19; i++) STORE(f2, (int64_t)i * NF + a, fi[i] - (fi[i] - fe[i]) * it);         \
    }                                                                                                        \
    if (acc) {                                                                                               \
      ACC_ADD(acc, a, ux);                                                                                   \
      ACC_ADD(acc, NF + a, uy);                                                                              \
      ACC_ADD(acc, 2 * NF + a, uz);                                                                          \
      ACC_ADD(acc, 3 * NF + a, rho);                                                                         \
    }                                                                                                        \
  }

#define LD(arr, idx) ((double)(arr)[idx])
#define ST32(arr, idx, val) ((arr)[idx] = (float)(val))
#define ST64(arr, idx, val) ((arr)[idx] = (double)(val))
#define ACC32(arr, idx, val) ((arr)[idx] = (float)((double)(arr)[idx] + (val)))
#define ACC64(arr, idx, val) ((arr)[idx] = (arr)[idx] + (val))

static void step_range32(const orc_sim* S, const float* f, float* f2, float* acc, const float* ns, int64_t a0,
                        int64_t a1) {
  CPU_STEP_BODY(LD, ST32, ACC32)
}
static void step_range64(const orc_sim* S, const double* f, double* f2, double* acc, const float* ns, int64_t a0,
                        int64_t a1) {
  CPU_STEP_BODY(LD, ST64, ACC64)
}

void orc_cpu_step(const orc_sim* S, const float* f, float* f2, float* acc, const float* ns) {
  init_wt();
  step_range32(S, f, f2, acc, ns, 0, S->NF);
}
void orc_cpu_step_f64(const orc_sim* S, const double* f, double* f2, double* acc, const float* ns) {
  init_wt();
  step_range64(S, f, f2, acc, ns, 0, S->NF);
}
/* The same step split into nthreads contiguous cell ranges run on OpenMP
 * threads (the host CPU baseline); results are identical to orc_cpu_step. */
void orc_cpu_step_threads(const orc_sim* S, const float* f, float* f2, float* acc, const float* ns, int nthreads) {
  init_wt();
  if (nthreads < 1) nthreads = 1;
  const int64_t NF = S->NF;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int t = 0; t < nthreads; t++)
    step_range32(S, f, f2, acc, ns, NF * t / nthreads, NF * (t + 1) / nthreads);
}

/* ---- probeCPU + progressInfo (engine.js:324-339) ---------------------- */
int orc_probe(const orc_sim* S, const float* f, int step, double out[3]) {
  const int64_t NF = S->NF, nM = S->nM, nT = S->nT;
  float* vy = (float*)malloc(sizeof(float) * (size_t)(nM + nT + 1));
  for (int64_t n = 0; n < nM + nT; n++) {
    int64_t a = n < nM ? S->layerM[n] : S->layerT[n - nM];
    double r = 0, my = 0;
    for (int i = 0; i < 19; i++) {
      double v = f[(int64_t)i * NF + a];
      r += v;
      my += CY[i] * v;
    }
    vy[n] = (float)(my / r);
  }
  double sw = 0;
  for (int64_t n = 0; n < nM; n++) sw += -(double)vy[n];
  double wm = sw / (double)nM, dev = 0;
  for (int64_t n = 0; n < nM; n++) dev += fabs(-(double)vy[n] - wm);
  double qo = 0;
  for (int64_t n = nM; n < nM + nT; n++) qo += -(double)vy[n];
  free(vy);
  out[2] = qo;
  if (!isfinite(qo))
    return fail("Rechnung instabil bei Schritt %d. Zellgröße vergrößern oder Viskosität (tau) erhöhen.", step);
  out[0] = 1 - dev / (2 * (double)nM * wm);
  out[1] = qo / (S->Ain * S->u0);
  return 0;
}

/* ---- simulate (engine.js:341-358), up to the averaged acc ------------- */
int orc_simulate(const orc_sim* S, int steps, int avg, int every, const float* ns, float* acc_out, double* hist,
                 int* nhist, int64_t step_limit) {
  init_wt();
  const int64_t NF = S->NF;
  float* f = (float*)malloc(sizeof(float) * 19 * (size_t)NF);
  float* f2 = (float*)calloc(19 * (size_t)NF, sizeof(float));
  for (int i = 0; i < 19; i++)
    for (int64_t a = 0; a < NF; a++) f[(int64_t)i * NF + a] = (float)WT[i];
  memset(acc_out, 0, sizeof(float) * 4 * (size_t)NF);
  int nacc = 0, nh = 0, rc = 0;
  for (int step = 1; step <= steps && (step_limit <= 0 || step <= step_limit); step++) {
    int doAcc = step > steps - avg;
    orc_cpu_step(S, f, f2, doAcc ? acc_out : NULL, ns);
    float* t = f;
    f = f2;
    f2 = t;
    if (doAcc) nacc++;
    if (step % every == 0 || step == steps) {
      double o[3];
      if (orc_probe(S, f, step, o)) {
        rc = -1;
        break;
      }
      if (hist) {
        hist[3 * nh] = step;
        hist[3 * nh + 1] = o[0];
        hist[3 * nh + 2] = o[1];
      }
      nh++;
    }
  }
  if (!rc) {
    double d = nacc > 1 ? nacc : 1;
    for (int64_t x = 0; x < 4 * NF; x++) acc_out[x] = (float)((double)acc_out[x] / d);
  }
  if (nhist) *nhist = nh;
  free(f);
  free(f2);
  return rc;
}

/* ---- evaluate (engine.js:632-691) ------------------------------------- */
static double Uc(const orc_sim* S, const float* acc, int64_t c, int comp) {
  int32_t a = S->cid[c];
  return a < 0 ? NAN : (double)acc[(int64_t)comp * S->NF + a];
}

int orc_evaluate(const orc_geom* g, const orc_sim* S, double u0, const float* acc, const float* ns, orc_metrics* M,
                 float* fmap, float* planInlet, float* planLow, float* lon, uint8_t* fluid) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t N = g->N;
  const double h = g->h;
  for (int64_t c = 0; c < N; c++) {
    if (ns) {
      int32_t a = S->cid[c];
      int nsbig = (a >= 0) ? ((double)ns[a] > 0.5) : 0;
      fluid[c] = (g->fluid[c] && !nsbig) ? 1 : 0;
    } else
      fluid[c] = g->fluid[c];
  }
#define IDX(i, j, k) (((int64_t)(i) * ny + (j)) * nz + (k))
  const int jm = g->jm;
  for (int64_t x = 0; x < (int64_t)nx * nz; x++) fmap[x] = NAN;
  double sw = 0;
  int n = 0;
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      int64_t c = IDX(i, jm, k);
      if (fluid[c]) {
        double w = -Uc(S, acc, c, 1);
        fmap[(int64_t)i * nz + k] = (float)w;
        sw += w;
        n++;
      }
    }
  double wm = sw / n, dev = 0, cdev = 0, csum = 0, mxr = 0;
  int cn = 0, low = 0, high = 0, rev = 0;
  uint8_t* core = (uint8_t*)calloc((size_t)nx * nz, 1);
  for (int i = 1; i < nx - 1; i++)
    for (int k = 1; k < nz - 1; k++) {
      if (fluid[IDX(i, jm, k)] && fluid[IDX(i - 1, jm, k)] && fluid[IDX(i + 1, jm, k)] && fluid[IDX(i, jm, k - 1)] &&
          fluid[IDX(i, jm, k + 1)])
        core[(int64_t)i * nz + k] = 1;
    }
  for (int64_t x = 0; x < (int64_t)nx * nz; x++)
    if (core[x]) {
      csum += fmap[x];
      cn++;
    }
  double cm = csum / cn;
  for (int64_t x = 0; x < (int64_t)nx * nz; x++) {
    double w = fmap[x];
    if (isnan(w)) continue;
    dev += fabs(w - wm);
    if (w < 0) rev++;
    mxr = js_max(mxr, w / wm);
    if (core[x]) {
      cdev += fabs(w - cm);
      if (w < 0.5 * cm) low++;
      if (w > 1.5 * cm) high++;
    }
    fmap[x] = (float)(w / wm);
  }
  free(core);
  double pin = 0, pb = 0;
  int pinN = 0, pbN = 0;
  int jb = firstGE_f32(g->dc, ny, g->pleat + 2 * h);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      if (g->opening[(int64_t)i * ny + j]) {
        int64_t c = IDX(i, j, 1);
        if (fluid[c]) {
          pin += Uc(S, acc, c, 3);
          pinN++;
        }
      }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      int64_t c = IDX(i, jb, k);
      if (fluid[c]) {
        pb += Uc(S, acc, c, 3);
        pbN++;
      }
    }
  double zeta = ((pin / pinN - pb / pbN) / 3) / (0.5 * u0 * u0);
  double si = 0, sj = 0, so = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      if (g->opening[(int64_t)i * ny + j]) {
        si += i;
        sj += j;
        so++;
      }
  int ic = (int)orc_js_round(si / so), jc = (int)orc_js_round(sj / so);
  int jlow = firstGE_f32(g->dc, ny, g->pleat + 12);
  float* planes[2] = {planInlet, planLow};
  int pj[2] = {jc, jlow};
  for (int pl = 0; pl < 2; pl++) {
    float *s = planes[pl], *pa = planes[pl] + (int64_t)nx * nz, *pbp = planes[pl] + 2 * (int64_t)nx * nz;
    for (int64_t x = 0; x < (int64_t)nx * nz; x++) {
      s[x] = NAN;
      pa[x] = 0;
      pbp[x] = 0;
    }
    int j = pj[pl];
    for (int i = 0; i < nx; i++)
      for (int k = 0; k < nz; k++) {
        int64_t c = IDX(i, j, k);
        if (!fluid[c]) continue;
        double ux = Uc(S, acc, c, 0) / u0, uy = Uc(S, acc, c, 1) / u0, uz = Uc(S, acc, c, 2) / u0;
        s[(int64_t)i * nz + k] = (float)orc_js_hypot3(ux, uy, uz);
        pa[(int64_t)i * nz + k] = (float)ux;
        pbp[(int64_t)i * nz + k] = (float)uz;
      }
  }
  {
    float *s = lon, *la = lon + (int64_t)nz * ny, *lb = lon + 2 * (int64_t)nz * ny;
    for (int64_t x = 0; x < (int64_t)nz * ny; x++) {
      s[x] = NAN;
      la[x] = 0;
      lb[x] = 0;
    }
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        int64_t c = IDX(ic, j, k);
        if (!fluid[c]) continue;
        double ux = Uc(S, acc, c, 0) / u0, uy = Uc(S, acc, c, 1) / u0, uz = Uc(S, acc, c, 2) / u0;
        s[(int64_t)k * ny + j] = (float)orc_js_hypot3(ux, uy, uz);
        la[(int64_t)k * ny + j] = (float)uz;
        lb[(int64_t)k * ny + j] = (float)uy;
      }
  }
#undef IDX
  M->gamma = 1 - dev / (2 * n * wm);
  M->gammaCore = 1 - cdev / (2 * cn * cm);
  M->lowCore = (double)low / cn;
  M->highCore = (double)high / cn;
  M->reverse = (double)rev / n;
  M->maxToMean = mxr;
  M->zeta = zeta;
  M->inlet_cm2 = S->Ain * h * h / 100;
  M->filter_cm2 = S->Af * h * h / 100;
  M->ic = ic;
  M->jc = jc;
  M->jlow = jlow;
  M->jb = jb;
  M->rhoRef = pb / pbN;
  return 0;
}

/* ---- field3d (engine.js:698-719) -------------------------------------- */
int orc_field3d(const orc_geom* g, const orc_sim* S, double u0, const float* acc, const uint8_t* fluid, double rhoRef,
                float* spd, float* pst, float* loss, int* kB_out, double* pin_out) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t NF = S->NF, N = g->N;
  const double q = 0.5 * u0 * u0;
#define PTOT(a)                                                                                    \
  ((double)acc[3 * NF + (a)] / 3 +                                                                 \
   0.5 * ((double)acc[a] * acc[a] + (double)acc[NF + (a)] * acc[NF + (a)] +                        \
          (double)acc[2 * NF + (a)] * acc[2 * NF + (a)]))
  int kw = findIndex_ge(g->zc, nz, g->mn[2]);
  int kB = (kw - 2) > 2 ? (kw - 2) : 2;
  double pin = 0, win = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      if (g->opening[(int64_t)i * ny + j]) {
        int64_t c = ((int64_t)i * ny + j) * nz + kB;
        int32_t a = S->cid[c];
        if (a < 0 || !fluid[c]) continue;
        double w = js_max(acc[2 * NF + a], 0);
        pin += w * PTOT(a);
        win += w;
      }
  pin /= js_max(win, 1e-12);
  for (int64_t c = 0; c < N; c++) {
    spd[c] = NAN;
    pst[c] = NAN;
    loss[c] = NAN;
    int32_t a = S->cid[c];
    if (a < 0 || !fluid[c] || c % nz < kB) continue;
    double ux = acc[a], uy = acc[NF + a], uz = acc[2 * NF + a];
    spd[c] = (float)(orc_js_hypot3(ux, uy, uz) / u0);
    pst[c] = (float)(((double)acc[3 * NF + a] - rhoRef) / 3 / q);
    loss[c] = (float)((pin - PTOT(a)) / q);
  }
#undef PTOT
  if (kB_out) *kB_out = kB;
  if (pin_out) *pin_out = pin;
  return 0;
}

/* ---- streamlines (engine.js:721-758) ---------------------------------- */
typedef struct {
  const orc_geom* g;
  const orc_sim* S;
  const float* acc;
  const uint8_t* fluid;
} sl_ctx;

static double sl_vel(const sl_ctx* X, double x, double y, double z, double out[3]) {
  const int nx = X->g->nx, ny = X->g->ny, nz = X->g->nz;
  const int64_t NF = X->S->NF;
  double i0 = floor(x), j0 = floor(y), k0 = floor(z), fx = x - i0, fy = y - j0, fz = z - k0;
  out[0] = out[1] = out[2] = 0;
  double wsum = 0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int c = 0; c < 2; c++) {
        double i = i0 + a, j = j0 + b, k = k0 + c;
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) continue;
        double w = (a ? fx : 1 - fx) * (b ? fy : 1 - fy) * (c ? fz : 1 - fz);
        int64_t cc = ((int64_t)i * ny + (int64_t)j) * nz + (int64_t)k;
        int32_t q = X->S->cid[cc];
        if (q < 0 || !X->fluid[cc]) continue;
        out[0] += w * X->acc[q];
        out[1] += w * X->acc[NF + q];
        out[2] += w * X->acc[2 * NF + q];
        wsum += w;
      }
  return wsum;
}
static int sl_inside(const sl_ctx* X, double x, double y, double z) {
  const int nx = X->g->nx, ny = X->g->ny, nz = X->g->nz;
  double i = orc_js_round(x), j = orc_js_round(y), k = orc_js_round(z);
  if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return 0;
  return X->fluid[((int64_t)i * ny + (int64_t)j) * nz + (int64_t)k] != 0;
}

int64_t orc_streamlines(const orc_geom* g, const orc_sim* S, double u0, const float* acc, const uint8_t* fluid,
                        int maxSeeds, float* pts, int64_t pts_cap, float* spd, int64_t spd_cap, uint32_t* off,
                        int64_t off_cap, int64_t* npts_out) {
  sl_ctx X = {g, S, acc, fluid};
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const double h = g->h;
  frame_t fr = {g->dAx, g->iAx, g->oAx, g->sd, g->si, g->seal, g->cy, g->sy};
  int kw = findIndex_ge(g->zc, nz, g->mn[2]);
  int kS = (kw - 3) > 2 ? (kw - 3) : 2;
  int64_t nseed = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      if (g->opening[(int64_t)i * ny + j]) nseed += 4;
  double* seeds = (double*)malloc(sizeof(double) * 3 * (size_t)(nseed > 0 ? nseed : 1));
  const double offs[4][2] = {{-0.25, -0.25}, {0.25, -0.25}, {-0.25, 0.25}, {0.25, 0.25}};
  int64_t ns_ = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      if (g->opening[(int64_t)i * ny + j])
        for (int d = 0; d < 4; d++) {
          seeds[3 * ns_] = i + offs[d][0];
          seeds[3 * ns_ + 1] = j + offs[d][1];
          seeds[3 * ns_ + 2] = kS;
          ns_++;
        }
  double stride = js_max(1, (double)nseed / maxSeeds);
  const double ds = 0.35;
  double* P = (double*)malloc(sizeof(double) * 3 * 6001);
  double* V = (double*)malloc(sizeof(double) * 6001);
  int64_t npts = 0, nlines = 0;
  if (off_cap > 0) off[0] = 0;
  for (double s = 0; s < nseed; s += stride) {
    int64_t si = (int64_t)floor(s);
    double x = seeds[3 * si], y = seeds[3 * si + 1], z = seeds[3 * si + 2];
    int64_t np = 0, nv = 0;
    double v1[3], v2[3];
    for (int st = 0; st < 6000; st++) {
      if (!sl_inside(&X, x, y, z) || sl_vel(&X, x, y, z, v1) < 0.3) break;
      double m1 = orc_js_hypot3(v1[0], v1[1], v1[2]);
      if (m1 < 0.01 * u0) break;
      if (st % 3 == 0) {
        P[np++] = x;
        P[np++] = y;
        P[np++] = z;
        V[nv++] = m1 / u0;
      }
      double xm = x + 0.5 * ds * v1[0] / m1, ym = y + 0.5 * ds * v1[1] / m1, zm = z + 0.5 * ds * v1[2] / m1;
      if (sl_vel(&X, xm, ym, zm, v2) < 0.3) break;
      double m2 = orc_js_hypot3(v2[0], v2[1], v2[2]);
      if (m2 < 0.01 * u0) break;
      x += ds * v2[0] / m2;
      y += ds * v2[1] / m2;
      z += ds * v2[2] / m2;
      if (y < g->top + 0.5) {
        P[np++] = x;
        P[np++] = y;
        P[np++] = z;
        V[nv++] = m2 / u0;
        break;
      }
    }
    if (np < 6) continue;
    for (int64_t q = 0; q < np; q += 3) {
      double m[3];
      to_model(&fr, g->xc[0] + P[q] * h, g->dc[0] + P[q + 1] * h, g->zc[0] + P[q + 2] * h, m);
      if (3 * npts + 2 < pts_cap) {
        pts[3 * npts] = (float)m[0];
        pts[3 * npts + 1] = (float)m[1];
        pts[3 * npts + 2] = (float)m[2];
      }
      if (npts < spd_cap) spd[npts] = (float)V[q / 3];
      npts++;
    }
    nlines++;
    if (nlines < off_cap) off[nlines] = (uint32_t)npts;
  }
  free(P);
  free(V);
  free(seeds);
  if (npts_out) *npts_out = npts;
  return nlines;
}
