// Library plumbing: error state, version, device handle, STL reader.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "abx_common.hpp"

namespace abx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

void set_errorf(const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

const char* last_error() { return g_err.c_str(); }

}  // namespace abx

using namespace abx;

extern "C" const char* abx_last_error(void) { return last_error(); }

extern "C" const char* abx_version(void) { return "airbox-hip 0.1 (gfx950)"; }

extern "C" int abx_device_open(int ordinal, abx_dev** dev, char* name, size_t name_len) {
  if (!dev) {
    set_error("abx_device_open: null argument");
    return -1;
  }
  *dev = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_errorf("no HIP GPU available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return -1;
  }
  if (ordinal < 0 || ordinal >= n) {
    set_errorf("GPU %d not present (%d devices)", ordinal, n);
    return -1;
  }
  hipDeviceProp_t p;
  ABX_HIP(hipGetDeviceProperties(&p, ordinal));
  if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    set_errorf("GPU %d is %s; this library is built for gfx950 (MI355X) only", ordinal, p.gcnArchName);
    return -1;
  }
  ABX_HIP(hipSetDevice(ordinal));
  abx_dev* d = new abx_dev();
  d->ordinal = ordinal;
  snprintf(d->name, sizeof d->name, "%s (%s, %d CUs)", p.name[0] ? p.name : "AMD Instinct GPU", p.gcnArchName,
           p.multiProcessorCount);
  d->ncu = p.multiProcessorCount;
  d->xcds = 8;
  d->lds_per_cu = p.maxSharedMemoryPerMultiProcessor > 0 ? (int)p.maxSharedMemoryPerMultiProcessor : 160 * 1024;
  if (name && name_len) snprintf(name, name_len, "%s", d->name);
  *dev = d;
  return 0;
}

extern "C" void abx_device_close(abx_dev* dev) { delete dev; }

// ---- parseSTL (engine.js:15-28) -------------------------------------------
namespace {

bool js_space(unsigned char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }
bool num_char(unsigned char c) {
  return (c >= '0' && c <= '9') || c == '-' || c == '+' || c == '.' || c == 'e' || c == 'E';
}

// Number(token) for a token of [-+0-9.eE]: a decimal literal, else NaN.
double to_number(const char* s, size_t n) {
  if (n == 0 || n > 120) return NAN;
  std::string b(s, n);
  size_t i = 0;
  if (b[i] == '+' || b[i] == '-') i++;
  size_t d0 = i;
  while (i < n && b[i] >= '0' && b[i] <= '9') i++;
  size_t digits = i - d0;
  if (i < n && b[i] == '.') {
    i++;
    size_t f0 = i;
    while (i < n && b[i] >= '0' && b[i] <= '9') i++;
    digits += i - f0;
  }
  if (digits == 0) return NAN;
  if (i < n && (b[i] == 'e' || b[i] == 'E')) {
    i++;
    if (i < n && (b[i] == '+' || b[i] == '-')) i++;
    size_t e0 = i;
    while (i < n && b[i] >= '0' && b[i] <= '9') i++;
    if (i == e0) return NAN;
  }
  if (i != n) return NAN;
  return strtod(b.c_str(), nullptr);
}

}  // namespace

extern "C" int abx_parse_stl(const uint8_t* buf, int64_t len, float* out, int64_t cap, int64_t* nfloat) {
  if ((!buf && len > 0) || !nfloat || len < 0) {
    set_error("abx_parse_stl: bad argument");
    return -1;
  }
  uint32_t n = 0;
  if (len >= 84) memcpy(&n, buf + 80, 4);
  if (n > 0 && 84 + 50 * (int64_t)n == len) {
    *nfloat = 9 * (int64_t)n;
    if (out)
      for (int64_t t = 0; t < (int64_t)n; ++t)
        for (int k = 0; k < 9; ++k)
          if (9 * t + k < cap) memcpy(out + 9 * t + k, buf + 84 + 50 * t + 12 + 4 * k, 4);
    return 0;
  }
  // ASCII: every "vertex <ws> num <ws> num <ws> num" record, left to right
  int64_t cnt = 0;
  for (int64_t p = 0; p + 6 <= len;) {
    if (memcmp(buf + p, "vertex", 6) != 0) {
      ++p;
      continue;
    }
    int64_t q = p + 6, ts[3], tn[3];
    bool ok = true;
    for (int c = 0; c < 3 && ok; ++c) {
      int64_t w = q;
      while (q < len && js_space(buf[q])) ++q;
      if (q == w) ok = false;
      int64_t s = q;
      while (ok && q < len && num_char(buf[q])) ++q;
      if (ok && q == s) ok = false;
      ts[c] = s;
      tn[c] = q - s;
    }
    if (!ok) {
      ++p;
      continue;
    }
    for (int c = 0; c < 3; ++c) {
      if (out && cnt < cap) out[cnt] = (float)to_number((const char*)buf + ts[c], (size_t)tn[c]);
      ++cnt;
    }
    p = q;
  }
  if (cnt < 9) {
    set_error("STL enthält keine Dreiecke.");
    return -1;
  }
  *nfloat = cnt;
  return 0;
}
