// Shared definitions of the gfx950 AirboxCalc core: error state, device
// handle, D3Q19 lattice constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "../../include/airbox_hip.h"
#include "abx_error.hpp"

namespace abx {

#define ABX_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::abx::set_errorf("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                      \
      return -1;                                                                        \
    }                                                                                   \
  } while (0)

// D3Q19 velocity set in the reference's numbering (engine.js:4-9): rest,
// +-x, +-y, +-z, the four xy, xz and yz diagonals.
constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
// Weights as double literals 1/3, 1/18, 1/36 (engine.js:7-8).
__host__ __device__ constexpr double wt(int q) { return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }

}  // namespace abx

struct abx_dev {
  int ordinal;
  char name[256];
  int ncu;
  int xcds;
  int lds_per_cu;  // bytes of LDS per CU (the workgroups resident on a CU share it)
};
