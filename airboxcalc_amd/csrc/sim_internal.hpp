// Read-only view of a simulation handle for the post-processing kernels.
#pragma once

#include "abx_common.hpp"

namespace abx {

struct SimView {
  int ordinal;
  hipStream_t stream;
  int B;
  int nx, ny, nz, nyp, nzp;
  int64_t plane;
  size_t ssz;              // bytes per stored value (4 or 8)
  const void* acc;         // [B][4][plane]
  const float* ns;         // [B][plane] or null
  const uint32_t* fluid_bits;
  double u0;
  int Ain, Af;
};

SimView sim_view(abx_sim* s);

}  // namespace abx
