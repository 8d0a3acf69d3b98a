// Read-only view of a simulation handle for the post-processing kernels.
#pragma once

#include <vector>

#include "abx_common.hpp"

namespace abx {

struct SimView {
  int ordinal;
  hipStream_t stream;
  int B;
  int nx, ny, nz;
  const int32_t* rowid;    // nx*ny: row of (i, j), -1 without fluid
  const int4* rowinfo;     // per row {rowoff, kmin, kmax, i*ny + j}: cell k at rowoff + k
  int64_t plane;
  size_t ssz;              // bytes per stored value (4 or 8)
  const void* acc;         // [B][4][plane]
  const float* ns;         // [B][plane] or null
  const uint32_t* fluid_bits;
  double u0;
  int Ain, Af;
};

SimView sim_view(abx_sim* s);

struct StepArgs;
// Launch arguments of the handle's next step (its current lattice pointers).
StepArgs sim_step_args(abx_sim* s, int doacc);
// Stored form of the populations (0 plain, 1 AA swapped, 2 AA streamed).
int sim_form(abx_sim* s);
void sim_set_form(abx_sim* s, int form);
// Bring an AA-pattern handle to the swapped form (host round trip if needed).
int sim_normalize(abx_sim* s);
// Host copy of the handle's active 64-position segments (ascending).
std::vector<int32_t> sim_segments(abx_sim* s);

}  // namespace abx
