// Batched solves: many simulations of different geometries and parameters
// (uneven grid sizes) stepped together by one launch per step.
//
// The reference batches B variants of one geometry (GpuSim, engine.js:495-545,
// used by its self-test).  Here a batch joins independent simulation handles:
// their active 64-position segments form one work list of (member, segment)
// items, each member's part padded to whole workgroups, and a workgroup reads
// its member's launch arguments from a device table.  Every cell is computed by
// the same device code as a single-handle launch (seg_cell), so each member's
// result is bit-identical to solving it alone.  Launch shape (segments per
// workgroup, persistent grid, consecutive workgroups per XCD) is chosen by
// abx_batch_tune; chunks of steps can be captured once as a HIP graph and
// replayed.
#include <algorithm>
#include <cstring>
#include <map>
#include <utility>
#include <string>
#include <tuple>
#include <vector>

#include "abx_common.hpp"
#include "lbm_device.hpp"
#include "sim_internal.hpp"

namespace abx {

template <typename S, typename M, bool ACC, int MODE, int SPB, bool PERSIST>
__global__ __launch_bounds__(64 * SPB) void k_batch_step(const StepArgs* __restrict__ args,
                                                         const int2* __restrict__ work, int nblocks) {
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  for (int blk = blockIdx.x; blk < nblocks; blk += PERSIST ? gridDim.x : nblocks) {
    const int2 w = work[blk * SPB + wv];  // (member, segment); segment -1 = padding
    const int m = __builtin_amdgcn_readfirstlane(w.x), seg = __builtin_amdgcn_readfirstlane(w.y);
    // non-temporal accesses as a single handle's default: the even AA step's
    // populations and averaged fields (float storage)
    if (seg >= 0) seg_cell<S, M, ACC, false, MODE, (sizeof(S) == 4 && MODE == 1) ? 19 : 0>(args[m], seg, l, 0);
    if (!PERSIST) break;
  }
}

}  // namespace abx

using namespace abx;

struct abx_batch {
  abx_dev* dev = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<abx_sim*> members;
  int prec = ABX_F32;
  int spb = 4, persistent = 0, use_graph = 0, xcd_group = 4;
  int occ_even = 16;  // waves per CU of the even AA step (0: no cap), as a single handle's default
  StepArgs* d_args[2] = {nullptr, nullptr};  // member launch arguments without / with averaging
  std::vector<std::vector<int32_t>> segs;    // every member's active segments (lattice order)
  // work lists by (segments per workgroup, workgroups per XCD group): the
  // device list and its number of blocks
  std::map<std::pair<int, int>, std::pair<int2*, int>> lists;
  int64_t cells = 0;
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;  // (n, doacc, start form)
};

namespace {

// Work list for spb segments per workgroup and XCD groups of g workgroups:
// every member's active segments in lattice order, in member order, each
// member's part padded to whole workgroups.  Workgroup b runs on XCD b % 8 as
// that XCD's s-th workgroup (s = b / 8); it takes list block
// (s / g) 8 g + (b % 8) g + s % g, so each XCD steps through groups of g
// consecutive blocks and all XCDs sweep the lattices together (the
// single-handle launch's block order).  Built once per shape, before any
// launch of it (so a captured graph never uploads).
int ensure_list(abx_batch* bt, int spb, int g) {
  const auto key = std::make_pair(spb, g);
  if (bt->lists.count(key)) return 0;
  std::vector<int2> lin;
  for (size_t m = 0; m < bt->segs.size(); ++m) {
    for (int32_t sg : bt->segs[m]) lin.push_back(make_int2((int)m, sg));
    while (lin.size() % (size_t)spb) lin.push_back(make_int2((int)m, -1));
  }
  const size_t G = (size_t)g, nlin = lin.size() / (size_t)spb, nblk = (nlin + 8 * G - 1) / (8 * G) * (8 * G);
  std::vector<int2> w(nblk * (size_t)spb, make_int2(0, -1));
  for (size_t blk = 0; blk < nblk; ++blk) {
    const size_t sq = blk / 8, src = (sq / G) * 8 * G + (blk % 8) * G + sq % G;
    if (src < nlin)
      for (size_t i = 0; i < (size_t)spb; ++i) w[blk * (size_t)spb + i] = lin[src * (size_t)spb + i];
  }
  int2* d = nullptr;
  if (hipMalloc((void**)&d, std::max<size_t>(w.size(), 1) * sizeof(int2)) != hipSuccess ||
      hipMemcpy(d, w.data(), w.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
    if (d) (void)hipFree(d);
    set_error("abx_batch: work list upload failed");
    return -1;
  }
  bt->lists[key] = std::make_pair(d, (int)(w.size() / (size_t)spb));
  return 0;
}

void drop_graphs(abx_batch* bt) {
  for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
  bt->graphs.clear();
}

// Launch arguments of every member (one lattice each, fixed for the batch's
// lifetime), without and with averaging.  Uploaded once, so a step launch
// carries no copy and captures cleanly into a graph.
int upload_args(abx_batch* bt) {
  for (int doacc = 0; doacc < 2; ++doacc) {
    std::vector<StepArgs> h(bt->members.size());
    for (size_t m = 0; m < bt->members.size(); ++m) h[m] = sim_step_args(bt->members[m], doacc);
    if (hipMalloc((void**)&bt->d_args[doacc], h.size() * sizeof(StepArgs)) != hipSuccess ||
        hipMemcpy(bt->d_args[doacc], h.data(), h.size() * sizeof(StepArgs), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("abx_batch_create: argument table upload failed");
      return -1;
    }
  }
  return 0;
}

template <typename S, typename M, bool ACC, int MODE>
void launch_b(abx_batch* bt) {
  const StepArgs* args = bt->d_args[ACC ? 1 : 0];
  // the even step's occupancy cap, as reserved LDS (no more workgroups fit on
  // a CU; the kernel uses no LDS)
  const int wgs = MODE == 1 && bt->occ_even > 0 ? std::max(1, bt->occ_even / bt->spb) : 0;
  const size_t lds = wgs ? (size_t)(bt->dev->lds_per_cu / (wgs + 1) + 1024) : 0;
  const auto& L = bt->lists.at(std::make_pair(bt->spb, bt->xcd_group));
  const int2* w = L.first;
  const int nb = L.second;
  const unsigned grid = bt->persistent ? (unsigned)std::min(nb, bt->dev->ncu * 8) : (unsigned)nb;
#define ABX_BL(SPB)                                                                                                \
  do {                                                                                                             \
    if (bt->persistent)                                                                                            \
      hipLaunchKernelGGL((k_batch_step<S, M, ACC, MODE, SPB, true>), dim3(grid), dim3(64 * SPB), lds, bt->stream,   \
                         args, w, nb);                                                                             \
    else                                                                                                           \
      hipLaunchKernelGGL((k_batch_step<S, M, ACC, MODE, SPB, false>), dim3(grid), dim3(64 * SPB), lds, bt->stream,  \
                         args, w, nb);                                                                             \
  } while (0)
  if (bt->spb == 2)
    ABX_BL(2);
  else if (bt->spb == 4)
    ABX_BL(4);
  else
    ABX_BL(8);
#undef ABX_BL
}

template <typename S, typename M>
void launch_bt(abx_batch* bt, int doacc, int mode) {
  if (doacc) {
    if (mode == 1) launch_b<S, M, true, 1>(bt);
    else launch_b<S, M, true, 2>(bt);
  } else {
    if (mode == 1) launch_b<S, M, false, 1>(bt);
    else launch_b<S, M, false, 2>(bt);
  }
}

// One batched step: AA even when the members hold the streamed form, odd otherwise.
int batch_step(abx_batch* bt, int doacc) {
  if (bt->lists.at(std::make_pair(bt->spb, bt->xcd_group)).second == 0) return 0;
  const int form = sim_form(bt->members[0]);
  const int mode = form == 2 ? 1 : 2;
  if (bt->prec == ABX_F32)
    launch_bt<float, float>(bt, doacc, mode);
  else if (bt->prec == ABX_F32_F64)
    launch_bt<float, double>(bt, doacc, mode);
  else
    launch_bt<double, double>(bt, doacc, mode);
  ABX_HIP(hipGetLastError());
  for (abx_sim* s : bt->members) sim_set_form(s, form == 2 ? 1 : 2);
  return 0;
}

}  // namespace

extern "C" int abx_batch_create(abx_dev* dev, abx_sim* const* sims, int n, abx_batch** out) {
  if (!dev || !sims || n < 1 || !out) {
    set_error("abx_batch_create: bad argument");
    return -1;
  }
  *out = nullptr;
  abx_batch* bt = new abx_batch();
  bt->dev = dev;
  ABX_HIP(hipSetDevice(dev->ordinal));
  for (int m = 0; m < n; ++m) {
    abx_sim* s = sims[m];
    if (!s) {
      delete bt;
      set_error("abx_batch_create: null member");
      return -1;
    }
    const SimView v = sim_view(s);
    abx_sim_info I;
    abx_sim_info_get(s, &I);
    if (m == 0) bt->prec = I.precision;
    if (I.precision != bt->prec || v.B != 1 || v.ns) {
      delete bt;
      set_error("abx_batch_create: members need one precision, one variant each and no ns");
      return -1;
    }
    // every member on the AA pattern, in the same (swapped) form
    if (abx_sim_set_launch(s, 4, 0, 0) || sim_normalize(s)) {
      delete bt;
      return -1;
    }
    bt->members.push_back(s);
    bt->cells += I.NF;
  }
  // every member's segments, and the default shape's work list
  bt->segs.resize((size_t)n);
  for (int m = 0; m < n; ++m) bt->segs[(size_t)m] = sim_segments(bt->members[(size_t)m]);
  if (ensure_list(bt, bt->spb, bt->xcd_group) || upload_args(bt)) {
    abx_batch_destroy(bt);
    return -1;
  }
  if (hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&bt->ev0) != hipSuccess || hipEventCreate(&bt->ev1) != hipSuccess) {
    set_error("abx_batch_create: allocation failed");
    abx_batch_destroy(bt);
    return -1;
  }
  // the members' own streams must be idle before the batch stream touches their lattices
  for (abx_sim* s : bt->members) abx_sim_sync(s);
  *out = bt;
  return 0;
}

extern "C" int abx_batch_set_launch(abx_batch* bt, int spb, int persistent, int graph) {
  if (!bt || (spb != 2 && spb != 4 && spb != 8)) {
    set_error("abx_batch_set_launch: spb must be 2, 4 or 8");
    return -1;
  }
  if (ensure_list(bt, spb, bt->xcd_group)) return -1;
  bt->spb = spb;
  bt->persistent = persistent ? 1 : 0;
  bt->use_graph = graph ? 1 : 0;
  drop_graphs(bt);
  return 0;
}

extern "C" int abx_batch_set_xcd_group(abx_batch* bt, int group) {
  if (!bt || (group != 1 && group != 2 && group != 4 && group != 8 && group != 16)) {
    set_error("abx_batch_set_xcd_group: group must be 1, 2, 4, 8 or 16");
    return -1;
  }
  if (ensure_list(bt, bt->spb, group)) return -1;
  bt->xcd_group = group;
  drop_graphs(bt);
  return 0;
}

extern "C" int abx_batch_get_xcd_group(const abx_batch* bt) { return bt ? bt->xcd_group : -1; }

extern "C" int abx_batch_set_occupancy(abx_batch* bt, int even_waves_per_cu) {
  if (!bt || even_waves_per_cu < 0 || even_waves_per_cu > 64) {
    set_error("abx_batch_set_occupancy: waves per CU must be 0 (no cap) to 64");
    return -1;
  }
  bt->occ_even = even_waves_per_cu;
  drop_graphs(bt);
  return 0;
}

extern "C" int abx_batch_get_occupancy(const abx_batch* bt) { return bt ? bt->occ_even : -1; }

extern "C" int abx_batch_steps(abx_batch* bt, int n, int doacc) {
  if (!bt || n < 0) {
    set_error("abx_batch_steps: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(bt->dev->ordinal));
  doacc = doacc ? 1 : 0;
  // the members' own streams (probes, reads, resets) must be done first
  for (abx_sim* s : bt->members)
    if (abx_sim_sync(s)) return -1;
  if (!bt->use_graph || n == 0) {
    for (int i = 0; i < n; ++i)
      if (batch_step(bt, doacc)) return -1;
    return 0;
  }
  // the launch arguments only change with the stored form, so a captured chunk
  // of n steps replays exactly
  const int form0 = sim_form(bt->members[0]);
  const auto key = std::make_tuple(n, doacc, form0);
  auto it = bt->graphs.find(key);
  if (it == bt->graphs.end()) {
    hipGraph_t g;
    ABX_HIP(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int i = 0; i < n && !rc; ++i) rc = batch_step(bt, doacc);
    hipError_t e = hipStreamEndCapture(bt->stream, &g);
    if (rc || e != hipSuccess) {
      if (!rc) set_errorf("abx_batch_steps: graph capture failed: %s", hipGetErrorString(e));
      return -1;
    }
    hipGraphExec_t ex;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      set_errorf("abx_batch_steps: graph instantiation failed: %s", hipGetErrorString(e));
      return -1;
    }
    // capture recorded the form changes; put them back before replaying
    for (abx_sim* s : bt->members) sim_set_form(s, form0);
    it = bt->graphs.emplace(key, ex).first;
  }
  ABX_HIP(hipGraphLaunch(it->second, bt->stream));
  const int form_end = (n % 2) ? (form0 == 2 ? 1 : 2) : form0;
  for (abx_sim* s : bt->members) sim_set_form(s, form_end);
  return 0;
}

extern "C" int abx_batch_time_steps(abx_batch* bt, int n, int doacc, double* ms) {
  if (!bt || !ms) {
    set_error("abx_batch_time_steps: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(bt->dev->ordinal));
  ABX_HIP(hipEventRecord(bt->ev0, bt->stream));
  if (abx_batch_steps(bt, n, doacc)) return -1;
  ABX_HIP(hipEventRecord(bt->ev1, bt->stream));
  ABX_HIP(hipEventSynchronize(bt->ev1));
  float f = 0;
  ABX_HIP(hipEventElapsedTime(&f, bt->ev0, bt->ev1));
  *ms = f;
  return 0;
}

extern "C" int abx_batch_sync(abx_batch* bt) {
  if (!bt) {
    set_error("abx_batch_sync: null argument");
    return -1;
  }
  ABX_HIP(hipStreamSynchronize(bt->stream));
  return 0;
}

extern "C" int abx_batch_probe(abx_batch* bt, int step, double* out) {
  if (!bt || !out) {
    set_error("abx_batch_probe: bad argument");
    return -1;
  }
  if (abx_batch_sync(bt)) return -1;
  for (size_t m = 0; m < bt->members.size(); ++m)
    if (abx_sim_probe(bt->members[m], step, out + 3 * m)) return -1;
  return 0;
}

extern "C" int abx_batch_tune(abx_batch* bt, int trial_steps, int* best_spb, int* best_persistent, double* best_ms) {
  if (!bt || trial_steps < 2) {
    set_error("abx_batch_tune: bad argument");
    return -1;
  }
  if (abx_batch_sync(bt)) return -1;
  trial_steps += trial_steps % 2;  // whole AA pairs
  const int graph = bt->use_graph;
  double best = 1e300;
  int bs = 4, bp = 0, bg = 4;
  for (int g : {1, 2, 4, 8, 16})
    for (int spb : {2, 4, 8})
      for (int pers : {0, 1}) {
        if (abx_batch_set_xcd_group(bt, g) || abx_batch_set_launch(bt, spb, pers, 0) || abx_batch_steps(bt, 2, 0))
          return -1;
        double ms = 0;
        if (abx_batch_time_steps(bt, trial_steps, 0, &ms)) return -1;
        if (ms < best) {
          best = ms;
          bs = spb;
          bp = pers;
          bg = g;
        }
      }
  if (abx_batch_set_xcd_group(bt, bg) || abx_batch_set_launch(bt, bs, bp, 0)) return -1;
  // the best shape without and with the even step's occupancy cap
  const int occ0 = bt->occ_even;
  int bo = occ0;
  for (int occ : {occ0 == 0 ? 16 : 0}) {
    if (abx_batch_set_occupancy(bt, occ) || abx_batch_steps(bt, 2, 0)) return -1;
    double ms = 0;
    if (abx_batch_time_steps(bt, trial_steps, 0, &ms)) return -1;
    if (ms < best) {
      best = ms;
      bo = occ;
    }
  }
  if (abx_batch_set_occupancy(bt, bo) || abx_batch_set_launch(bt, bs, bp, graph)) return -1;
  // trials advanced the members: start them again from f0 with empty averages
  for (abx_sim* s : bt->members)
    if (abx_sim_reset(s)) return -1;
  if (best_spb) *best_spb = bs;
  if (best_persistent) *best_persistent = bp;
  if (best_ms) *best_ms = best / trial_steps;
  return 0;
}

extern "C" void abx_batch_destroy(abx_batch* bt) {
  if (!bt) return;
  if (bt->stream) (void)hipStreamSynchronize(bt->stream);
  drop_graphs(bt);
  for (auto& kv : bt->lists)
    if (kv.second.first) (void)hipFree(kv.second.first);
  for (auto* p : bt->d_args)
    if (p) (void)hipFree(p);
  if (bt->ev0) (void)hipEventDestroy(bt->ev0);
  if (bt->ev1) (void)hipEventDestroy(bt->ev1);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  delete bt;
}
