// Host stages of geometry preparation (the reference's prepare(p),
// engine.js:110-234).  abx_prepare runs them around its GPU stages:
//
//   prep_grid         local frame, bounding box, grid      (engine.js:111-125)
//   [GPU]             voxelize, 27-subcell vote, slice kf  (engine.js:126-144)
//   prep_openings     openings of the first material layer (engine.js:141-171)
//   prep_flood_inputs duct depth kw0, seal row top, floor rows
//   [GPU]             duct, floor, flood fill, leak check  (engine.js:173-195)
//   prep_finish       lid plenum and outlet, filter plane, points, summary
//                     (engine.js:196-234)
//
// Plain C++ without the HIP headers: the stages build with a host compiler on
// their own (the sanitizer build of the tests runs them between CPU stand-ins
// of the GPU stages).  Every stage returns 0, or -1 with abx_last_error set to
// the reference's message.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/airbox_hip.h"

namespace abx {

// Model frame <-> local frame (X along the filter, D depth below the seal
// plane, Z along the inlet direction), makeFrame (engine.js:31-43).
struct Frame {
  int dAx, iAx, oAx, sd, si;
  double seal, cy, sy, yaw;
  void to_local(const double q[3], double L[3]) const {
    const double a = q[oAx], b = -si * q[iAx];
    L[0] = cy * a + sy * b;
    L[1] = -sd * (q[dAx] - seal);
    L[2] = -sy * a + cy * b;
  }
  void to_model(double X, double D, double Z, double q[3]) const {
    q[oAx] = cy * X - sy * Z;
    q[dAx] = seal - sd * D;
    q[iAx] = -si * (sy * X + cy * Z);
  }
};

// Host side of a geometry: the reference's g object without the fine grid.
struct GeomHost {
  abx_geom_info info;
  Frame fr;
  std::vector<uint8_t> solid, fluid, opening, outlet;
  std::vector<float> xc, dc, zc;
  std::vector<double> openPts, filterPts, outletPts;
};

// Number#toString for the cell sizes the leak message quotes.
std::string js_num(double v);
// Number#toFixed(1) for positive values.
std::string js_fixed1(double x);
// First index with a[i] >= v; -1 (or the last index when or_last) if none.
int first_ge(const std::vector<float>& a, double v, bool or_last);

// Frame, triangles in the local frame (T: nfloat floats), bounding box and
// grid of g.info; checks the axes, the grid and p->max_cells.
int prep_grid(const float* T0, int64_t nfloat, const abx_prep_params* p, GeomHost& g, std::vector<float>& T);
// Index of the fine-grid slice whose empty cells make the openings.
int prep_slice_k(const GeomHost& g);
// Cell-centre coordinates, then the opening mask from the slice: empty holds
// nxf*nyf bytes, 1 where the fine voxel (i, j, kf) is not material.
int prep_openings(const abx_prep_params* p, const uint8_t* empty, GeomHost& g);
// The flood's inputs: info.kw0, info.top (error if the seal lies outside the
// model) and the floor rows (ny bytes, 1 where a row lies beyond the floor).
int prep_flood_inputs(GeomHost& g, std::vector<uint8_t>& floor_row);
// After the flood: g.solid and g.fluid hold its result, nfluid its cell count
// and leaked whether air reached an outer face.
int prep_finish(const abx_prep_params* p, int64_t nfluid, bool leaked, GeomHost& g);

}  // namespace abx
