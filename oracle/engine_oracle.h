/*
 * ORACLE - test infrastructure only.
 *
 * A plain-C restatement of the reference compute core (src/engine.js) used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (airboxcalc_amd/, include/) includes, links
 * or calls this code.
 *
 * Arithmetic follows JavaScript: every scalar is an IEEE double, typed-array
 * stores round to float32, no fused multiply-add (built with
 * -ffp-contract=off), evaluation order as written in engine.js.
 *
 * The only deliberate difference: the reference derives cos/sin of the inlet
 * yaw with V8's Math.cos/Math.sin (engine.js:36); the caller passes them in
 * (cy, sy) so a fixture recorded under Node reproduces bit for bit.
 */
#ifndef ABX_ENGINE_ORACLE_H
#define ABX_ENGINE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double dx;
  int sealAxis, sealSign;
  double seal;
  int inletAxis, inletSign;
  double yaw_deg, cy, sy;
  int has_click;
  double click[3];
  double pleat;
  int has_outlet;
  double out_center[3], out_diameter, out_height;
  double max_cells; /* 4.5e6 in engine.js:124; <= 0 lifts the cap */
} orc_prep_params;

typedef struct {
  int nx, ny, nz, nxf, nyf, nzf;
  int64_t N;
  double h, origin[3], mn[3], mx[3];
  int top, jm, kw0, nHoles, nOpen, nOut, Af, leak;
  int64_t nf;
  uint8_t *fine, *solid, *fluid, *opening, *outlet;
  float *xc, *dc, *zc, *T;
  double *openPts, *filterPts, *outletPts;
  int64_t nOpenPts, nFilterPts, nOutletPts; /* number of points (x3 doubles) */
  int oAx, dAx, iAx, sd, si;
  double seal, cy, sy, yaw_deg, pleat;
  int has_outlet;
} orc_geom;

typedef struct {
  int64_t NF;
  int32_t *cid, *cells, *nb, *jsOf;
  uint32_t *flag;
  int32_t *layerM, *layerT;
  int64_t nM, nT;
  int Ain, Af, Lp;
  double u0, tau0, sig, aniso, cs2s;
} orc_sim;

typedef struct {
  double gamma, gammaCore, lowCore, highCore, reverse, maxToMean, zeta, inlet_cm2, filter_cm2;
  int ic, jc, jlow, jb, kB;
  double rhoRef, pin_field;
} orc_metrics;

const char* orc_last_error(void);

int orc_parse_stl(const uint8_t* buf, int64_t len, float** out, int64_t* nfloat);
void orc_free(void* p);

int orc_voxelize(const float* T, int64_t ntri, const double origin[3], int nx, int ny, int nz, double h,
                 uint8_t* vox);
/* column hit counts of the count pass (watertightness check) */
int orc_voxel_counts(const float* T, int64_t ntri, const double origin[3], int nx, int ny, double h,
                     int32_t* cnt);

int orc_prepare(const float* tris, int64_t nfloat, const orc_prep_params* p, orc_geom* g);
void orc_geom_free(orc_geom* g);

int orc_setup_sim(const orc_geom* g, double K, double aniso, double u, double tau, orc_sim* S);
void orc_sim_free(orc_sim* S);

/* one cpuStep (engine.js:266-321); f32 storage, f64 arithmetic. acc/ns may be NULL */
void orc_cpu_step(const orc_sim* S, const float* f, float* f2, float* acc, const float* ns);
/* the same step with f64 storage (FP64 reference for the double kernel) */
void orc_cpu_step_f64(const orc_sim* S, const double* f, double* f2, double* acc, const float* ns);
/* orc_cpu_step on nthreads OpenMP threads (contiguous cell ranges); same result. */
void orc_cpu_step_threads(const orc_sim* S, const float* f, float* f2, float* acc, const float* ns, int nthreads);
/* probeCPU + progressInfo reduction (engine.js:324-339): out = {gamma, massRatio, qo} */
int orc_probe(const orc_sim* S, const float* f, int step, double out[3]);

/* simulate (engine.js:341-358) up to the averaged acc; hist rows (step,gamma,mr) */
int orc_simulate(const orc_sim* S, int steps, int avg, int every, const float* ns, float* acc_out,
                 double* hist, int* nhist, int64_t step_limit);

/* evaluate (engine.js:632-691) scalar part + maps */
int orc_evaluate(const orc_geom* g, const orc_sim* S, double u0, const float* acc, const float* ns,
                 orc_metrics* m, float* fmap, float* planInlet /*3*nx*nz*/, float* planLow /*3*nx*nz*/,
                 float* lon /*3*nz*ny*/, uint8_t* fluid_eff /*N*/);
/* field3d (engine.js:698-719) */
int orc_field3d(const orc_geom* g, const orc_sim* S, double u0, const float* acc, const uint8_t* fluid,
                double rhoRef, float* spd, float* pst, float* loss, int* kB_out, double* pin_out);
/* streamlines (engine.js:721-758); returns the number of lines.  pts capacity
   in floats, spd capacity in points, off capacity in lines+1 */
int64_t orc_streamlines(const orc_geom* g, const orc_sim* S, double u0, const float* acc, const uint8_t* fluid,
                        int maxSeeds, float* pts, int64_t pts_cap, float* spd, int64_t spd_cap, uint32_t* off,
                        int64_t off_cap, int64_t* npts_out);

double orc_js_hypot3(double a, double b, double c);
double orc_js_round(double x);

#ifdef __cplusplus
}
#endif
#endif
