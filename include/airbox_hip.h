/*
 * airbox_hip.h - C ABI of the MI355X (gfx950) AirboxCalc compute core.
 *
 * The reference compute core is src/engine.js: a browser/Node module whose
 * backend API is   prepare(p) -> g,   simulate*(g, p, onProgress, ns) -> result,
 * evaluate(g, p, acc, cells, cid, hist, Ain, Af, ns) -> result
 * (engine.js:110, 341, 613, 632; exported at engine.js:803).  This library is
 * the native layer beneath a drop-in backend beside the reference's WebGPU
 * (GpuSim, engine.js:493-594) and CPU (cpuStep, engine.js:266-321) backends.
 * Every entry point below names the reference interface it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Inputs are borrowed for the duration of
 *    the call; outputs are caller-allocated with explicit lengths.
 *  - Every function returning int returns 0 on success and -1 on failure;
 *    abx_last_error() then holds the message of the calling thread.  Messages
 *    that the reference raises are reproduced with the same text.
 *  - Grids use the reference's flat index c = (i*ny + j)*nz + k (k fastest,
 *    engine.js:128-134); per-fluid-cell arrays use its compact order
 *    ("cid order": fluid cells by ascending c, engine.js:241-242) and its
 *    structure-of-arrays layout f[q*NF + a], acc[comp*NF + a]
 *    (engine.js:274, 319).
 *  - One HIP stream per simulation handle; calls on one handle are not
 *    thread-safe, calls on different handles are.
 */
#ifndef AIRBOX_HIP_H
#define AIRBOX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct abx_dev abx_dev;
typedef struct abx_geom abx_geom;
typedef struct abx_sim abx_sim;
typedef struct abx_batch abx_batch;
typedef struct abx_surrogate abx_surrogate;

/* Message of the last failed call on this thread ("" if none). */
const char* abx_last_error(void);

/* Library version string. */
const char* abx_version(void);

/* ---- device ---------------------------------------------------------------
 * Replaces gpuDevice() (engine.js:479-491): opens GPU `ordinal`, checks that
 * it is a gfx950 part and writes its name ("AMD Instinct MI355X ...") into
 * name[name_len].  Fails (no silent fallback) when no such GPU exists. */
int abx_device_open(int ordinal, abx_dev** dev, char* name, size_t name_len);
void abx_device_close(abx_dev* dev);

/* Practical HBM ceiling (no reference counterpart): GB/s of a plain float4
 * copy of `bytes` (read + write counted), averaged over iters launches. */
int abx_copy_bandwidth(abx_dev* dev, int64_t bytes, int iters, double* gbps, double* ms_per_copy);

/* Ceiling of the stream-collide's access pattern (no reference counterpart):
 * GB/s (152 B per position) of an in-place pass over 19 planes of `positions`
 * floats that reads slot q and writes slot opp q at every position with
 * non-temporal accesses, the AA even step without its collision, averaged
 * over iters launches. */
int abx_aa_pattern_bandwidth(abx_dev* dev, int64_t positions, int iters, double* gbps, double* ms_per_pass);

/* ---- STL ------------------------------------------------------------------
 * Replaces parseSTL (engine.js:15-28).  Binary when 84 + 50*n == len, else
 * the ASCII "vertex x y z" records.  Writes at most cap floats to out (may be
 * NULL with cap 0 to query) and the number of floats (9 per triangle) to
 * *nfloat.  Error text as engine.js:26. */
int abx_parse_stl(const uint8_t* buf, int64_t len, float* out, int64_t cap, int64_t* nfloat);

/* ---- voxelizer ------------------------------------------------------------
 * Replaces voxelize (engine.js:45-93), run on the GPU: parity fill along Z of
 * the triangle soup T (9 floats per triangle, local frame) on the grid
 * nx*ny*nz of spacing h at origin.  vox_out (nx*ny*nz bytes, i-major, k
 * fastest) receives 0/1, bit-identical to engine.js. */
int abx_voxelize(abx_dev* dev, const float* T, int64_t ntri, const double origin[3], int nx, int ny, int nz,
                 double h, uint8_t* vox_out);

/* ---- geometry preparation -------------------------------------------------
 * Replaces prepare(p) (engine.js:110-234).  Parameters as the reference's
 * prepare object (ui.html:477-479): cy/sy are cos/sin of the inlet yaw
 * (engine.js:36); pass NaN to have the library compute them.  max_cells is
 * the grid-size cap (engine.js:124 uses 4.5e6); <= 0 lifts it. */
typedef struct {
  double dx;
  int sealAxis, sealSign;
  double seal;
  int inletAxis, inletSign;
  double inletYaw, cy, sy;
  int has_click;
  double click[3];
  double pleat;
  int has_outlet;
  double out_center[3], out_diameter, out_height;
  double max_cells;
} abx_prep_params;

/* Scalar part of the reference's geometry object g and its summary
 * (engine.js:226-233). */
typedef struct {
  int nx, ny, nz;
  int64_t N, nf;
  int top, jm, kw0;
  double h, origin[3], mn[3], mx[3];
  int nHoles, nOpen, nOut, Af, leak, pleatOk, has_outlet;
  double volume_l, inlet_cm2, filter_cm2, outlet_cm2, outlet_cm2_real;
  int64_t nOpenPts, nFilterPts, nOutletPts;
  int dAx, iAx, oAx, sd, si;
  double seal, yaw, cy, sy, pleat;
} abx_geom_info;

int abx_prepare(abx_dev* dev, const float* tris, int64_t nfloat, const abx_prep_params* p, abx_geom** g);
int abx_geom_info_get(const abx_geom* g, abx_geom_info* out);

/* Arrays of g: ABX_G_SOLID/FLUID/OUTLET (N bytes), ABX_G_OPENING (nx*ny
 * bytes), ABX_G_XC/DC/ZC (nx/ny/nz floats), ABX_G_OPENPTS/FILTERPTS/
 * OUTLETPTS (3*n doubles, model frame), ABX_G_FINE (the fine voxel mask,
 * 27*N bytes).  cap is in elements. */
enum {
  ABX_G_SOLID = 0, ABX_G_FLUID = 1, ABX_G_OPENING = 2, ABX_G_OUTLET = 3, ABX_G_XC = 4, ABX_G_DC = 5,
  ABX_G_ZC = 6, ABX_G_OPENPTS = 7, ABX_G_FILTERPTS = 8, ABX_G_OUTLETPTS = 9, ABX_G_FINE = 10
};
int abx_geom_read(const abx_geom* g, int what, void* out, int64_t cap);
void abx_geom_free(abx_geom* g);

/* ---- simulation -----------------------------------------------------------
 * The lattice a simulation runs on, as plain arrays.  abx_geom_grid() fills
 * it from a prepared geometry; a caller holding the reference's g object
 * fills it from g.fluid, g.opening, g.outlet, g.dc, g.top, g.jm, g.pleat. */
typedef struct {
  int nx, ny, nz;
  const uint8_t* fluid;   /* N */
  const uint8_t* opening; /* nx*ny */
  const uint8_t* outlet;  /* N */
  const float* dc;        /* ny */
  int top, jm;
  double pleat;
} abx_grid;

int abx_geom_grid(const abx_geom* g, abx_grid* out);

/* Run parameters of setupSim (engine.js:239-264, p = {K, aniso, u, tau}). */
typedef struct {
  double K, aniso, u, tau;
} abx_run_params;

/* Arithmetic of the stream-collide kernel:
 *  ABX_F32        float storage, float arithmetic (the WGSL lbm_step,
 *                 engine.js:378-450);
 *  ABX_F32_F64    float storage, double arithmetic in the reference's
 *                 operation order (cpuStep, engine.js:266-321); bit-identical
 *                 to the CPU backend;
 *  ABX_F64        double storage and arithmetic (error reference). */
enum { ABX_F32 = 0, ABX_F32_F64 = 1, ABX_F64 = 2 };

/* setupSim scalars (engine.js:246-263). */
typedef struct {
  int64_t NF;
  int Ain, Af, Lp, nM, nT;
  double u0, tau0, sig, aniso, cs2s;
  int precision, B;
  int64_t device_bytes;
  /* the device lattice (no reference counterpart): positions per population
   * plane (the grid rows that hold fluid, each stored from its first to its
   * last fluid cell, one after another), the number of such rows, and the
   * number of 64-position segments that hold fluid */
  int64_t plane;
  int nrows, nseg;
} abx_sim_info;

/* Replaces setupSim + new GpuSim(dev, S, B, ns) (engine.js:239-264,
 * 495-545): builds the dense device lattice, B variants that share the
 * geometry, all at the cold start f = WT (engine.js:260-261).  ns is NULL or
 * B*NF floats in cid order (partial bounce-back, engine.js:238). */
int abx_sim_create(abx_dev* dev, const abx_grid* grid, const abx_run_params* rp, int precision, int B,
                   const float* ns, abx_sim** sim);
int abx_sim_info_get(const abx_sim* sim, abx_sim_info* out);

/* Replaces GpuSim.steps(n, doacc) (engine.js:552-557): n fused
 * stream-collide steps on every variant; with doacc the step adds the
 * physical velocity and density to acc (engine.js:319, 445-449).
 * Asynchronous on the handle's stream. */
int abx_sim_steps(abx_sim* sim, int n, int doacc);

/* Replaces GpuSim.probe() + progressInfo (engine.js:566-572, 330-339) for
 * variant 0: out = {gamma, massRatio, qo}.  Fails with the reference's
 * instability message (engine.js:335) when the outflow is not finite. */
int abx_sim_probe(abx_sim* sim, int step, double out[3]);

/* Replaces GpuSim.read(acc) and the division by nacc (engine.js:627-628):
 * acc of variant b in cid order (4*NF), each value divided by divisor in
 * double and stored as float (as the reference's Float32Array does). */
int abx_sim_read_acc(abx_sim* sim, int b, double divisor, float* out);
/* The same in double (ABX_F64 handles keep acc in double). */
int abx_sim_read_acc_f64(abx_sim* sim, int b, double divisor, double* out);

/* Replaces GpuSim.read(curBuf()) (engine.js:580-585, 603): populations of
 * variant b in cid order (19*NF); float for ABX_F32/ABX_F32_F64, double for
 * ABX_F64 handles. */
int abx_sim_read_f(abx_sim* sim, int b, void* out);
/* Replaces resetFrom (engine.js:561-565): overwrite the populations of
 * variant b (b < 0: every variant) from 19*NF values in cid order. */
int abx_sim_write_f(abx_sim* sim, int b, const void* in);
/* Equilibrium start f_q = feq_q(rho, u) from rho_u = {rho[NF], ux[NF],
 * uy[NF], uz[NF]} in cid order, for every variant (feq as engine.js:296).
 * With rho = 1, u = 0 this is the cold start f0 = WT exactly. */
int abx_sim_init_eq(abx_sim* sim, const float* rho_u);
/* Equilibrium start from a predicted field on a coarser lattice of the same
 * box (the surrogate's warm start; the reference starts from rest,
 * engine.js:260-261).  vol (device, bf16, vd x vh x vw cells, cstride
 * channels per cell, channels 0..3 = ux, uy, uz, rho as v) holds the coarse
 * lattice in its first nxc x nyc x nzc cells; coarse_fluid (device,
 * nxc*nyc*nzc bytes, k fastest) marks its fluid cells.  Each fluid cell of
 * sim takes the trilinear average over the fluid coarse cells around its
 * centre (weights renormalised; rest if none), mapped to lattice units as
 * phys_c = affine[c] * v_c + affine[4 + c].  Cell centres are origin +
 * (i + 0.5) h on both lattices (engine.js:137-139).  The caller makes vol
 * and coarse_fluid ready before the call; the call returns when sim is set. */
int abx_sim_init_eq_resample(abx_sim* sim, const uint16_t* vol, int cstride, int vd, int vh, int vw, int nxc,
                             int nyc, int nzc, const uint8_t* coarse_fluid, const double origin_c[3], double h_c,
                             const double origin_f[3], double h_f, const float affine[8]);
/* Replaces clearAcc (engine.js:559). */
int abx_sim_clear_acc(abx_sim* sim);
/* Cold start again: every variant back to f0 = WT (engine.js:260-261) and
 * the averaged fields cleared. */
int abx_sim_reset(abx_sim* sim);
/* Replaces setNS (engine.js:558): B*NF floats in cid order, or NULL for 0. */
int abx_sim_set_ns(abx_sim* sim, const float* ns);
/* Wait for the handle's stream. */
int abx_sim_sync(abx_sim* sim);
/* Tuning knobs of the stream-collide launch (no reference counterpart).
 * kernel: 4 = AA pattern in one lattice (the default; one position per lane,
 *             four active 64-position segments per workgroup),
 *         2 = the same segments with two lattices and pull streaming.
 * xcd_map: order of the workgroups over the 8 XCDs.  G > 1 = groups of G
 * consecutive blocks per XCD, all XCDs moving through the lattice together
 * (4, the default, also taken for any negative value); 0 = blocks in lattice
 * order, round robin over the XCDs; 1 = each XCD sweeps one contiguous part
 * of the lattice.
 * nt (AA pattern, float storage): non-temporal accesses (streamed through
 * the caches without keeping the lines), flags of the even step in bits 0-7
 * and of the odd step in bits 8-15: 1 / 2 = a cell's own population loads /
 * stores (all of the even step's, the odd step's rest population), 4 / 8 =
 * the odd step's neighbour loads / stores, 16 = the averaged fields, 32 = the
 * odd step's meta words.  Compiled sets: even 0, 3, 19; odd 0, 1, 2, 3, 8,
 * 11, 16, 19, 32, 48.  Default 19 (the even step's populations and averaged
 * fields non-temporal, the odd step's accesses cached: its gathers reuse
 * neighbouring rows' lines).  Every choice computes the same values bit for
 * bit. */
int abx_sim_set_launch(abx_sim* sim, int kernel, int xcd_map, int nt);
/* Ceiling of the stream-collide's own memory pattern (no reference
 * counterpart): average ms of the even and of the odd AA step's loads and
 * stores on this handle's lattice, in its launch order, with the collision
 * left out (every population moves as it came), over iters pairs of steps.
 * Default non-temporal flags.  Needs a single float-storage handle on the AA
 * pattern; it moves the populations (a streaming without collision), so the
 * state afterwards is not a solve's: use a handle made for it, or reset. */
/* Occupancy caps of the AA steps, in waves per CU (0: no cap; default 16 for
 * the even step, none for the odd step): a launch reserves LDS it does not
 * use so that no more workgroups fit on a CU.  The even step streams 38
 * accesses per lane and measured 2.5 % faster at 16 waves per CU than at the
 * 32 its registers allow.  Results are the same bit for bit.  No reference
 * counterpart (a launch knob). */
int abx_sim_set_occupancy(abx_sim* sim, int even_waves_per_cu, int odd_waves_per_cu);

int abx_sim_pattern_time(abx_sim* sim, int iters, double* ms_even, double* ms_odd);
/* Capture each chunk of abx_sim_steps once as a HIP graph and replay it
 * (fewer launch overheads on small grids).  Results are unchanged. */
int abx_sim_set_graph(abx_sim* sim, int enable);
/* Average duration in ms of the stream-collide launches of the last
 * abx_sim_steps call with timing on (HIP events on the handle's stream). */
int abx_sim_time_steps(abx_sim* sim, int n, int doacc, double* ms_total);
/* Replaces GpuSim.destroy() (engine.js:593). */
void abx_sim_destroy(abx_sim* sim);

/* ---- batched solves ---------------------------------------------------------
 * Generalises GpuSim's B variants of one geometry (engine.js:495-545) to n
 * independent handles of different geometries, grid sizes and parameters,
 * stepped by one launch per step.  Members must share the precision, hold one
 * variant each and no ns; they are switched to the AA-pattern launch.  Each
 * member's result is bit-identical to stepping it alone.  The member handles
 * stay usable (probe, read, evaluate) after abx_batch_sync. */
int abx_batch_create(abx_dev* dev, abx_sim* const* sims, int n, abx_batch** out);
/* n steps of every member; with doacc every member accumulates its fields. */
int abx_batch_steps(abx_batch* batch, int n, int doacc);
/* out[3*m .. 3*m+2] = {gamma, massRatio, qo} of member m (abx_sim_probe). */
int abx_batch_probe(abx_batch* batch, int step, double* out);
int abx_batch_sync(abx_batch* batch);
/* Launch shape: spb = 64-position segments per workgroup (2, 4 or 8),
 * persistent grid, graph = capture each chunk of steps once as a HIP graph
 * and replay it. */
int abx_batch_set_launch(abx_batch* batch, int spb, int persistent, int graph);
/* Workgroups per XCD group (1, 2, 4, 8 or 16; default 4): consecutive
 * blocks of the work list that one XCD takes in turn, all XCDs sweeping the
 * members together (the single-handle launch's block order, abx_sim_set_launch
 * xcd_map).  abx_batch_set_launch leaves it as it is. */
int abx_batch_set_xcd_group(abx_batch* batch, int group);
int abx_batch_get_xcd_group(const abx_batch* batch);
/* Occupancy cap of the batch's even AA steps in waves per CU (0: no cap;
 * default 16, as abx_sim_set_occupancy's). */
int abx_batch_set_occupancy(abx_batch* batch, int even_waves_per_cu);
int abx_batch_get_occupancy(const abx_batch* batch);
/* Autotuner: times every launch shape (segments per workgroup x persistent
 * grid x XCD group, 30 shapes) over trial_steps steps, then the fastest with
 * the even step's occupancy cap toggled, keeps the fastest (the group and the
 * cap through abx_batch_get_xcd_group / abx_batch_get_occupancy), then resets
 * every member to the cold start (abx_sim_reset). */
int abx_batch_tune(abx_batch* batch, int trial_steps, int* best_spb, int* best_persistent, double* best_ms);
/* Time n steps with HIP events on the batch's stream. */
int abx_batch_time_steps(abx_batch* batch, int n, int doacc, double* ms);
void abx_batch_destroy(abx_batch* batch);

/* ---- evaluation -------------------------------------------------------------
 * The geometry fields evaluate/field3d/streamlines read from the reference's
 * g object (engine.js:633, 699, 722): grid, rows top and jm, pleat depth,
 * inlet plane mn[2], opening mask, cell centres and the local frame. */
typedef struct {
  int nx, ny, nz, top, jm;
  double h, pleat, mn_z;
  const uint8_t* opening; /* nx*ny */
  const float *xc, *dc, *zc;
  int oAx, dAx, iAx, sd, si;
  double seal, cy, sy;
} abx_eval_geom;

int abx_geom_eval_view(const abx_geom* g, abx_eval_geom* out);

/* Result metrics of evaluate (engine.js:682-685) and the section indices and
 * reference values it derives (engine.js:657-665, 704-710). */
typedef struct {
  double gamma, gammaCore, lowCore, highCore, reverse, maxToMean, zeta, inlet_cm2, filter_cm2;
  int ic, jc, jlow, jb, kB;
  double rhoRef, pin;
  int n, cn;
} abx_metrics;

/* Replaces evaluate + field3d (engine.js:632-719) on the GPU, reading the
 * averaged fields of variant b (acc / divisor) in place.  Cells with ns > 0.5
 * leave the fluid set as in engine.js:634.  Outputs (any may be NULL):
 * fmap nx*nz; planInlet, planLow 3*nx*nz {speed, ua, ub}; lon 3*nz*ny
 * {speed, ua, ub}; spd, pst, loss N each (reference flat order, NaN outside
 * the fluid), all three or none. */
int abx_evaluate(abx_sim* sim, const abx_eval_geom* g, int b, double divisor, abx_metrics* m, float* fmap,
                 float* planInlet, float* planLow, float* lon, float* spd, float* pst, float* loss);

/* No reference counterpart: evaluate's metrics (and, if fmap is not NULL,
 * the normalised filter map of nx*nz floats) exactly as abx_evaluate computes
 * them, without allocating device memory, for callers that cannot allocate
 * (code under HIP graph capture, serving loops that poll a running solve).
 * The working buffers live in one fixed device array of 8 MiB per GPU, held
 * until the process ends; a grid needs abx_evaluate_metrics_noalloc_bytes
 * (nx, ny, nz) bytes of it (about 5*nx*nz + nx*ny) and larger grids fail with
 * an error.  Every call on a GPU uses that GPU's one array: calls that
 * overlap on one GPU (different threads, or streams of different handles)
 * race and must be serialised by the caller. */
int64_t abx_evaluate_metrics_noalloc_bytes(int nx, int ny, int nz);
int abx_evaluate_metrics_noalloc(abx_sim* sim, const abx_eval_geom* g, int b, double divisor, abx_metrics* m,
                                 float* fmap);

/* Replaces streamlines (engine.js:721-758): RK2 traces of the averaged field
 * from up to maxSeeds seeds over the inlet opening, one GPU thread per seed.
 * pts (3 floats per point, model frame), spd (one per point) and off (line
 * starts, nlines+1 entries) as the reference's {pts, spd, off}.  The seed
 * stride gives at most maxSeeds+1 lines of at most 2002 points
 * (engine.js:738, 745-748).  *npts and *nlines are always set; the call fails
 * when a non-NULL output is shorter than they need (a NULL output with cap 0
 * only counts). */
int abx_streamlines(abx_sim* sim, const abx_eval_geom* g, int b, double divisor, int maxSeeds, float* pts,
                    int64_t pts_cap, float* spd, int64_t spd_cap, uint32_t* off, int64_t off_cap,
                    int64_t* npts, int64_t* nlines);

/* ---- time to steady ----------------------------------------------------------
 * No reference counterpart: the reference always runs a fixed number of steps
 * averaging the last `avg` (engine.js:341-358, 613-630).  Runs variant 0 of
 * sim in windows of `window` steps, each averaged on the GPU; after each
 * window the probe runs (failing with the reference's instability message,
 * engine.js:335) and evaluate's metrics of the window's average are taken
 * (engine.js:632-691).  Steady: gamma-core moved by less than tol_gamma and
 * zeta by at most tol_zeta * |zeta| between two consecutive windows; the
 * result is then the mean of those two windows' metrics (section indices and
 * counts of the last).  Stops unsteady after max_steps (windows that fit).
 * step0 is added to the step numbers reported (steps already run).  cb, if
 * not NULL, gets every window on the calling thread and stops the solve by
 * returning nonzero.  On return the handle holds the last window's average in
 * acc (divisor window) and its populations after out->steps more steps. */
typedef struct {
  int window, max_steps, step0;
  double tol_gamma, tol_zeta;
} abx_steady_params;
typedef struct {
  int step;                 /* step0 + steps run at the end of the window */
  double gamma, massRatio;  /* the probe (progressInfo, engine.js:330-339) */
  abx_metrics metrics;      /* evaluate of the window's average */
} abx_steady_window;
typedef struct {
  int converged, steps, windows;
  abx_metrics metrics;
} abx_steady_result;
typedef int (*abx_steady_cb)(void* user, const abx_steady_window* w);
int abx_sim_solve_to_steady(abx_sim* sim, const abx_eval_geom* g, const abx_steady_params* p, abx_steady_cb cb,
                            void* user, abx_steady_result* out);

/* ---- Flow surrogate (3-D U-Net) kernels ---------------------------------
 * The reference has no surrogate: every solve starts from rest, f = w_q
 * (engine.js:260-261).  These kernels train and run a U-Net whose predicted
 * (rho, u) seeds abx_sim_init_eq.  All launch on `stream` (a hipStream_t,
 * NULL = the default stream) and return at once; errors are argument errors
 * found before the launch.
 *
 * Tensors are bf16 (uint16_t bit patterns) in NDHWC order: voxel (n, z, y, x)
 * of an n x d x h x w grid at ((n*d + z)*h + y)*w + x, channels contiguous,
 * `*s` voxel strides in elements.  Convolution channel counts are multiples of
 * 16; strides and channel offsets multiples of 8; pointers 16-byte aligned.
 * Convolutions are 'same' (zero padding), ksize 3 (27 taps) or 1. */

/* FP32 weights [co][taps][ci] -> bf16 for abx_conv3d_forward; with
 * for_input_grad the flipped transpose [ci][taps][co] (tap t <- taps-1-t),
 * which makes abx_conv3d_forward over dZ compute the input gradient. */
int abx_conv3d_pack(void* stream, const float* w, int co, int ksize, int ci, int for_input_grad, uint16_t* out);

/* abx_conv3d_pack of several layers in one launch, both layouts of each: fwd
 * as for_input_grad = 0, bwd as for_input_grad = 1 (taps 1 or 27).  At most
 * ABX_PACK_MAX layers.  No reference counterpart. */
#define ABX_PACK_MAX 32
typedef struct {
  const float* w;
  uint16_t* fwd;
  uint16_t* bwd;
  int co, taps, ci;
} abx_pack_desc;
int abx_conv3d_pack_all(void* stream, int n, const abx_pack_desc* desc);

/* y = act(conv(x, w) + bias) [+ y if accumulate]; act = ReLU if relu.  bias
 * may be NULL.  MFMA implicit GEMM, FP32 accumulation, bf16 result. */
int abx_conv3d_forward(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int ci,
                       const uint16_t* wpack, const float* bias, int co, int ksize, int relu, int accumulate,
                       uint16_t* y, int ys);

/* abx_conv3d_forward, then every output value is set to zero where gate (a
 * tensor of co channels at per-voxel stride gs, the output of a ReLU layer)
 * is not positive: the input gradient of a convolution with the backward of
 * the ReLU before it folded into the store, equal bit for bit to
 * abx_conv3d_forward followed by abx_relu_backward(y, ys, gate, gs, ...).
 * No reference counterpart. */
int abx_conv3d_forward_gated(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int ci,
                             const uint16_t* wpack, const float* bias, int co, int ksize, int relu, int accumulate,
                             uint16_t* y, int ys, const uint16_t* gate, int gs);

/* abx_conv3d_forward (3x3x3, no gate) over a concatenation held as two
 * tensors: input channels [0, ci1) are channels of x (stride xs), [ci1, ci)
 * channels 0.. of x2 (stride xs2); output channels [0, co1) go to y (stride
 * ys), [co1, co) to channels 0.. of y2 (stride ys2).  ci1 and co1 are
 * multiples of 16; with ci1 = ci (co1 = co) x2 (y2) is unused.  Equal bit for
 * bit to abx_conv3d_forward on the concatenated tensors.  No reference
 * counterpart (the U-Net's skip connections). */
int abx_conv3d_forward_parts(void* stream, const uint16_t* x, int xs, int ci1, const uint16_t* x2, int xs2, int n,
                             int d, int h, int w, int ci, const uint16_t* wpack, const float* bias, int co, int relu,
                             int accumulate, uint16_t* y, int ys, int co1, uint16_t* y2, int ys2);

/* Process-wide choice of convolution kernels, for comparison and timing
 * (no reference counterpart).  flags bit 0: 3x3x3 forward convolutions use the
 * 16-channel-tile kernel (16x4x8 voxel tiles, weights in LDS) instead of the
 * general one; bit 1: the tap-group weight gradients build their MFMA
 * fragments with gfx950's transposing LDS reads (ds_read_b64_tr_b16) instead
 * of gathering 16-bit values; bit 2: 3x3x3 weight gradients use those
 * tap-group kernels instead of the every-tap kernel (two-part inputs always
 * take the every-tap kernel); bit 3: the 16-channel forward kernel stages its
 * input brick and weights through registers instead of loading them straight
 * into LDS; bit 6: it reads the weights from LDS for every MFMA (three
 * workgroups per CU; bit 4: two) instead of holding them in registers; bit
 * 5: the every-tap weight gradient stages its
 * tiles through registers everywhere (else it loads them straight into LDS
 * where its channel blocking then fits more workgroups per CU, which may
 * group its partial sums in more rows).  Bits 3, 4 and 6 change timing only,
 * not a bit of the output.  Default 3.  The 16-channel and general forward
 * kernels, and the two tap-group weight gradients, compute the same sums; the
 * every-tap weight gradient sums its partials in another grouping. */
int abx_conv3d_set_kernel(int flags);

/* Bytes of FP32 workspace abx_conv3d_wgrad needs for this shape (-1 on a bad
 * shape). */
int64_t abx_conv3d_wgrad_workspace(int n, int d, int h, int w, int ci, int co, int ksize);

/* dw[co][taps][ci] (FP32) = sum over voxels of dz[v][co] * x[v + tap][ci];
 * db[co] = sum dz[v][co] (db may be NULL); both added to the old contents if
 * accumulate.  Deterministic: per-workgroup partial sums in `workspace`,
 * reduced in a fixed order. */
int abx_conv3d_wgrad(void* stream, const uint16_t* x, int xs, const uint16_t* dz, int dzs, int n, int d, int h,
                     int w, int ci, int co, int ksize, float* workspace, int64_t ws_bytes, float* dw, float* db,
                     int accumulate);

/* abx_conv3d_wgrad (3x3x3) over an input held as two tensors, as in
 * abx_conv3d_forward_parts (channels [0, ci1) of x, [ci1, ci) of x2); the
 * workspace of abx_conv3d_wgrad_workspace for the whole ci suffices.  No
 * reference counterpart. */
int abx_conv3d_wgrad_parts(void* stream, const uint16_t* x, int xs, int ci1, const uint16_t* x2, int xs2,
                           const uint16_t* dz, int dzs, int n, int d, int h, int w, int ci, int co, float* workspace,
                           int64_t ws_bytes, float* dw, float* db, int accumulate);

/* 2x2x2 max pooling of an n x d x h x w grid (even sizes), c channels
 * (multiple of 8), and its gradient (to the first maximum in z, y, x order;
 * dx overwritten, or added to if accumulate). */
int abx_maxpool2(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int c, uint16_t* y, int ys);
int abx_maxpool2_backward(void* stream, const uint16_t* x, int xs, const uint16_t* dy, int dys, int n, int d, int h,
                          int w, int c, uint16_t* dx, int dxs, int accumulate);
/* abx_maxpool2_backward where x is the output of a ReLU: dx is also zeroed
 * wherever x is not positive (the ReLU's backward, as abx_relu_backward(dx,
 * dxs, x, xs, ...) after it, bit for bit).  No reference counterpart. */
int abx_maxpool2_backward_gated(void* stream, const uint16_t* x, int xs, const uint16_t* dy, int dys, int n, int d,
                                int h, int w, int c, uint16_t* dx, int dxs, int accumulate);

/* Nearest-neighbour 2x upsampling of an n x d x h x w grid and its gradient
 * (sum of the 8 fine values, FP32). */
int abx_upsample2(void* stream, const uint16_t* x, int xs, int n, int d, int h, int w, int c, uint16_t* y, int ys);
int abx_upsample2_backward(void* stream, const uint16_t* dy, int dys, int n, int d, int h, int w, int c, uint16_t* dx,
                           int dxs);
/* abx_upsample2_backward, then dx = 0 wherever gate (the coarse output of a
 * ReLU, c channels at stride gs) is not positive, bit for bit as
 * abx_relu_backward after it.  No reference counterpart. */
int abx_upsample2_backward_gated(void* stream, const uint16_t* dy, int dys, int n, int d, int h, int w, int c,
                                 uint16_t* dx, int dxs, const uint16_t* gate, int gs);

/* dy *= (y > 0) in place over nvox voxels of c channels. */
int abx_relu_backward(void* stream, uint16_t* dy, int dys, const uint16_t* y, int ys, int64_t nvox, int c);

/* Training's head, loss and head backward in one pass (no reference
 * counterpart: the surrogate network's output layer).  x: the 16-channel
 * ReLU output that feeds the head, a 1x1x1 convolution (wpack: its forward
 * pack, bf16 [co >= c][16]; bias may be NULL) of which the first c (1-4)
 * outputs enter the loss.  Per voxel: out_j = bf16(sum_k W[j][k] x_k + b_j),
 * the loss and its gradient d_j = bf16(dloss / dout_j) as abx_mse_loss
 * computes them from out, and dx_k = bf16(sum_j W[j][k] d_j), zero where
 * x_k <= 0 (the ReLU's backward); dw[j][k] = sum over voxels of d_j x_k and
 * db[j] = sum of d_j for j < 4 (rows from c on get zeros; rows 4.. of a wider
 * head are not written).  loss_parts: abx_head_loss_partials(nvox) floats
 * whose sum is the loss.  Deterministic: per-workgroup partial sums in the
 * workspace (abx_head_loss_workspace(nvox) bytes), reduced in a fixed
 * order.  Numerically the head convolution, abx_mse_loss and the head's
 * gated input and weight gradients, up to the order of the FP32 sums. */
int64_t abx_head_loss_partials(int64_t nvox);
int64_t abx_head_loss_workspace(int64_t nvox);
int abx_head_loss_backward(void* stream, const uint16_t* x, int xs, int ci, const uint16_t* wpack, const float* bias,
                           int c, const uint16_t* target, int ts, const uint8_t* mask, int64_t nvox,
                           const float* chan_w, float scale, float* loss_parts, uint16_t* dx, int dxs,
                           float* workspace, int64_t ws_bytes, float* dw, float* db);

/* Number of per-workgroup loss partial sums abx_mse_loss writes. */
int64_t abx_mse_partials(int64_t nvox);

/* Masked, channel-weighted squared error over the first c channels:
 * loss = scale * sum_v mask[v] sum_c chan_w[c] (pred - target)^2, written as
 * abx_mse_partials(nvox) partial sums; dpred (all dps channels) = its
 * gradient, zero outside the mask and for channels >= c. */
int abx_mse_loss(void* stream, const uint16_t* pred, int ps, const uint16_t* target, int ts, const uint8_t* mask,
                 int64_t nvox, int c, const float* chan_w, float scale, float* loss_parts, uint16_t* dpred, int dps);

/* Input volume of the surrogate for one box (no reference counterpart):
 * the lattice nx x ny x nz of a prepared box (device copies of abx_grid's
 * fluid, opening, outlet and dc; top and pleat) placed at the origin of a
 * d x h x w volume (nx <= d, ny <= h, nz <= w, each <= 1024), voxels beyond
 * it solid.  Writes c >= 7 bf16 channels per voxel (stride xs): solid,
 * porous, inlet, K/6 on porous cells, signed Euclidean wall distance in cells
 * / 8 clipped to [-4, 4] (exact, positive in the fluid), outlet, aniso/20 on
 * porous cells, then zeros.  Cell types follow setupSim's precedence
 * (engine.js:250-255).  workspace: abx_surrogate_features_workspace(d, h, w)
 * bytes of device memory. */
int64_t abx_surrogate_features_workspace(int d, int h, int w);
int abx_surrogate_features(void* stream, const uint8_t* fluid, const uint8_t* opening, const uint8_t* outlet,
                           const float* dc, int nx, int ny, int nz, int top, double pleat, double K, double aniso,
                           int d, int h, int w, int32_t* workspace, int64_t ws_bytes, uint16_t* x, int xs, int c);

/* The trained network as a handle (no reference counterpart): reads an
 * inference weight file (.abxw: "ABXUNET1", shape, cell size of the
 * network's volume, target standardisation, FP32 weights; written by
 * airboxcalc_amd/surrogate.py write_weights), packs the weights to bf16 and
 * allocates one pass's activations on dev.  abx_surrogate_info gives the
 * volume d x h x w and the cell size (mm) the network reads boxes at. */
int abx_surrogate_load(abx_dev* dev, const char* path, abx_surrogate** out);
int abx_surrogate_info(const abx_surrogate* net, int dims[3], double* family_dx);
void abx_surrogate_free(abx_surrogate* net);
/* Warm start (the reference starts from rest, engine.js:260-261): prepares
 * the box at the network's cell size (prepare as abx_prepare from tris and p
 * with dx replaced, the cap lifted; skipped when g_fine already has that cell
 * size, and then tris and p may be NULL), builds the input volume
 * (abx_surrogate_features, with K and aniso of the run), runs the network on
 * the library's kernels and seeds every variant of sim, whose lattice is
 * g_fine's, at the equilibrium of the prediction (abx_sim_init_eq_resample).
 * Returns when sim is set.  Calls on one network handle must not overlap. */
int abx_surrogate_warm_start(abx_surrogate* net, abx_sim* sim, const abx_geom* g_fine, const float* tris,
                             int64_t nfloat, const abx_prep_params* p, double K, double aniso);

/* One Adam step (bias-corrected, step >= 1) on n FP32 weights. */
int abx_adam(void* stream, float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
             float eps, int step);

#ifdef __cplusplus
}
#endif
#endif
