// amcafe_dev.h -- internal interface between the g++-compiled host code and the
// hipcc-compiled device code of libamcafe.  Plain C structs and functions only.
//
// Device layout of one CA instance (a "domain"): structure-of-arrays over the
// nX*nY*nZ voxels in the reference's linear order (x fastest, S/Partition.C:281-288):
//   vs    uint8   voxel state (0 empty, 1 liquid, 2 mushy, 3 solid) in bits 0-1,
//                 bit 6 = member of the active set this step, bit 7 = T >= tL
//   gid   int32   grain id (1-based, 0 = none)
//   ext   double  octahedron half-diagonal ("extents")
//   cx/cy/cz      octahedron centre ("centroidOct"), one array per component
//   T     double  temperature of this step
//   bj    uint8   sticky melt flag (TempField::BuildjID)
//   amap  int32   voxel -> index in the active list, -1 otherwise
// The active list (interface cells of this step, ascending voxel id) and the
// growing list (active cells that can capture) hold their own SoA arrays.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMCD_NB 26  // Moore neighbours

// constant per-domain parameters (from the deck, computed once by the host)
typedef struct {
  int nx, ny, nz;
  double dx, dy, dz;
  double hx, hy;     // LX[0]/2., LX[1]/2.
  double tL, tS, T0, xi;  // xi = tL*(1+DBL_EPSILON)
  double a1[6];      // melt-pool semi-axes, TempField::a1 (S/TempField.C:71-76)
  double bmDX0;      // DelT*bmV
  double DelT;
  double rX;         // rNmax*pow(dx*1e6,3.) from host libm (S/VoxelsCA.C:344)
  double vhat_tS, vhat_T0;  // Avel*pow(tL-T,nvel) at the two sub-liquidus analytic temperatures
  double Avel, nvel;
  unsigned int seed1;  // VoxelsCA::seed1
  double bmV;          // beam speed (EASM source path)
} amcd_params;

// per-step inputs (host computes them from the Grid / TempField state)
typedef struct {
  double lx, ly, lz;   // laser position (S/TempField.C:152-154)
  double dirp;         // scan direction (S/TempField.C:167-171)
  double cg, sg;       // cos(gth), sin(gth) from host libm
  int ilaserLoc;
  int tInd;            // Grid::tInd, seeds the step's random numbers
  int n_pool;          // nucleation orientations available in the pool
  int nGrain;          // grains before the step (VoxelsCA::nGrain)
  int skip;            // 1: this instance does not advance (finished run)
  int pad;
} amcd_step;

// per-step results written by the device
typedef struct {
  int status;          // 0 ok, 1 needs more nucleation orientations, <0 error
  int any_melt;        // some voxel at T >= tL
  int n_active, n_mushy0, n_events, n_captures, n_nucleations, n_slots;
  int nGrain;          // after the step
  int pool_used;
  long long draws;     // engine draws consumed (2 per uniform)
  double tinc;
  int overflow;        // active list exceeded capacity
  int err_state;        // a growing cell carried no valid grain id (the domain state is corrupt)
  long long slot_visits;  // growing cells with live candidates, summed over the loop's scans
  long long cand_evals;   // capture times evaluated, summed over the loop's scans
  long long interface_events;  // growing-list length summed over the loop's events
  long long loop_cycles;       // shader-clock cycles of the capture loop, summed over its launches
} amcd_out;

typedef struct amcd_batch amcd_batch;

// lifecycle
int amcd_create(const amcd_params* p, int n, amcd_batch** out, char* err, int errlen);
void amcd_destroy(amcd_batch* b);
int amcd_device_name(char* buf, int len);

// state transfer (reference layouts: int32 / double, centroid interleaved xyz)
int amcd_upload(amcd_batch* b, int i, const int* gid, const int* vstate, const double* ext, const double* cen,
                const int* bj);
int amcd_download(amcd_batch* b, int i, int* gid, int* vstate, double* ext, double* cen, int* bj, double* T);
// grain table: 9 doubles (row-major 3x3, local -> global) per grain, grains 1..n
int amcd_set_rot(amcd_batch* b, int i, const double* rot9, int ngrain);
int amcd_get_rot(amcd_batch* b, int i, double* rot9, int ngrain);
int amcd_set_pool(amcd_batch* b, int i, const double* rot9, int n);
// VoxelsCA::SetBuild on the device (grain ids 1..ngrain); Build = 1 on the first n_built voxels;
// Build as ints
int amcd_set_build(amcd_batch* b, int i, int ngrain);
int amcd_build_init(amcd_batch* b, int i, long long n_built);
int amcd_get_build(amcd_batch* b, int i, int* out);
// every instance's pool in one staged copy (rot9[i] NULL or n[i] = 0: instance i unchanged)
int amcd_set_pools(amcd_batch* b, const double* const* rot9, const int* n);
// graph launches of the step so far (all batches of the process)
long long amcd_graph_launches(void);
// the chunking the autotuner picked (0: not yet), how often it tuned, the growing-list size of
// the last tuning and each candidate's time (ms) there; returns 1 once tuned
int amcd_tuning(amcd_batch* b, int* list_cap, int* n_tunings, int* cells, float* ms, int nms);

// the time step: all instances of the batch advance together.
// stage 1: temperature only (AnalyticTempCurr seam)
int amcd_temperature(amcd_batch* b, const amcd_step* steps);
// stage 2: UpdateVoxels (SetLiquid3 .. ConvertSolid1(0)); requires stage 1 of the same step
int amcd_update_voxels(amcd_batch* b, const amcd_step* steps, amcd_out* outs);
// copy the per-step records to the host (synchronises the stream)
int amcd_read_outs(amcd_batch* b, amcd_out* outs);
// resume the event loop of instances that reported status 1 after a pool refill
int amcd_resume(amcd_batch* b, const amcd_step* steps, amcd_out* outs);

// per-layer support (S/VoxelsCA.C:800-839, 887-982)
int amcd_grain_presence(amcd_batch* b, int i, int ngrain, int* present);
int amcd_remap_grains(amcd_batch* b, int i, const int* remap, int ngrain);
// base plate (S/BasePlate.C:41-97 + S/VoxelsCA.C:125-161): Voronoi of the nzh bottom planes
// over ng sites (xyz interleaved), gID = site + 1, vState 3 (2 in the top plate plane),
// centroids at voxel centres; the rest of the domain is left as created (empty)
int amcd_init_plate(amcd_batch* b, int i, int nzh, const double* xs, int ng);
// AddLayer1's Voronoi, renumbering and voxel writes (S/VoxelsCA.C:920-981): the layer
// iz1..ilaserLoc-1 takes grains gbase + 1..; *n_used = sites that label a voxel
int amcd_add_layer(amcd_batch* b, int i, int iz1, int ilaserLoc, int gbase, const double* xs, int ng, int* n_used);

// differentiable temperature field (heat_adjoint.hip) for nb instances on one grid and track
// (theta, target, grad: 4 per instance; loss: one per instance): mode 0 loss (FP32), 1 loss
// and gradient (FP32), 2 loss (FP64); amcafe_heat_desc is declared in include/amcafe.h
struct amcafe_heat_desc_s;
unsigned long long amcd_heat_last_evals(void);
// the FP32 batch loss and gradient with every device buffer in a static device array
// (no allocation): 0 ok, -1 device error, -2 the call needs more than that array holds
int amcd_heat_noalloc(const struct amcafe_heat_desc_s* d, int nb, const double* theta, const double* target,
                      int step0, int nsteps, double* loss, double* grad, char* err, int errlen);
size_t amcd_heat_noalloc_bytes(const struct amcafe_heat_desc_s* d, int nb);
size_t amcd_heat_noalloc_max(void);
int amcd_heat(const struct amcafe_heat_desc_s* d, int nb, const double* theta, const double* target, int step0,
              int nsteps, int mode, double* loss, double* grad, float* ms, char* err, int errlen);
// EASM temperature (S/TempField.C:79-137) instead of the analytic one: beam power Q,
// layer thickness; outs[i].any_melt = 1 when some voxel is below tL; *n_evals (may be
// NULL) = source-integral evaluations summed over the batch
int amcd_easm(amcd_batch* b, const amcd_step* steps, double Q, double layerT, amcd_out* outs,
              unsigned long long* n_evals);
// temperature + UpdateVoxels in one pass (SetLiquid3 fused into the temperature kernel)
int amcd_step_fused(amcd_batch* b, const amcd_step* steps, amcd_out* outs);
// grain-table capacity (grains + nucleation pool) for the coming step
int amcd_reserve_grains(amcd_batch* b, int i, int need);

// per-kernel timing of the last step (ms): temperature, solid->mushy, flags+scan+compact,
// active setup, event loop, scatter+mushy->solid; recorded only when enabled
void amcd_set_timing(amcd_batch* b, int on);
int amcd_last_timings(amcd_batch* b, double* ms, int n);
const char* amcd_last_error(void);
// capture-loop phase timers of a -DAMCD_PHASE_PROFILE build (zeros otherwise); resets them
int amcd_phase_profile(unsigned long long* out32);

// known-answer helpers: libstdc++ uniform_real_distribution<double> over
// minstd_rand0(seed), sequentially and through the device's jump-ahead path
int amcd_uniform_stream(unsigned long long seed, int n, double* out);
int amcd_uniform_jump(unsigned long long seed, int n, double* out);
// host build of the device's shape factor pow(cos th sin ph, .5), n vectors xyz
int amcd_host_shape_factor(const double* dnx, int n, double* out);
double amcd_host_pow(double x, double y);
int amcd_host_pow_n(const double* x, const double* y, int n, double* out);

#ifdef __cplusplus
}
#endif
