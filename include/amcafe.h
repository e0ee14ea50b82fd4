/* amcafe.h -- C ABI of the MI355X-native CAFE time step.
 *
 * Two levels, both plain C (pointers, sizes, int status codes; no C++ types):
 *
 *  1. Seam level: a device-resident CA domain that replaces the per-step member
 *     functions the reference's driver calls,
 *       TempField::AnalyticTempCurr   src/src-mpi/TempField.h:22, TempField.C:139-217
 *       VoxelsCA::UpdateVoxels        src/src-mpi/VoxelsCA.h:35,  VoxelsCA.C:232-655
 *     and, with the whole domain in one GPU's HBM, the halo exchange they rely on,
 *       Partition::PassInformation    src/src-mpi/Partition.h:13-14, Partition.C:414-449
 *     State is uploaded / downloaded in the reference's layouts: int32 gID and
 *     vState, double extents, double centroidOct (xyz interleaved), int32 BuildjID,
 *     all indexed by the linear voxel id x + nX*y + nX*nY*z (1 MPI rank).
 *
 *  2. Driver level: the reference's main loop (src/src-mpi/main1.C:22-98) for one
 *     or many input decks at once -- deck parsing (Grid.C:165-318), base plate
 *     (BasePlate.C:41-97), powder layers (VoxelsCA.C:887-982), CleanLayer
 *     (VoxelsCA.C:800-839), SetBuild (VoxelsCA.C:76-115) and the HDF5 output
 *     layout of WriteToHDF1 (VoxelsCA.C:1592-1856).  Several decks advance as one
 *     batch: one launch per stage for all of them.
 *
 * Every function returns AMCAFE_OK (0) or a negative code; amcafe_last_error()
 * gives a message for the calling thread.  No C++ exception crosses this ABI.
 * A handle must be used by one host thread at a time.
 */
#ifndef AMCAFE_H
#define AMCAFE_H

#ifdef __cplusplus
extern "C" {
#endif

#define AMCAFE_OK 0
#define AMCAFE_ERR_ARG (-1)     /* bad argument or call order */
#define AMCAFE_ERR_DEVICE (-2)  /* HIP runtime failure or no device */
#define AMCAFE_ERR_DECK (-3)    /* deck could not be parsed / is incomplete */
#define AMCAFE_ERR_STATE (-4)   /* state the path cannot handle (see message); a domain or
                                   simulation whose step failed holds a partial step: upload its
                                   state again (seam level) or recreate it before stepping on */
#define AMCAFE_ERR_IO (-5)      /* output file could not be written */

const char* amcafe_last_error(void);
/* device description, e.g. "AMD Instinct MI355X (gfx950, 256 CUs)" */
int amcafe_device_name(char* buf, int len);

/* ---------------------------------------------------------------------------
 * 1. Seam level
 */
typedef struct amcafe_domain amcafe_domain;

/* Grid / TempField / VoxelsCA members the time step reads (Grid.h:42-70). */
typedef struct {
  int nX[3];
  double dX;           /* isotropic spacing, deck keyword dx (Grid.C:200-206) */
  double tL, tS, T0;   /* liquidus, solidus, ambient (T0 = 300 in the reference) */
  double meltparam[4]; /* deck keyword meltparam a b c d (TempField.C:107-114) */
  double bmV, bmDelT;  /* beam speed, time step */
  double rNmax;        /* deck keyword r */
  double Avel, nvel;   /* deck keyword dendritegrowth2 */
} amcafe_domain_desc;

/* Laser state for one temperature evaluation (TempField.C:146-171, Grid.h:42-70). */
typedef struct {
  double x, y, z;      /* lcoor[2*ispvec[isp]]-offset[0], ...+1]-offset[1], ilaserLoc*dX-offset[2] */
  double dirp;         /* scan direction as in TempField.C:167-171 */
  double gth;          /* grid rotation angle (Grid::gth) */
  int ilaserLoc;       /* Grid::ilaserLoc */
} amcafe_laser;

/* Per-step counters of UpdateVoxels. */
typedef struct {
  int n_active;        /* interface cells (the reference's Na) */
  int n_mushy0;        /* mushy cells at the start of the capture loop */
  int n_events;        /* capture-loop iterations that captured or nucleated */
  int n_captures, n_nucleations;
  int n_growing;       /* cells that could capture at some point of the step */
  int nGrain;          /* grains after the step */
  long long draws;     /* minstd_rand0 outputs consumed */
  double tinc;         /* time advanced by the capture loop */
  long long slot_visits; /* growing cells with live candidates, summed over the loop's scans */
  long long cand_evals;  /* capture times evaluated, summed over the loop's scans */
  long long interface_events; /* growing cells summed over the capture loop's events (the
                                 cells the reference scans per event, S/VoxelsCA.C:386-400) */
  long long loop_cycles;      /* the capture loop's duration in shader-clock cycles (read by
                                 its engine wave), summed over its launches */
} amcafe_stats;

int amcafe_domain_create(const amcafe_domain_desc* desc, amcafe_domain** out);
void amcafe_domain_destroy(amcafe_domain* d);
/* any pointer may be NULL (that field is left unchanged) */
int amcafe_domain_upload(amcafe_domain* d, const int* gID, const int* vState, const double* extents,
                         const double* centroidOct, const int* BuildjID);
int amcafe_domain_download(amcafe_domain* d, int* gID, int* vState, double* extents, double* centroidOct,
                           int* BuildjID, double* TempCurr);
/* grain orientations cTheta (omega, axis) per grain, as VoxelsCA::cTheta */
int amcafe_domain_set_grains(amcafe_domain* d, const double* cTheta, int nGrain);
int amcafe_domain_get_grains(amcafe_domain* d, double* cTheta, int* nGrain);
/* TempField::AnalyticTempCurr: TempCurr and BuildjID on the device; *any_melt
 * is the end-of-scan input "some voxel at or above tL" (TempField.C:202-205). */
int amcafe_analytic_temp_curr(amcafe_domain* d, const amcafe_laser* laser, int* any_melt);
/* TempField::EASM (S/TempField.C:79-137; Utilities.C:60-140, 228-249) in place of
 * AnalyticTempCurr: the Gaussian moving source (absorptivity 0.56, radius 50 um, power bmP)
 * integrated over the last 8e-4 s with the 25-point Gauss-Legendre rule, 316L properties by a
 * damped Picard iteration (tolerance 10 K, at most 100 iterations), FP64 on the device.  0 K
 * above the laser plane, 0.99 T0 ahead of the laser within the top layer (layerT: the layer
 * thickness, Grid::layerT, S/TempField.C:103).  *all_melt: every
 * voxel at T >= tL (the end-of-scan test of S/TempField.C:117-126); *n_integrals: source
 * integrals evaluated.  Either may be NULL.  The reference cannot select this model from a deck
 * (its EASM keyword never matches, S/Grid.C:198-199,276).  amcafe_update_voxels after it grows
 * every cell at its own temperature, vhat = Avel (tL - T)^nvel with pow rounded to nearest. */
int amcafe_easm_temp_curr(amcafe_domain* d, const amcafe_laser* laser, double bmP, double layerT,
                          int* all_melt, long long* n_integrals);
/* VoxelsCA::UpdateVoxels on the current TempCurr; tInd = Grid::tInd. */
int amcafe_update_voxels(amcafe_domain* d, int tInd, int ilaserLoc, amcafe_stats* stats);

/* ---------------------------------------------------------------------------
 * 2. Driver level
 */
typedef struct amcafe_sim amcafe_sim;

typedef struct {
  int nX[3];
  int nGrain;
  int tInd;            /* Grid::tInd (completed steps) */
  int isp, indlayer, ilaserLoc, nlayerTot;
  int done;            /* the run's layer loop has finished (main1.C:75) */
  int inewscanflg, inewlayerflg;
  double time;
  double gth;          /* current grid rotation (Grid::gth) */
} amcafe_info;

/* n_decks deck texts (the .in file contents); they advance together. */
int amcafe_sim_create(const char* const* deck_texts, int n_decks, amcafe_sim** out);
void amcafe_sim_destroy(amcafe_sim* s);
/* one iteration of main1.C:62-93 for every unfinished deck; *n_running gets the
 * number of decks still running afterwards */
int amcafe_sim_step(amcafe_sim* s, int* n_running);
int amcafe_sim_info(amcafe_sim* s, int inst, amcafe_info* info);
int amcafe_sim_stats(amcafe_sim* s, int inst, amcafe_stats* stats);
/* The launch autotuner's state (no counterpart in the reference): the candidate-list pass's
 * chunking (blocks per instance) is picked by timing the pass at each candidate inside a step,
 * between the pass and the capture loop (the pass only rebuilds the lists from their inputs),
 * once the largest growing list reaches 256 cells and again each time it doubles
 * (AMCAFE_AUTOTUNE=0 keeps the default, 2048).  Results never depend on it. */
typedef struct {
  int list_blocks;    /* chunking in use (0: not tuned yet) */
  int tunings;        /* times the autotuner ran */
  int growing_cells;  /* largest growing list at the last tuning */
  int candidates[5];  /* the chunkings tried */
  float ms[5];        /* each candidate's time per launch at the last tuning */
} amcafe_tuning;
int amcafe_sim_tuning(amcafe_sim* s, amcafe_tuning* out);
/* reference layouts; NULL pointers are skipped; Build is VoxelsCA::Build */
int amcafe_sim_get_fields(amcafe_sim* s, int inst, int* gID, int* vState, double* TempCurr, double* extents,
                          double* centroidOct, int* BuildjID, int* Build);
int amcafe_sim_get_grains(amcafe_sim* s, int inst, double* cTheta, int* nGrain);
/* VoxelsCA::WriteToHDF1 (S/VoxelsCA.C:1592-1856) of one instance's current state to the file
 * `path` (the reference names it CA3D<deck stem><tInd>.h5, S/main1.C:53,80-84).  Layout:
 * group Step0 with dims int32[3], VoxelDX float32[3], gID, vState, Build int32[N],
 * Temperature float32[N], IPFz, IPFx, IPFy float32[3N], angleAxis float32[4 nGrain]; root
 * attribute NumSteps = 1.  Needs the HDF5 (>= 1.10) shared library at run time, loaded with
 * dlopen from $AMCAFE_HDF5_LIB or libhdf5.so.103; AMCAFE_ERR_IO when it is absent. */
int amcafe_sim_write_hdf5(amcafe_sim* s, int inst, const char* path);
/* The reference driver's own output schedule (S/main1.C:76-84, 95-97): from now on every
 * output step (tInd % outint == 0, the last step, or a new layer with outNL 0) writes
 * <prefix><tInd>.h5, and the end of the run writes it again after the final CleanLayer.
 * prefix NULL stops writing. */
int amcafe_sim_set_output(amcafe_sim* s, int inst, const char* prefix);
/* WriteToHDF1 on host buffers in the reference layouts (no device involved): nX, dX of the
 * grid, gID / vState / Build int32[N], TempCurr double[N], cTheta double[4 nGrain]. */
int amcafe_write_hdf5_fields(const char* path, const int* nX, const double* dX, const int* gID,
                             const int* vState, const int* Build, const double* TempCurr,
                             const double* cTheta, int nGrain);
/* The IPF colours WriteToHDF1 gives a grain (S/VoxelsCA.C:1597-1817): 9 floats per grain,
 * RGB of the build direction (IPFz), of x (IPFx) and of y (IPFy). */
int amcafe_ipf_colours(const double* cTheta, int nGrain, float* rgb9);
/* per-stage device time of the last step in ms when enabled: temperature, solid->mushy,
 * active list, growing-cell setup, capture loop, scatter + mushy->solid */
int amcafe_sim_set_timing(amcafe_sim* s, int on);
int amcafe_sim_timings(amcafe_sim* s, double* ms, int n);

/* ---------------------------------------------------------------------------
 * 3. Differentiable temperature field (no counterpart in the reference; for calibration
 *    callers that fit beam parameters): the moving Gaussian source of Utilities::Temp_Fun_G
 *    (S/Utilities.C:228-249) with constant properties on a straight +x track, integrated over
 *    the last t_hist seconds with the 25-point Gauss-Legendre rule of Utilities::Integral, and
 *    the loss sum over voxels and steps of (s(T) - s(T*))^2, s(T) = 1/(1+exp(-(T-tL)/width)),
 *    T* the field of the target parameters.  theta = {P [W], v [m/s], eta, sigma [m]}.
 *    Steps step0+1 .. step0+nsteps at t = n dt; each launch covers `checkpoint` steps (the field
 *    of a step is recomputed from theta and t, nothing is stored between steps).  *ms: device
 *    time (may be NULL).
 * ------------------------------------------------------------------------- */
typedef struct amcafe_heat_desc_s {
  int nX[3];
  double dX;
  double x_start, y_track, z_track; /* source at (x_start + v t, y_track, z_track) */
  double dt, t_hist;                /* step length; source-history window (8e-4 s in the reference) */
  double T0, tL, width;             /* ambient, liquidus, indicator width */
  double k, alpha;                  /* conductivity [W/m/K], diffusivity [m^2/s] */
  int checkpoint;                   /* steps per launch (0: 100) */
} amcafe_heat_desc;
/* FP32 forward with FP32 accumulation */
int amcafe_heat_loss_f32(const amcafe_heat_desc* d, const double* theta, const double* target, int step0,
                         int nsteps, double* loss, float* ms);
/* FP32 loss and its gradient with respect to theta (hand-written adjoint) */
int amcafe_heat_loss_grad_f32(const amcafe_heat_desc* d, const double* theta, const double* target, int step0,
                              int nsteps, double* loss, double* grad, float* ms);
/* FP64 forward (finite differences) */
int amcafe_heat_loss_f64(const amcafe_heat_desc* d, const double* theta, const double* target, int step0,
                         int nsteps, double* loss, float* ms);
/* n instances on the same grid and track in one launch per checkpoint segment (the calibration
 * form): theta, target and grad hold 4 values per instance, loss one (1 <= n <= 65535).  Each
 * instance's result is bit-identical to amcafe_heat_loss_grad_f32 on its own parameters.
 * All FP32 entry points sum in a fixed order (per thread over steps, per 256-voxel block in a
 * fixed tree, blocks across segments in launch order, block slots in a fixed tree), so a call
 * repeated with the same arguments returns the same bits.  They skip voxels whose temperature
 * bound stays below tL - 40 width at every step for both parameter sets (indicator below e^-40,
 * loss term below 2e-35); the FP64 forward evaluates every voxel. */
int amcafe_heat_loss_grad_batch_f32(const amcafe_heat_desc* d, int n, const double* theta, const double* target,
                                    int step0, int nsteps, double* loss, double* grad, float* ms);
/* voxel-steps (per instance) the last heat call of this thread evaluated, after pruning */
long long amcafe_heat_last_evaluations(void);
/* amcafe_heat_loss_grad_batch_f32 without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  Same kernels and the same bits as the batch entry point; no
 * timing.  Its device buffers (parameters, block sums) live in one static device array of
 * amcafe_heat_noalloc_max_bytes() (8 MiB) per GPU, held until the process ends: a call needing
 * more (amcafe_heat_noalloc_bytes: about 20 bytes per 256 voxels and instance, so 4.2 M voxels
 * for up to 25 instances) fails with AMCAFE_ERR_ARG before touching the device.  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (two host threads, or two
 * streams) race and return garbage. */
int amcafe_heat_loss_grad_batch_f32_noalloc(const amcafe_heat_desc* d, int n, const double* theta,
                                            const double* target, int step0, int nsteps, double* loss,
                                            double* grad);
long long amcafe_heat_noalloc_bytes(const amcafe_heat_desc* d, int n);
long long amcafe_heat_noalloc_max_bytes(void);

#ifdef __cplusplus
}
#endif
#endif
