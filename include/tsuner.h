/*
 * tsuner.h -- C ABI of libtsuner, the MI355X (gfx950) implementation of the
 * TSUNER per-step shallow-water + sediment solver.
 *
 * Plain C types only: pointers, sizes, POD structs.  Every entry returns an
 * int status (TSU_OK = 0); tsu_last_error() gives a message for the calling
 * thread.  State planes are planar SoA in double precision on the host:
 * V[f*nvol + j*npx + i] with f = 0 h, 1 qx, 2 qy, 3 H (depth below datum,
 * bed z = -H), 4 hc (reference layout, GPU/main_TSUNER.cu:104-111).
 *
 * Reference interfaces replaced (paths relative to the TSUNER source tree):
 *   TSUNER()            TSUNER.h:39, GPU/main_TSUNER.cu:21-222
 *   checkCUDAsupport()  TSUNER.h:41, GPU/CheckCUDASupport.cu:49-73
 *   initializeGPU()     TSUNER.h:42, GPU/InitializeGPU.cu:21-27
 *   tsu_state_* / tsu_step / tsu_init_dt split the TSUNER() time loop
 *     (GPU/main_TSUNER.cu:142-197) into re-entrant calls on an opaque handle.
 */
#ifndef TSUNER_AMD_TSUNER_H
#define TSUNER_AMD_TSUNER_H

#ifdef __cplusplus
extern "C" {
#endif

#if defined(TSU_BUILDING_LIBRARY)
#define TSU_API __attribute__((visibility("default")))
#else
#define TSU_API
#endif

/* CUDA's double2 as the reference's host code uses it (TSUNER.h:39). */
typedef struct { double x, y; } tsu_double2;

/* Same field order and types as the reference SimParams (TSUNER.h:10-30),
 * so a caller's existing struct can be passed unchanged.  FR, Q, Opar and
 * HQB hold the derived values of Problem_loader.cxx:337-342. */
typedef struct {
    double xmin, xmax, ymin, ymax;
    double total_time, save_time;
    int npx, npy, nvol;
    double CFL;
    double Q, R;
    double porosity;
    double DS;
    double FR;
    double Opar;
    double OC;
    double Dsg;
    double HQB;
    double vmax;
    double epsilon_h;
    double dx;
    double dy;
} tsu_sim_params;

/* The reference's compile-time switches (Sim_parameters.hxx:40-54) as
 * runtime fields, plus the selectable laws.  bc[] order: left (x = xmin),
 * right (x = xmax), bottom (y = ymin), top (y = ymax). */
enum { TSU_BC_FREE = 0, TSU_BC_WALL = 1, TSU_BC_PERIODIC = 2, TSU_BC_INFLOW = 3 };
enum { TSU_LIM_VANLEER = 0, TSU_LIM_MINMOD = 1 };
enum { TSU_BEDLOAD_ASHIDA_MICHIUE = 0, TSU_BEDLOAD_GRASS = 1 };
enum { TSU_FRICTION_QUADRATIC = 0, TSU_FRICTION_MANNING = 1 };

typedef struct {
    int order;          /* 1 forward Euler, 2 two-stage RK (Sim_parameters.hxx:48) */
    int bedload;        /* BEDLOAD 0/1 (Sim_parameters.hxx:45) */
    int suspension;     /* SUSPENSION 0/1 (Sim_parameters.hxx:46) */
    int bc[4];          /* TSU_BC_* per side (WBC, Read_files: Sim_parameters.hxx:51,54) */
    int limiter;        /* TSU_LIM_* ; reference: van Leer (Reconstructor.cu:19-31) */
    int bedload_law;    /* TSU_BEDLOAD_* ; reference: Ashida-Michiue (SedFlux.cuh:150-168) */
    int friction_law;   /* TSU_FRICTION_* ; reference: quadratic (Reconstructor.cu:45-53) */
    int dt_narrow;      /* 1: dt passes through float for nvol > 4096 (Reduction_kernel_min.cu:46-94) */
    int extrema;        /* 1: keep running hmax, max H, eta_max (Volume_kernel.cu:79-110) */
    double grass_ag;    /* Grass coefficient A_g (bedload_law == GRASS) */
    double manning_n;   /* Manning n (friction_law == MANNING) */
} tsu_scheme;

enum { TSU_FP64 = 0, TSU_FP32 = 1 };

enum {
    TSU_OK = 0,
    TSU_ERR_DEVICE_MEMORY = 1, /* reference code 1: device allocation failed */
    TSU_ERR_HOST = 2,          /* reference code 2: host allocation / output init failed */
    TSU_ERR_ARG = 3,
    TSU_ERR_HIP = 4,
    TSU_ERR_UNSUPPORTED = 5
};

/* ---- reference-compatible entries ------------------------------------- */

/* Same arguments, meaning and return codes as TSUNER.h:39 (the reference's
 * double& tiempo is a pointer at ABI level).  Runs the reference
 * configuration (second order, bedload and suspension on; inflow series on
 * all sides when CC/tCC are given, walls otherwise) in FP64 and writes
 * <input_prefix>.nc.  d, dyrup, time_dyrup and the tides arguments are
 * accepted and unused, as in the reference build (SPL = 0). */
TSU_API int TSUNER(double *dVh, tsu_double2 *dVhu, double *dVH, int *fixH, tsu_sim_params params,
                   char *input_prefix, double *d, double *CC, double *tCC, unsigned long timeCC,
                   double *dyrup, float *time_dyrup, double *time_gpu, int device,
                   double *tides_bd, float *tides_time, int tidesNPT);

/* GPU/CheckCUDASupport.cu:49-73: 0 ok, 1 no GPU, 2 no supported GPU
 * (here: not gfx950), 3 bad device index. */
TSU_API int checkCUDAsupport(int device);
TSU_API int tsu_device_check(int device);
/* GPU/InitializeGPU.cu:21-27: warm-up launch. */
TSU_API void initializeGPU(void);

/* TSUNER() with an explicit scheme and precision (TSU_FP64 / TSU_FP32). */
TSU_API int tsu_run(const double *V, const int *fixH, const tsu_sim_params *params, const tsu_scheme *scheme,
                    int precision, const char *output_prefix, const double *CC, const double *tCC,
                    unsigned long timeCC, double *time_gpu, int device);

/* Host-only output writer (GPU/netcdf.cu:38-435 variables and attributes):
 * writes <output_prefix>.nc with nrec snapshots (times[nrec]; V_records holds
 * nrec consecutive 5*nvol state blocks) and, where non-null, hmax, bmin
 * (= -bmax, GPU/main_TSUNER.cu:207-209) and eta_max. */
TSU_API int tsu_write_output(const char *output_prefix, const tsu_sim_params *params, int nrec,
                             const double *times, const double *V_records, const double *hmax,
                             const double *bmax, const double *eta_max);

/* Reads the reference's positional data.dat (data.dat:1-45, parsed as
 * Problem_loader.cxx:264-335) into params with the derived fields of
 * Problem_loader.cxx:337-342 and dx = (xmax-xmin)/npx, dy likewise.  prefix
 * receives the output prefix; files the bathymetry, eta/CC, fixed-bed and
 * tides file names separated by newlines. */
TSU_API int tsu_parse_data_dat(const char *path, tsu_sim_params *params, char *prefix, int prefix_len,
                               char *files, int files_len, double *sea_level_correction);

/* ---- real-data input (the reference's Read_files = 1 build) ------------- */

/* Number of time levels of the free-surface series file (the length of its
 * time/TIME dimension, Problem_loader.cxx:153-161): sizes CC and tCC for
 * tsu_load_input_files.  NetCDF classic and NetCDF-4 (HDF5) files are read;
 * the latter through the HDF5 C library loaded at run time (libhdf5.so, or
 * the path in TSU_HDF5_LIB). */
TSU_API int tsu_input_time_levels(const char *eta_file, unsigned long *time_levels);

/* Replaces LoadProblemData's file branch (Problem_loader.cxx:355-624, with
 * interpolations.cxx:46-95,368-1114): the bathymetry grid (variable
 * "bathy", coordinates lon/x/LON and lat/y/LAT, depth positive down), the
 * free-surface series eta_file (ha_lc/HA/ha, cm) with its siblings whose
 * 4th character from the end is 'u' and 'v' (ua_lc/UA/ua, va_lc/VA/va,
 * cm/s) and an optional fixed-bed map (NULL, "" or "0" for none; 1 = bed may
 * move) become, on params->npx x params->npy cells, the Cartesian initial
 * state h, hu (= 0), H, fixH and the boundary series CC
 * ((2*npx+2*npy)*4*time_levels doubles, VolumeSolver.cu:172-202 layout) and
 * tCC (time_levels doubles).  All buffers are caller-owned.  On return
 * params holds xmin = ymin = 0, xmax, ymax in metres, dx, dy and nvol; its
 * other fields are untouched.  Errors: TSU_ERR_ARG for shapes and sizes,
 * TSU_ERR_HOST for unreadable files (tsu_last_error() names the file). */
TSU_API int tsu_load_input_files(const char *bathy_file, const char *eta_file, const char *fixed_bed_file,
                                 double sea_level_correction, tsu_sim_params *params, double *dVh,
                                 tsu_double2 *dVhu, double *dVH, int *fixH, double *CC, double *tCC,
                                 unsigned long time_levels);

/* ---- handle API --------------------------------------------------------- */

typedef struct tsu_state tsu_state;

TSU_API int tsu_state_create(tsu_state **out, const tsu_sim_params *params, const tsu_scheme *scheme,
                             int precision, int device);
TSU_API int tsu_state_destroy(tsu_state *st);
/* V: 5*nvol doubles (host); fixH: nvol ints in {0, 1} (host). */
TSU_API int tsu_state_upload(tsu_state *st, const double *V, const int *fixH);
TSU_API int tsu_state_download(tsu_state *st, double *V);
/* CC: (2*npx+2*npy)*4*timeCC doubles, tCC: timeCC doubles, layout of
 * VolumeSolver.cu:172-202; needed when any side is TSU_BC_INFLOW. */
TSU_API int tsu_state_set_inflow(tsu_state *st, const double *CC, const double *tCC, unsigned long timeCC);
/* dt of the current state (GPU/main_TSUNER.cu:147-148), kept on device. */
TSU_API int tsu_init_dt(tsu_state *st);
/* Enqueue nsteps full time steps; no host synchronisation. */
TSU_API int tsu_step(tsu_state *st, int nsteps);
TSU_API int tsu_synchronize(tsu_state *st);
TSU_API int tsu_get_time(tsu_state *st, double *t, double *dt, long long *steps);
TSU_API int tsu_set_time(tsu_state *st, double t, double dt);
/* Steps enqueued after the simulation time reaches stop_time do nothing. */
TSU_API int tsu_set_stop_time(tsu_state *st, double stop_time);
/* on != 0: the step that would cross stop_time is shortened to end on it
 * exactly (dt = stop_time - t).  Off by default, as in the reference loop
 * (GPU/main_TSUNER.cu:157-197), whose last step may overshoot. */
TSU_API int tsu_set_clamp(tsu_state *st, int on);
/* FP64 sums of h and of (H + hc) over all cells. */
TSU_API int tsu_mass(tsu_state *st, double out[2]);
TSU_API int tsu_get_extrema(tsu_state *st, double *hmax, double *bmax, double *eta_max);
/* Rows marched by one wave (tile height); 0 picks the default. */
TSU_API int tsu_set_tile_rows(tsu_state *st, int rows);
TSU_API int tsu_get_tile_rows(tsu_state *st);
/* Replay steps through a captured HIP graph of chunk steps (0 disables). */
TSU_API int tsu_set_graph_chunk(tsu_state *st, int chunk);
/* hipEvent timing of every stage-kernel launch enqueued while on (eager
 * steps only; graph replays are not timed).  tsu_get_timing synchronises and
 * returns the summed milliseconds and launch counts per RK stage (arrays of
 * 2); reset != 0 clears the accumulators. */
TSU_API int tsu_set_timing(tsu_state *st, int on);
TSU_API int tsu_get_timing(tsu_state *st, double stage_ms[2], long long launches[2], int reset);
/* The stream all work of this handle is enqueued on (a hipStream_t). */
TSU_API void *tsu_stream(tsu_state *st);
/* Device pointer of the U^n planes (element size 8 or 4 by precision). */
TSU_API void *tsu_device_state(tsu_state *st);

/* ---- ensembles ------------------------------------------------------------
 * n independent scenarios (grids of any size) advanced together: one launch
 * per RK stage covers every member, each member keeps its own dt and time,
 * and whole-ensemble steps replay from a captured HIP graph.  The reference
 * runs one scenario per TSUNER() call (GPU/main_TSUNER.cu:21-222); members
 * here follow the same arithmetic.  params and scheme are arrays of n; all
 * members share the RK order.  A member with an inflow side takes its own
 * boundary series (tsu_ensemble_set_inflow) before stepping. */
typedef struct tsu_ensemble tsu_ensemble;

TSU_API int tsu_ensemble_create(tsu_ensemble **out, int n, const tsu_sim_params *params,
                                const tsu_scheme *scheme, int precision, int device);
TSU_API int tsu_ensemble_destroy(tsu_ensemble *e);
TSU_API int tsu_ensemble_upload(tsu_ensemble *e, int member, const double *V, const int *fixH);
TSU_API int tsu_ensemble_download(tsu_ensemble *e, int member, double *V);
TSU_API int tsu_ensemble_init_dt(tsu_ensemble *e);
/* Boundary series of one member, in the layout of tsu_state_set_inflow
 * (VolumeSolver.cu:172-202 read per member at that member's time). */
TSU_API int tsu_ensemble_set_inflow(tsu_ensemble *e, int member, const double *CC, const double *tCC,
                                    unsigned long timeCC);
/* Members stop (their steps become no-ops) after stop_step steps. */
TSU_API int tsu_ensemble_set_stop_step(tsu_ensemble *e, long long stop_step);
TSU_API int tsu_ensemble_step(tsu_ensemble *e, int nsteps);
TSU_API int tsu_ensemble_synchronize(tsu_ensemble *e);
TSU_API int tsu_ensemble_get_time(tsu_ensemble *e, int member, double *t, double *dt, long long *steps);
TSU_API int tsu_ensemble_mass(tsu_ensemble *e, int member, double out[2]);
/* Tile height per member (n ints), e.g. chosen by tsuner_amd.autotune. */
TSU_API int tsu_ensemble_set_tile_rows(tsu_ensemble *e, const int *rows);
TSU_API int tsu_ensemble_set_graph_chunk(tsu_ensemble *e, int chunk);
/* restore == 0 saves every member's state and step control; 1 restores it
 * (second-order schemes). */
TSU_API int tsu_ensemble_snapshot(tsu_ensemble *e, int restore);
/* hipEvent-timed wall time of nsteps eager whole-ensemble steps. */
TSU_API int tsu_ensemble_time_steps(tsu_ensemble *e, int nsteps, double *ms);
/* Per-member stop times (n doubles); clamp as in tsu_set_clamp. */
TSU_API int tsu_ensemble_set_stop_time(tsu_ensemble *e, const double *stop_time, int clamp);
/* Running maxima of h, H and eta of member (schemes with extrema on),
 * updated after every step as in GPU/Volume_kernel.cu:79-137. */
TSU_API int tsu_ensemble_get_extrema(tsu_ensemble *e, int member, double *hmax, double *bmax, double *eta_max);
TSU_API void *tsu_ensemble_stream(tsu_ensemble *e);
/* Times and step counts of all members (n each; either may be NULL), one
 * synchronisation. */
TSU_API int tsu_ensemble_get_times(tsu_ensemble *e, double *t, long long *steps);

/* ---- learned coarse-grid correction (no counterpart in the reference) ------
 * A 4-layer 3x3 convolutional network (8 -> 32 -> 32 -> 32 -> 4 channels,
 * ReLU) run on BF16 MFMA with FP32 accumulation.  Activations are NHWC BF16
 * (uint16 bit patterns), n images of h x w pixels.  Weights: BF16
 * [9 taps][out][in], tap = 3*(dy+1) + (dx+1); the output layer holds 16 rows
 * (4 used).  All pointers are device pointers; stream is a hipStream_t (NULL:
 * the null stream).  Calls only enqueue work. */
enum {
    TSU_NN_FWD_IN = 0,     /* x[8] -> relu(conv + b)[32]                       */
    TSU_NN_FWD_MID = 1,    /* [32] -> relu(conv + b)[32]                        */
    TSU_NN_FWD_OUT = 2,    /* [32] -> conv + b, FP32 [4]                        */
    TSU_NN_DGRAD_OUT = 3,  /* dy[8] (4 used) -> dx[32] * (act > 0); w: [9][32][8] */
    TSU_NN_DGRAD_MID = 4,  /* dy[32] -> dx[32] * (act > 0); w: [9][32][32]      */
    TSU_NN_WGRAD_IN = 5,   /* dW[9][32][8], db[32] from x[8], dy[32]            */
    TSU_NN_WGRAD_MID = 6,  /* dW[9][32][32], db[32] from x[32], dy[32]          */
    TSU_NN_WGRAD_OUT = 7,  /* dW[9][mrows][32] rows 0..7, db[8] from x[32], dy[8] */
    TSU_NN_W64 = 16        /* OR-ed with a kind: hidden width 64 instead of 32  */
};
/* Convolution kinds 0-4 (optionally | TSU_NN_W64).  act: the layer input's
 * activations (ReLU mask of the data-gradient kinds), else unused. */
TSU_API int tsu_nn_conv(int kind, const void *in, const void *w, const float *bias, const void *act, void *out,
                        int n, int h, int w_, void *stream);
/* Weight-gradient kinds 5-7 (optionally | TSU_NN_W64; FP32 dW, db);
 * workspace: FP32 scratch of tsu_nn_wgrad_workspace(kind) floats. */
TSU_API long long tsu_nn_wgrad_workspace(int kind);
TSU_API int tsu_nn_wgrad(int kind, const void *x, const void *dy, float *dw, float *db, int mrows, int n, int h,
                         int w_, float *workspace, long long workspace_floats, void *stream);
/* No counterpart in the reference.  tsu_nn_wgrad for callers that cannot
 * allocate (code under HIP graph capture, training or serving loops that
 * forbid allocation): the same kernels and result, with the workspace in a
 * fixed device array of 9,453,568 floats (37.8 MB, the largest kind's need)
 * that the library holds on each device until the process ends.  No memory is
 * allocated by the call.  Calls on one device share that array: calls that
 * overlap on one device (different streams, or threads) race. */
TSU_API int tsu_nn_wgrad_noalloc(int kind, const void *x, const void *dy, float *dw, float *db, int mrows, int n,
                                 int h, int w_, void *stream);
/* dy[p][0..3] = BF16(scale*(out - target)), dy[p][4..7] = 0 for npix pixels
 * of 4 FP32 channels; loss_part[b] (nblocks doubles) = partial sums of the
 * squared error. */
TSU_API int tsu_nn_mse(const float *out, const float *target, void *dy, double *loss_part, int nblocks,
                       long long npix, float scale, void *stream);
/* Adam step on n FP32 parameters; lr_t includes the bias correction. */
TSU_API int tsu_nn_adam(float *p, const float *g, float *m, float *v, long long n, float lr_t, float beta1,
                        float beta2, float eps, void *stream);
/* FP32 [9][m][k] -> BF16 same layout (wb) and, if wt != NULL, the per-tap
 * transpose wt[9][k][mt] of rows m < mt. */
TSU_API int tsu_nn_pack(const float *w, void *wb, void *wt, int m, int k, int mt, void *stream);
/* Network inputs of every member of a two-stage ensemble whose members share
 * one grid: x[member][y][x][8] BF16 = (h, qx, qy, z = -H, dz/dx, dz/dy, eta if
 * h > wet_h else 0, wet flag) times scale_in[0..7]. */
TSU_API int tsu_ensemble_nn_features(tsu_ensemble *e, const float scale_in[8], float wet_h, void *x);
/* Adds out[member][y][x][0..3] * scale_out to (h, qx, qy, H) of the wet
 * cells (h > wet_h) of the members whose step count equals expect_step; h is
 * clipped at 0, and q and hc vanish where it reaches 0.  Dry cells are left to
 * the solver unless x (the network inputs of this window, as written by
 * tsu_ensemble_nn_features) is given: then a dry cell with a wet 4-neighbour
 * takes the depth and discharge corrections if they leave it deeper than
 * shore_h (its bed is unchanged).  resid_max (n floats on the device, or
 * NULL) is raised to the largest |correction of h - H| applied to a wet cell
 * of each member (the fallback indicator).  The corrected members' dt for
 * their next step is recomputed from the corrected state, so that step keeps
 * to the CFL limit of what it advances. */
TSU_API int tsu_ensemble_nn_apply(tsu_ensemble *e, const float scale_out[4], float wet_h, const float *out,
                                  long long expect_step, float *resid_max, const void *x, float shore_h);
/* coarse member i = free-surface restriction of fine member i (fine grids 4x
 * the coarse grid in each direction, same precision), then the coarse dt.
 * Per 4x4 block: H = block mean; h = max(mean of h - H over fine cells with
 * h > 1e-3 m, plus that H, 0), 0 if none; qx, qy, hc = block totals times
 * h / (block total of h). */
TSU_API int tsu_ensemble_restrict(tsu_ensemble *fine, tsu_ensemble *coarse);
/* target[i][y][x][0..3] = (restrict(fine_i) - coarse_i) / scale_out for
 * (h, qx, qy, H), FP32. */
TSU_API int tsu_ensemble_nn_target(tsu_ensemble *fine, tsu_ensemble *coarse, const float scale_out[4],
                                   float *target);

/* Achievable HBM bandwidth: times reps launches of a plain device copy of
 * `bytes` bytes and returns (read + write bytes) / mean launch time in GB/s. */
TSU_API int tsu_bandwidth_probe(int device, long long bytes, int reps, double *gbs);

TSU_API const char *tsu_last_error(void);
TSU_API const char *tsu_version(void);

/* ---- known-answer entries (device functions on arrays of inputs) -------- */

/* n edges; left/right: 5 doubles each (h, qx, qy, H, hc, unrotated);
 * normal: 2 ints each, one of (1,0) (-1,0) (0,1) (0,-1); fixed: n ints.
 * Dm, Dp: 5 doubles each, unrotated (VolumeSolver.cu:19-170). */
TSU_API int tsu_kat_edge_flux(const tsu_sim_params *params, const tsu_scheme *scheme, int n,
                              const double *left, const double *right, const int *normal,
                              const int *fixed, double *Dm, double *Dp, int device);
/* The same with edges 2k and 2k+1 solved together, as the stage kernels
 * solve a row's x and y edges (one pass through the solver's phases with
 * wave-wide tests over both); n even, otherwise TSU_ERR_ARG. */
TSU_API int tsu_kat_edge_flux_pairs(const tsu_sim_params *params, const tsu_scheme *scheme, int n,
                                    const double *left, const double *right, const int *normal,
                                    const int *fixed, double *Dm, double *Dp, int device);
/* what: 0 limiter(a,b) -> o1; 1 velocity(h=a,qx=b,qy=c) -> o1,o2,o3=|u|;
 * 2 state filter(a,b,c,d) -> o1..o4; 3 cell dt(a,b,c) -> o1;
 * 4 erosion rate(h=a,qx=b,qy=c,hc=d) -> o1; 5 bedload(h=a,un=b,ut=c) -> o1;
 * 6 shared-denominator division: o1 = 1/a, o2 = b/a (a in [2^-60, 2^60]),
 *   o3, o4 the same by IEEE division; 7 square root without operand scaling
 *   (a = 0 or in [2^-767, DBL_MAX]) -> o1, IEEE sqrt(a) -> o2. */
TSU_API int tsu_kat_point(const tsu_sim_params *params, const tsu_scheme *scheme, int what, int n,
                          const double *a, const double *b, const double *c, const double *d,
                          double *o1, double *o2, double *o3, double *o4, int device);

#ifdef __cplusplus
}
#endif
#endif
