/*
 * tsu_oracle.h -- CPU oracle for the TSUNER per-step solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial FP64 restatement of the
 * reference's kernel arithmetic, written in the reference's own operation
 * order.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker (or the timed CPU baseline).
 * The product library (libtsuner) never links or calls it.
 *
 * Reference files restated (paths relative to the TSUNER source tree):
 *   Sim_parameters.hxx:28-54       constants and scheme switches
 *   TSUNER.h:10-30                 SimParams
 *   Problem_loader.cxx:337-342     derived parameters (FR/8, HQB, Q, Opar)
 *   GPU/SedFlux.cuh:26-199         velocity, filters, sediment laws, Roe, eigenvalues
 *   GPU/Reconstructor.cu:19-356    MUSCL reconstruction, friction, erosion/deposition
 *   GPU/VolumeSolver.cu:19-408     edge Riemann flux, boundary edges, inflow series
 *   GPU/Volume_kernel.cu:19-43     per-cell CFL dt
 *   GPU/Volume_kernel.cu:79-137    running extrema
 *   GPU/Volume_kernel.cu:180-301   cell update and RK combination
 *   GPU/Reduction_kernel_min.cu:26-196  dt min reduction (float narrowing)
 *   GPU/main_TSUNER.cu:142-197     stage sequencing, time advance, dt handoff
 *
 * Built with gcc -O2 -ffp-contract=off (no FMA contraction, matching the
 * reference's -fmad=false, GPU/makefile:10).
 */
#ifndef TSU_ORACLE_H
#define TSU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as the reference SimParams (TSUNER.h:10-30). */
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
} orc_sim_params;

/* Compile-time switches of Sim_parameters.hxx:40-54 as runtime fields, plus
 * the selectable laws the benchmark configurations name.  Index order of
 * bc[]: left (x=0), right (x=max), bottom (y=0), top (y=max). */
enum { ORC_BC_FREE = 0, ORC_BC_WALL = 1, ORC_BC_PERIODIC = 2, ORC_BC_INFLOW = 3 };
enum { ORC_LIM_VANLEER = 0, ORC_LIM_MINMOD = 1 };
enum { ORC_BEDLOAD_AM = 0, ORC_BEDLOAD_GRASS = 1 };
enum { ORC_FRIC_CHEZY = 0, ORC_FRIC_MANNING = 1 };

typedef struct {
    int order;          /* 1: forward Euler, 2: two-stage RK (Sim_parameters.hxx:48) */
    int bedload;        /* BEDLOAD (Sim_parameters.hxx:45) */
    int suspension;     /* SUSPENSION (Sim_parameters.hxx:46) */
    int bc[4];          /* boundary type per side */
    int limiter;        /* ORC_LIM_* */
    int bedload_law;    /* ORC_BEDLOAD_* */
    int friction_law;   /* ORC_FRIC_* */
    int dt_narrow;      /* 1: dt min passes through float for nvol > 4096 */
    int extrema;        /* 1: update hmax / max H / eta_max after each step */
    double grass_ag;    /* Grass coefficient A_g (bedload_law == GRASS) */
    double manning_n;   /* Manning n (friction_law == MANNING) */
} orc_scheme;

typedef struct orc_sim orc_sim;

/* State planes are planar SoA: V[f*nvol + j*npx + i], f = h, qx, qy, H, hc. */
orc_sim *orc_create(const orc_sim_params *p, const orc_scheme *s,
                    const double *V, const int *fixH,
                    const double *CC, const double *tCC, long timeCC);
void   orc_destroy(orc_sim *sim);
/* Advance nsteps full time steps (all RK stages, t += dt, extrema, new dt). */
int    orc_step(orc_sim *sim, int nsteps);
/* One RK stage only (k = 0 or 1), leaving t and dt untouched. */
int    orc_stage(orc_sim *sim, int k);
void   orc_get_state(const orc_sim *sim, double *V);     /* the V planes */
void   orc_get_state0(const orc_sim *sim, double *V0);   /* the V0 planes */
void   orc_get_extrema(const orc_sim *sim, double *hmax, double *bmax, double *eta_max);
double orc_time(const orc_sim *sim);
double orc_dt(const orc_sim *sim);
void   orc_set_time(orc_sim *sim, double t, double dt);
long   orc_steps(const orc_sim *sim);

/* Point functions, exported for known-answer tests. */
double orc_limiter(int law, double a, double b);
void   orc_velocity(double h, double qx, double qy, double *ux, double *uy);
void   orc_state_filter(double *h, double *qx, double *qy, double *hc);
double orc_bedload_flux(const orc_sim_params *p, const orc_scheme *s, double h, double un, double ut);
void   orc_edge_flux(const orc_sim_params *p, const orc_scheme *s,
                     const double left[5], const double right[5],
                     int normal_x, int normal_y, int fixed,
                     double Dm[5], double Dp[5]);
double orc_cell_dt(const orc_sim_params *p, double h, double qx, double qy);
double orc_min_reduction(const double *dtv, long n, int narrow);
void   orc_erosion_deposition(const orc_sim_params *p, const orc_scheme *s,
                              double h, double qx, double qy, double hc, double *phi);
/* One cell's filtered faces along one axis from (minus neighbour, cell,
 * plus neighbour), 5 values each in the axis frame; returns deta. */
double orc_cell_faces(int limiter, const double c3[15], double Fm[5], double Fp[5]);
/* Derive Q, Opar, HQB and FR/8 from data.dat values (Problem_loader.cxx:337-342). */
void   orc_derive_params(orc_sim_params *p);
/* Test support (tsu_quotient.c): seeded check of the shared-denominator
 * quotient used by the FP64 GPU path against x / d; returns mismatches. */
long   orc_shared_quotient_mismatches(long n, unsigned long long seed);
/* The same for numerators from 2^-1074 to 2^-900, with or without the
 * device's fallback to x / d below 2^-968. */
long   orc_shared_quotient_tiny_mismatches(long n, unsigned long long seed, int guarded);

#ifdef __cplusplus
}
#endif
#endif
