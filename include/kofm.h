/*
 * kofm.h -- C ABI of the MI355X-native lattice-Boltzmann + k-omega time step.
 *
 * The reference (zmills04/kOmegaFoulingModel) has no FFI: its hot path is
 * reached through the C++ classes clVariablesLB, kOmega, shearStress,
 * clVariablesLS and Reducer, which drive OpenCL kernels.  Each entry point
 * below replaces the reference interface cited beside it (paths relative to
 * the reference checkout).  The C++ facades with the reference's class and
 * method names (komegafoulingmodel_amd/csrc/facade.hpp) call only these
 * functions, as does any other host language (ctypes, cgo, JNI, N-API).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative error code (the
 *    reference exit codes of ConstDef.h:588-611, negated); the process is
 *    never terminated.  kofm_last_error() describes the last failure.
 *  - Host buffers stay owned by the caller and are copied.  Host arrays use
 *    the reference layout: x + XsizeFull*y (+ XsizeFull*nY*k for planes),
 *    XsizeFull = ceil(nX/128)*128 (clProblem.cpp:469), directions in the
 *    order C,E,W,N,S,NE,SW,SE,NW (Heading.cl:197-215), FP64 unless stated.
 *  - A context owns all device memory and one HIP stream on the current
 *    device.  A context is not thread-safe; separate contexts may run
 *    concurrently.  All calls are asynchronous on that stream except the
 *    ones that return host data (download, reduce, counts, timing).
 */
#ifndef KOFM_H
#define KOFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kofm_ctx kofm_ctx;

/* error codes (ConstDef.h:588-611, negated, plus argument/device errors) */
#define KOFM_OK 0
#define KOFM_ERR_ARG (-1)
#define KOFM_ERR_DEVICE (-2)
#define KOFM_ERR_STATE (-3)
#define KOFM_ERR_INITIALIZING_VLS (-20)
#define KOFM_ERR_INITIALIZING_VLB (-21)
#define KOFM_ERR_INITIALIZING_KOMEGA (-22)
#define KOFM_ERR_SOLVER_NAN (-30)

/* population storage in HBM (all forms collide in FP32 except F64) */
#define KOFM_STORE_F32 0  /* FP32 deviations f_i - w_i */
#define KOFM_STORE_F16 1  /* FP16 scaled deviations (f_i - w_i) * 1024 / w_i */
#define KOFM_STORE_BF16 2 /* BF16 scaled deviations (f_i - w_i) * 1024 / w_i */
#define KOFM_STORE_F64 3  /* FP64 deviations, FP64 arithmetic (GPU FP64 oracle) */
#define KOFM_STORE_F32X 4 /* FP32 deviations, FP64 arithmetic, stochastically rounded
                             stores: FP32 bytes with long-run drift near FP64 (one step
                             per pass, no captured graphs, no ensemble or adjoint) */

/* options for kofm_step (Heading.cl:309) */
#define KOFM_OPTION_SAVE_MACRO_FIELDS 0x2

/* fields for kofm_download / kofm_reduce */
#define KOFM_FIELD_F 0        /* 9*D populations, reference FA/FB layout       */
#define KOFM_FIELD_RHO 1      /* D   Ro_array                                   */
#define KOFM_FIELD_UX 2       /* D   Ux_array (u_x + Fval/2)                    */
#define KOFM_FIELD_UY 3       /* D   Uy_array                                   */
#define KOFM_FIELD_KAPPA 4    /* D   Kappa_array                                */
#define KOFM_FIELD_OMEGA 5    /* D   Omega_array                                */
#define KOFM_FIELD_NUT 6      /* D   Nut_array                                  */
#define KOFM_FIELD_SXY 7      /* 3*D Sxy_array (Sxx, Sxy, Syy planes)           */
#define KOFM_FIELD_WALLD 8    /* D   WallD                                      */
#define KOFM_FIELD_TAU_NODE 9 /* nSS shearStress::Tau (per wall node)           */
#define KOFM_FIELD_TAU_BL 10  /* nBLshear  BL.Tau over the shear link range     */
#define KOFM_FIELD_DIFF_K 11  /* D   kOmega::Diff_K (k diffusivity; a node keeps its
                                 last value when a wall update makes it solid)   */
#define KOFM_FIELD_DIFF_OMEGA 12 /* D kOmega::Diff_Omega                           */
#define KOFM_FIELD_TEMP 13    /* D   clVariablesFD::tempArray (temperature)      */

/* reduction operators (ReduceGenerator.cpp:219-305) */
#define KOFM_RED_SUM 1
#define KOFM_RED_MAX 2
#define KOFM_RED_MIN 4

/* geometry arrays for kofm_geometry_array */
#define KOFM_GEOM_NTYPE 0     /* int16 [D]       clVariablesLS::nType            */
#define KOFM_GEOM_DX 1        /* double[4*D]     clVariablesLS::dXArr (E,W,N,S)  */
#define KOFM_GEOM_IBB 2       /* int32 [nIbb]    ibbArr (node + D*dir)           */
#define KOFM_GEOM_IBBQ 3      /* double[nIbb]    ibbDistArr                      */
#define KOFM_GEOM_SS 4        /* int32 [nSS]     ssArr                           */
#define KOFM_GEOM_SSMAP 5     /* int32 [D]       ssArrIndMap                     */
#define KOFM_GEOM_SSINDS 6    /* int32 [nX]      ssArrInds                       */
#define KOFM_GEOM_LSMAP 7     /* uint16[2*XF]    lsMap                           */
#define KOFM_GEOM_WALLD 8     /* double[D]       kOmega::WallD                   */
#define KOFM_GEOM_CFSS 9      /* double[16*nSS]  shearStress::shearCoeffs        */
#define KOFM_GEOM_SIND 10     /* int32 [4*nBLs]  shearStress::sInds              */
#define KOFM_GEOM_WEIGHTS 11  /* double[4*nBLs]  shearStress::ssWeights          */
#define KOFM_GEOM_BOUNDS 12   /* int32 [4]       BL_BOT_START, NUM_BL_BOT, BL_TOP_START, NUM_BL_TOP */
#define KOFM_GEOM_MASSES 13   /* double[2]       clVariablesLS::Masses           */
/* no reference counterpart: sensitivities used by kofm_adjoint */
#define KOFM_GEOM_IBB_SEG 14  /* int32 [nIbb]    wall segment (s, s+1) each link cuts */
#define KOFM_GEOM_IBB_DQDY 15 /* double[2*nIbb]  dq/dy of wall points s and s+1     */

typedef struct kofm_grid {
  int32_t nX;         /* "nX" */
  int32_t nY;         /* "nY" */
  int32_t storage;    /* KOFM_STORE_* */
  int32_t use_komega; /* "Use Turb Model" */
} kofm_grid;

typedef struct kofm_params {
  double tau0;          /* relaxation frequency 1/(3 nu + 0.5) (reference 'tau', #define tau0) */
  double Fval;          /* body-force velocity increment per step (FTERM_VAL) */
  double nu;            /* kinematic viscosity (MU_NUMBER) */
  double rho0;          /* "Rho" */
  double max_turb_visc; /* "Max Turb Visc" (MAX_TURB_VISC) */
  double ko_dt;         /* k-omega time step (kOmega::timeStep, DTFD) */
  double kappa_rel_tol; /* "Kappa Max Rel Tol" */
  double kappa_abs_tol; /* "Kappa Max Abs Tol" */
  int32_t kappa_max_iters; /* "Kappa Max Iterations" */
  double omega_rel_tol; /* "Omega Max Rel Tol" */
  double omega_abs_tol; /* "Omega Max Abs Tol" */
  int32_t omega_max_iters; /* "Omega Max Iterations" */
} kofm_params;

typedef struct kofm_geom_params {
  double pipe_radius;          /* "Pipe Radius" (Channel_Height = 2*(int)R) */
  double eps;                  /* clProblem::eps (1e-5) */
  double min_dx;               /* "Min dX Dist" (MINIMUM_DX_DIST) */
  int32_t wall_d_search_rad;   /* "Wall D Search Rad" */
  double x_release_pos;        /* "X Release Pos" */
  double x_min_val;            /* "X Min Val" (reference default X Release Pos - 1.25) */
  double x_stop_pos;           /* "X Stop Pos" */
  double cutoff_radius;        /* "Cutoff Radius" (SHEAR_CUTOFF_RADIUS) */
  int32_t index_radius_search; /* "Index Radius Search" */
} kofm_geom_params;

typedef struct kofm_geom_counts {
  int32_t nX, nY, XsizeFull;
  int64_t D;          /* XsizeFull * nY */
  int32_t nIbb;       /* boundary links (ibbArr entries) */
  int32_t nSS;        /* wall nodes (ssArr entries) */
  int32_t nBLshear;   /* boundary links in the shear range (NUM_BL_TOTAL) */
  int32_t nIbbNodes;  /* distinct nodes carrying interpolated links */
} kofm_geom_counts;

typedef struct kofm_solver_stats {
  int32_t kappa_iters;  /* sweeps of the last k solve */
  int32_t omega_iters;  /* sweeps of the last omega solve */
  double kappa_res0, kappa_res;
  double omega_res0, omega_res;
  int32_t nan_flag;     /* non-finite residuals seen since the last query (each one
                           halved the k-omega time step) */
  int32_t steps_per_lb; /* k-omega sub-steps per LB step (kOmega::stepsPerLB) */
  double time_step;     /* their time step (kOmega::timeStep) */
  int32_t batches;      /* batches of 8 k-omega steps replayed from one graph, with one
                           poll of the solver state, since the last query */
  int32_t batch_rollbacks; /* of those, batches in which a solve needed more
                              iterations than provisioned or met a non-finite
                              residual: they were rerun one step at a time */
  int32_t temp_iters;      /* temperature: iterations of the last solve */
  double temp_res0, temp_res; /* its first and last residual norms */
  int32_t temp_steps_per_lb;  /* clVariablesFD::stepsPerLB */
  double temp_time_step;      /* clVariablesFD::timeStep */
} kofm_solver_stats;

/* The temperature solve's constants (clVariablesFD::loadParams,
 * clVariablesFD.cpp:445-497; the facade derives them from RunParam.yaml).
 * alpha_fluid = MuVal / Pr, alpha_foul = (k_soot / rho_soot / cp_soot)
 * * DELTA_L^2 / DELTA_T (the fouling layer), pr_turb the turbulent Prandtl
 * number, t_in the inlet temperature "Tinlet Sim", t_initial "Tinitial",
 * dt = dTtfd ("LB Steps Per T"), and the BiCGStab tolerances: converged at
 * ||r|| <= rel_tol ||b|| or ||r|| <= abs_tol ||r0||. */
typedef struct kofm_thermal_params {
  double alpha_fluid, alpha_foul, pr_turb, t_in, t_initial, dt;
  double rel_tol, abs_tol;
  int32_t max_iters;
} kofm_thermal_params;

/* The Nusselt-number output (kofm_nusselt_init): height = the rows the bulk
 * temperature sums (FULLSIZEY = 2 * (int)"Pipe Radius"); the wavy section
 * ("Wavy Start Loc", "Period Length", "Num Periods"); the conductivities of
 * soot and air in lattice units (KSOOT, KAIR = k * DELTA_F / DELTA_T,
 * clVariablesTR.cpp:1912-1913). */
typedef struct kofm_nusselt_params {
  int32_t height, wavy_start, wavy_period, wavy_periods;
  double k_soot, k_air;
} kofm_nusselt_params;

/* TR_shear_removal's constants: particleProperties' per-size arrays
 * (host pointers; particleProperties.cpp:264-330) and the tracer's defines
 * (clVariablesTR.cpp:1900-1942). */
#define KOFM_MAX_PAR_SIZES 32
typedef struct kofm_removal_params {
  int32_t n_sizes;          /* NUM_PAR_SIZES: 1 to KOFM_MAX_PAR_SIZES, at most nBL */
  const double* tau_crit;   /* 2 * n_sizes: paramTauCrit, per size clean then fouled surface */
  const double* drag_coeff; /* n_sizes: paramDcoeff */
  const double* lift_coeff; /* n_sizes: paramLcoeff */
  const double* mass;       /* n_sizes: paramMp */
  double dt_wall;           /* DTTR_WALL: "LB Steps Per TR Wall" */
  double foul_size_switch;  /* FOUL_SIZE_SWITCH_SOOT2: "Foul Size Switch Soot" */
  int32_t loc_stride;       /* FULLSIZEX_TR_PADDED: the particle location's row length */
  int32_t timer_start;      /* PAR_TIMER_START: min("Time Before Re-release", INT_MAX - 1) */
} kofm_removal_params;

/* The tracer's particle arrays (vtr.P), device pointers on the context's
 * GPU. */
typedef struct kofm_particles {
  double* pos;              /* 2 per particle, x y (16-byte aligned, as the double2 P_pos) */
  const uint32_t* num_rep;  /* P_Num_rep */
  const int16_t* type;      /* P_type: the size index */
  int32_t* dep_flag;        /* P_Dep_Flag: >= 0 the link a particle sits on; -1 free; -2 gone */
  uint16_t* dep_timer;      /* P_Dep_timer */
  int32_t* loc;             /* P_loc */
  int32_t* timer;           /* P_timer */
} kofm_particles;

/* --- lifecycle ------------------------------------------------------------ */

/* Allocates device storage for an nX x nY lattice on the current HIP device.
 * Replaces clVariablesLB::allocateArrays/allocateBuffers
 * (clVariablesLB.cpp:47-73) and kOmega::allocateArrays (kOmega.cpp:76-108).
 * One population array (9 planes of XsizeFull * nY plus padding) must stay
 * below 4 GiB (about 119M nodes in FP32, 59M in FP64); larger lattices are
 * split over GPUs.  KOFM_ERR_ARG otherwise. */
int kofm_create(const kofm_grid* grid, kofm_ctx** out);

/* A context without device state, for geometry preparation on machines
 * without a GPU: kofm_build_geometry, kofm_geometry_counts,
 * kofm_geometry_array and kofm_set_params work; every other call returns
 * KOFM_ERR_STATE.  No counterpart in the reference, whose geometry build
 * needs the OpenCL context (clVariablesLS::ini runs after clEnv::iniOpenCL). */
int kofm_create_host(const kofm_grid* grid, kofm_ctx** out);

/* Releases the context and all its device memory. */
void kofm_destroy(kofm_ctx* ctx);

/* Last error text of this context (or, when ctx is NULL, of the last failed
 * kofm_create or kofm_reduce_device on this thread). */
const char* kofm_last_error(const kofm_ctx* ctx);

/* Relaxation, forcing and solver constants.  Replaces
 * clVariablesLB::loadParams/setSourceDefines (clVariablesLB.cpp:553-639,
 * 860-877), kOmega::loadParams/setSourceDefines (kOmega.cpp:352-378,
 * 660-680) and the MU_NUMBER define (clVariablesTR.cpp:1908). */
int kofm_set_params(kofm_ctx* ctx, const kofm_params* p);

/* --- geometry ------------------------------------------------------------- */

/* Builds the node-type map, wall distances dXArr, interpolated
 * bounce-back link list, wall-node list, lsMap, WallD and the wall-shear
 * coefficients from the wall polyline (load/lsc0.txt, nN points, x y
 * interleaved) and its links (load/lsbl.txt, nBL pairs), and uploads them.
 * Replaces clVariablesLS::ini (clVariablesLS.cpp:569-620),
 * clVariablesLS::iniIBBArrays (:1085-1164, lsKernels.cl:429-505),
 * kOmega::iniWallD (kOmega.cpp:330-349), clVariablesTR::iniBLandNodI
 * (clVariablesTR.cpp:476-658) and shearStress::iniShearCoeffs
 * (shearStress.cpp:135-139, trKernelsShear.cl:332-496). */
int kofm_build_geometry(kofm_ctx* ctx, const double* C0, int32_t nN, const int32_t* BL,
                        int32_t nBL, const kofm_geom_params* gp);

int kofm_geometry_counts(const kofm_ctx* ctx, kofm_geom_counts* out);

/* Moves the walls to the points C (nN points, x y interleaved, the same
 * points and links as kofm_build_geometry; a fouling layer grows between
 * the built walls and C).  The fouling-layer update of clVariablesFL::update
 * (clVariablesFL.cpp:745-780): clVariablesLS::update (clVariablesLS.cpp:
 * 1020-1083) with updateNType (fluid nodes between the built and the current
 * wall become solid, their rho 1, u, k, omega and nu_t 0), updateNodeBoundaryInfo
 * and updateBoundaryNodes (links and wall spacings, lsKernels.cl:47-634),
 * kOmega::updateWallD (kOmegaKernels.cl:813-840), and the wall-node list and
 * shear coefficients of shearStress (shearStress.cpp:135-139).  Node types,
 * links, spacings and wall distances are computed on the device (each node
 * finds its own links: the result does not depend on scheduling); the
 * populations keep their values.  kofm_geometry_array returns the updated
 * arrays; kofm_objective_run needs a rebuilt geometry afterwards. */
int kofm_update_walls(kofm_ctx* ctx, const double* C, int32_t nN);

/* Copies one host-side geometry array (KOFM_GEOM_*) into out. */
int kofm_geometry_array(const kofm_ctx* ctx, int32_t which, void* out);

/* --- state ---------------------------------------------------------------- */

/* Populations set to the product-form equilibrium of (rho, ux, uy) on
 * non-solid nodes (clVariablesLB::iniDistsFromVel, clVariablesLB.cpp:367-405,
 * feq of :77-117); NULL arrays give the rest state at rho0
 * (clVariablesLB::iniDists, :407-425).  Macro arrays are set as the
 * reference sets them (rho0 on fluid nodes, zero velocity) and the strain
 * tensor to zero. */
int kofm_init_state(kofm_ctx* ctx, const double* rho, const double* ux, const double* uy);

/* The initial velocity field of clVariablesLB::iniDists(Umax)
 * (clVariablesLB.cpp:427-496), "Ini Poiseuille": walking up each column from
 * its first fluid node, whose height above the wall below is its S-plane
 * wall spacing, one node at a time, u_x is the parabola
 * 1.5 Re nu / R (1 - (y - R)^2 / R^2) or, with turbulent_profile ("Ini Turb
 * Velocity"), the law of the wall u_tau u+(y+) with u_tau = sqrt(3 Re) nu / R;
 * with perturb ("Perturb Vel Field") streaks and wall-normal waves whose
 * amplitudes are Box-Muller draws (HelperFuncs.cpp:107-126) from the C
 * library generator seeded by srand(seed) (the reference calls srand(1),
 * clProblem.cpp:80).  rho is set to rho.  Host only (works on a
 * kofm_create_host context with geometry); the arrays [D] then go to
 * kofm_init_state.  Nodes outside the walked columns are left unchanged. */
typedef struct kofm_ini_flow {
  double Re;                 /* "Re" */
  double nu;                 /* MuVal */
  double pipe_radius;        /* "Pipe Radius" */
  double rho;                /* "Rho" */
  int32_t turbulent_profile; /* "Ini Turb Velocity" */
  int32_t perturb;           /* "Perturb Vel Field" */
  double duplus_mult;        /* "DUplus Perturb Multiplier" */
  double epsilon_mult;       /* "Epsilon Perturb Multiplier" */
  double beta_mult;          /* "Beta Perturb Multiplier" */
  double alpha_mult;         /* "Alpha Perturb Multiplier" */
  double sigma_mult;         /* "Sigma Perturb Multiplier" */
  uint32_t seed;             /* srand seed */
} kofm_ini_flow;
int kofm_ini_poiseuille(const kofm_ctx* ctx, const kofm_ini_flow* p, double* rho, double* ux,
                        double* uy);

/* k, omega and nu_t = k/omega on fluid nodes (kOmega::setInitialValues,
 * kOmega.cpp:530-570, fillByNodeType Array.h:903-918).  Values <= 0 select
 * the reference's turbulence-intensity estimate from the mean velocity
 * sum(u_x)/sum(rho), or, when that is zero, umean = Re nu / R / 4 *
 * geom_penalty_factor ("Geometric Penalty Factor", kOmega.cpp:541). */
int kofm_init_komega(kofm_ctx* ctx, double k_ini, double omega_ini, double Re,
                     double geom_penalty_factor);

/* Uploads populations (9*D, the reference FA layout as loaded from
 * load/lbf.bin, clVariablesLB.cpp:883-902) and, when non-NULL, k, omega
 * (load/lbkappa.bin, load/lbomega.bin, kOmega.cpp:685-710), nu_t and S (3*D).
 * Macro fields are rebuilt as Calc_Ro_J does on restart (lbKernels.cl:137-164). */
int kofm_upload_state(kofm_ctx* ctx, const double* f, const double* k, const double* omega,
                      const double* nut, const double* S);

/* --- time stepping -------------------------------------------------------- */

/* nsteps reference time steps (clVariablesLB::Solve, clVariablesLB.h:375-441:
 * collide-stream with bounce-back, interpolated bounce-back, and with k-omega
 * on the diffusivity/coefficient kernels and implicit omega and k solves of
 * kOmega::Solve, kOmega.cpp:31-73).  options: KOFM_OPTION_SAVE_MACRO_FIELDS
 * writes rho, u_x + F/2 and u_y every step (always on with k-omega). */
int kofm_step(kofm_ctx* ctx, int32_t nsteps, uint32_t options);

/* Runs nsteps as kofm_step and returns the mean device time per step in
 * *ms_per_step, measured with HIP events on the context stream around the
 * whole run.  With ms_per_kernel non-NULL, events also bracket every
 * collide-stream launch and *ms_per_kernel receives that kernel's mean time
 * per step (-1 for captured k-omega steps, which are timed whole); those
 * inner events then add their own cost to *ms_per_step, so time the step
 * and the kernel in separate calls. */
int kofm_time_steps(kofm_ctx* ctx, int32_t nsteps, uint32_t options, double* ms_per_step,
                    double* ms_per_kernel);

/* Wall shear at wall nodes and boundary links from the newest populations
 * and the stored macros (shearStress::calculateShear, shearStress.cpp:70-83:
 * TR_shear_1 and TR_shear_2, trKernelsShear.cl:174-316). */
int kofm_wall_shear(kofm_ctx* ctx);

int kofm_get_solver_stats(kofm_ctx* ctx, kofm_solver_stats* out);

/* The k-omega time step (k-omega contexts).  Each kofm_step runs
 * kOmega::Solve's sub-steps (kOmega.cpp:31-73): after 20 steps without a
 * change the time step doubles back towards its initial value, and a
 * non-finite residual in either solve restores k and omega to the start of
 * the step, halves the time step, doubles the sub-steps per LB step and
 * starts the sub-steps again (GMRESSolver::checkConvergence,
 * GMRESSolver.cpp:1535-1546).  The actions:
 *   KOFM_DT_QUERY       report only;
 *   KOFM_DT_HALVE       kOmega::halveTimeStep (kOmega.cpp:387-399): k and
 *                       omega back to the start of the last k-omega step
 *                       (when one has run), half the time step, twice the
 *                       sub-steps;
 *   KOFM_DT_DOUBLE      kOmega::doubleTimeStep (kOmega.cpp:401-414);
 *   KOFM_DT_INJECT_NAN  the next solve is taken to have a non-finite
 *                       residual (tests the restart path without one; no
 *                       reference counterpart).
 * *steps_per_lb and *time_step (either may be NULL) receive the state after
 * the action. */
#define KOFM_DT_QUERY 0
#define KOFM_DT_HALVE 1
#define KOFM_DT_DOUBLE 2
#define KOFM_DT_INJECT_NAN 3
int kofm_komega_time_step(kofm_ctx* ctx, int32_t action, int32_t* steps_per_lb,
                          double* time_step);

/* --- output --------------------------------------------------------------- */

/* Sum/min/max of one field over the zero-padded power-of-two domain,
 * accumulated in FP64 (Reducer::reduceSingle, Reducer.h:240-273). */
int kofm_reduce(kofm_ctx* ctx, int32_t field, int32_t op, double* out);

/* The reduction of kofm_reduce on n values at a device address, enqueued on
 * `stream` (a hipStream_t; NULL for the null stream) with nothing allocated
 * and nothing synchronised: the result goes to the device double *out.
 * elem_bytes is 4 (FP32 values) or 8 (FP64); n lies in
 * [1, KOFM_REDUCE_DEVICE_MAX_N].  The per-work-group partials live in a
 * fixed device array of KOFM_REDUCE_DEVICE_PARTIALS FP64 values (8 KiB per
 * GPU, held until the process ends); the work-group count is capped at that
 * number, so any n in range fits.  All calls on one GPU within a process
 * share that array and it keeps the last call's partials: calls that overlap
 * on one GPU (on different streams or from different threads) race and give
 * wrong results.  For a given n the result has the same bits as kofm_reduce
 * over the same values.  Errors are reported by kofm_last_error(NULL).
 * No counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops). */
#define KOFM_REDUCE_DEVICE_PARTIALS 1024
#define KOFM_REDUCE_DEVICE_MAX_N (((int64_t)1) << 40)
int kofm_reduce_device(const void* field, int32_t elem_bytes, int64_t n, int32_t op, double* out,
                       void* stream);

/* clVariablesFD::ini (clVariablesFD.cpp:230-283) for k-omega contexts with
 * a built geometry: T = t_initial on the nodes fluid in the clean geometry
 * (fouling layer included), 0 on the clean walls; stepsPerLB =
 * (int)(1 / dt).  Building the geometry again turns the temperature off
 * until the next call.  KOFM_ERR_ARG without k-omega (the rows take nu_t,
 * Update_T_Coeffs_Implicit tfdKernels.cl:122-280). */
int kofm_thermal_init(kofm_ctx* ctx, const kofm_thermal_params* params);

/* n calls of clVariablesFD::Solve (clVariablesFD.h:343-358) with the
 * current u and nu_t: per call, the time step control (doubling back after
 * 20 calls without a change), then stepsPerLB sub-steps of the implicit
 * temperature rows (Update_T_Coeffs_Implicit: rows on the nodes fluid in
 * the clean geometry, the fouling layer with its own diffusivity, inlet at
 * t_in, zero-gradient outlet) solved by BiCGStab; a non-finite or jumping
 * residual restores T, halves the time step and restarts the sub-steps
 * (halveTimeStep, clVariablesFD.cpp:202-213).  The reference calls it once
 * per "LB Steps Per T" LB steps, after the k-omega solve (LBLS.cpp:197,
 * 230, 315). */
int kofm_thermal_step(kofm_ctx* ctx, int32_t n);

/* KOFM_DT_* action on the temperature time step (clVariablesFD::
 * halveTimeStep / doubleTimeStep); reports the sub-steps and time step. */
int kofm_thermal_time_step(kofm_ctx* ctx, int32_t action, int32_t* steps_per_lb, double* time_step);

/* Replaces T (D doubles, reference layout), e.g. from load/temp.bin on a
 * restart (clVariablesFD.cpp:797-806). */
int kofm_set_temperature(kofm_ctx* ctx, const double* T);

/* clVariablesFD::iniNuCoeffs (clVariablesFD.cpp:362-442), after
 * kofm_thermal_init: the Nusselt number over the columns of the particle
 * domain of kofm_build_geometry's parameters, TR_X_IND_START =
 * floor(x_min_val) to TR_X_IND_STOP = floor(x_stop_pos) (clVariablesTR.cpp:
 * 1373), which must lie in [1, nX - 1].  Each column searches the boundary
 * links whose box reaches it: the first min(MAX_BL_BOT - MIN_BL_BOT rounded
 * up to 256, columns) links of each wall from MIN_BL_BOT / MIN_BL_TOP
 * (clVariablesTR.cpp:497-530, updateNuCoeff1's launch, clVariablesFD.cpp:
 * 164-165).  The wavy section is required, as in the reference (:364-366).
 * kofm_update_walls keeps it set up; building the geometry again turns it
 * off.  KOFM_ERR_ARG for parameters outside these ranges. */
int kofm_nusselt_init(kofm_ctx* ctx, const kofm_nusselt_params* params);

/* The sizes of kofm_nusselt's outputs: columns (TR_X_IND_STOP -
 * TR_X_IND_START) and period boundaries (Num Periods + 1). */
int kofm_nusselt_sizes(kofm_ctx* ctx, int32_t* columns, int32_t* locations);

/* One Nusselt record (clVariablesFD::updateTimeData, clVariablesFD.cpp:
 * 838-869) from the current T and the last saved velocity, all in FP64:
 * the nearest node of each column to each wall along the wall normal
 * (updateNuCoeff1/2, tfdKernels.cl:759-856; ties go to the link earlier in
 * the particle tracer's order, then the lower row), then FD_calc_Nu (:919-982) into nu[5 * columns]: per
 * column the bulk temperature (getTbm, :858-876, whose rows x + j * nY are
 * kept), then the wall-normal gradient and the wall temperature (getTandDT,
 * :878-917) of the bottom and of the top wall, -1000 where a wall has no
 * node; and getNuMean (:984-1033, the OpenCL 1.2 form) into
 * mean[2 * locations]: sum |u| T and sum |u| over the column at each
 * period boundary.  When non-NULL, nodes[2 * columns] receives each slot's
 * (row << 2) | type (bottom wall first; row -1 when none) and
 * vec[4 * columns] its vector from the wall (x y interleaved). */
int kofm_nusselt(kofm_ctx* ctx, double* nu, double* mean, int32_t* nodes, double* vec);

/* --- particle re-entrainment by the wall shear ------------------------------ */

/* TR_shear_removal (trKernelsShear.cl:13-121) on particles [offset, offset +
 * count) of the tracer's arrays, as the reference calls it after TR_shear_2
 * at the start of a tracer step (LBLS.cpp:236-246, shearStress.cpp:327-335):
 * a particle sitting on link b (dep_flag = b, the tracer's link index)
 * counts its deposit timer down; at zero it deposits for good
 * (bl_dep[b + n_sizes * size] += num_rep, the reference's index; flag and
 * location -2). Otherwise, when |tau_b| exceeds tau_crit of its size and the
 * link's surface type, it re-enters the flow: drag along the link and lift
 * L sqrt(tau) along its normal, moved tau / m * dt^2 / 2 from the link's
 * midpoint 0.001 off the wall, with flag -1, the new location (-2 outside
 * [X Release Pos - 0.5, X Stop Pos)) and timer timer_start. tau_b is the
 * link shear of the last kofm_wall_shear (0 on links outside its range);
 * the link's point, normal, length and surface type are the tracer's BL
 * arrays for the walls now (updateBoundaryNodes, lsKernels.cl:246-258;
 * kofm_removal_links). bl_dep is a device array of nBL * n_sizes counters.
 * Enqueued on the context's stream (kofm_synchronize waits for it); a
 * particle whose link or size index is out of range is left unchanged.
 * KOFM_ERR_ARG for missing arrays or constants. */
int kofm_shear_removal(kofm_ctx* ctx, const kofm_removal_params* params,
                       const kofm_particles* particles, int64_t offset, int64_t count,
                       uint32_t* bl_dep);

/* The tracer's link arrays kofm_shear_removal uses, for the walls now and
 * the surface-type threshold foul_size_switch, in the tracer's link order
 * (its top wall runs left to right, iniBLinksTop, clVariablesTR.cpp:
 * 600-603): first point p0[2 * nBL], unit normal vN[2 * nBL], length
 * len[nBL], surface type surf[nBL]; any may be NULL.  Host arrays. */
int kofm_removal_links(kofm_ctx* ctx, double foul_size_switch, double* p0, double* vN, double* len,
                       int16_t* surf);

/* Copies one field (KOFM_FIELD_*) to host FP64 in the reference layout. */
int kofm_download(kofm_ctx* ctx, int32_t field, double* host);

/* Writes lbf.bin / lbkappa.bin / lbomega.bin (raw FP64, Array.h:1178-1194,
 * clVariablesLB.cpp:752-765, GMRESSolver::saveCheckPoint) and the save2file
 * text dumps lbux.txt, lbuy.txt, lbro.txt (+ lbkappa.txt, lbomega.txt) in
 * the reference format (Array.h:2528-2546: one row per x, tab-separated
 * values per y, 6 significant digits) into dir. */
int kofm_save2file(kofm_ctx* ctx, const char* dir);

/* saveRestartFiles of clVariablesLB, kOmega and clVariablesLS
 * (clVariablesLB.cpp:752-765, kOmega.cpp:492-496, clVariablesLS.cpp:534-537):
 * lbf.bin, lbkappa.bin and lbomega.bin (k-omega contexts) and lsc.bin (the
 * wall points of the current geometry, x y interleaved, 2*nN doubles), raw
 * FP64 with no header (Array.h:1178-1194), into dir. */
int kofm_save_restart(kofm_ctx* ctx, const char* dir);

/* The debug dumps of clVariablesLS::saveDebug (clVariablesLS.cpp:545-556)
 * and kOmega::saveDebug (kOmega.cpp:422-480) in the save2file text layout:
 * nType.txt (integers), dXarr_0.txt .. dXarr_3.txt (E, W, N, S wall
 * spacings, one file per plane, Array.h:3193-3220) and, with k-omega,
 * lbnut.txt and wallD.txt, into dir. */
int kofm_save_debug(kofm_ctx* ctx, const char* dir);

/* --- launch shape ---------------------------------------------------------- */

/* Launch shape of the plain collide-stream step: work-group width (64, 128,
 * 256, 512 or 1024), lattice nodes per lane (1, or 2 neighbouring nodes when
 * nX is even), steps per pass (1, or 2 for the temporally blocked kernels:
 * k-omega off, nodes_per_lane 1, and block_x is then the tile's node count,
 * 512 (64 x 8), 832 (64 x 13), 640 (128 x 5) or 1856 (64 x 29) with one
 * node per lane and step, or 1024 (64 x 16), 1664 (64 x 26), 1536
 * (128 x 12) or, with 16- or 32-bit storage, 3584 (128 x 28) with two
 * neighbouring nodes per lane; the tile width must divide nX) and the
 * padding (elements,
 * multiple of 64) between the nine population planes in HBM.  The state is preserved and
 * every shape gives bit-identical results.  No counterpart in the reference,
 * whose work-group is fixed at 128 x 1 (ConstDef.h:70-79). */
int kofm_set_launch(kofm_ctx* ctx, int32_t block_x, int32_t nodes_per_lane,
                    int32_t steps_per_pass, int64_t plane_pad);

/* Times the collide-stream kernels over the candidate shapes (about `steps`
 * steps each, on scratch populations at rest) and keeps the fastest; returns
 * the choice and its time per step.  No counterpart in the reference. */
int kofm_autotune(kofm_ctx* ctx, int32_t steps, int32_t* block_x, int32_t* nodes_per_lane,
                  int32_t* steps_per_pass, int64_t* plane_pad, double* ms_per_step);

/* With on != 0 (the default) kofm_step replays a captured HIP graph of 32
 * collide-stream launches for each whole block of 32 plain (k-omega off)
 * steps; off launches every step separately.  Results are identical.  No
 * counterpart in the reference. */
int kofm_set_graphs(kofm_ctx* ctx, int32_t on);

/* Blocks until all work on the context stream is done. */
int kofm_synchronize(kofm_ctx* ctx);

/* --- sensitivity to the wall shape ----------------------------------------- */

/* J = mean over the last `window` steps of a run of
 *       pbar(x_in) - pbar(x_out) + shear_weight * sum_BL tau_BL * |BL|
 * with pbar(x) the mean of rho/3 over the fluid nodes of column x, tau_BL the
 * boundary-link wall shear (TR_shear_2, trKernelsShear.cl:300-316, from the
 * step's newest populations and its collision-time density and velocity)
 * and |BL| the link length.  The reference has no adjoint; this is the
 * objective of the wall-shape sensitivity. */
typedef struct kofm_objective {
  int32_t window;      /* 1 <= window <= nsteps */
  int32_t x_in;        /* pressure column 1 (0 for the channel inlet) */
  int32_t x_out;       /* pressure column 2 (nX-1 for the outlet) */
  int32_t flags;       /* KOFM_ADJOINT_* (0: the exact adjoint) */
  double shear_weight; /* weight of the wall-shear integral */
} kofm_objective;

/* With k-omega: kofm_adjoint holds the eddy viscosity of every step at its
 * recorded value (the collide-stream is differentiated with that per-node
 * relaxation frequency; the k-omega equations, the wall distances and the
 * spacings are not).  The reference's k-omega rows are not monotone next to
 * nodes very close to a curved wall: k and omega go negative, are clamped,
 * and the eddy viscosity sits at its 1e-16 floor, where the production term
 * (divided by the eddy viscosity) makes the exact tangent grow by orders of
 * magnitude per step.  This gives a bounded gradient there. */
#define KOFM_ADJOINT_FROZEN_TURBULENCE 1

/* Runs nsteps steps (with k-omega if the context has it) from the current
 * state, returns J and keeps the state every checkpoint_interval steps
 * (about sqrt(nsteps) keeps the checkpoints and the recomputed interval the
 * same size).  The state advances exactly as with kofm_step.  With k-omega
 * the traction uses each step's eddy viscosity, as kofm_wall_shear does. */
int kofm_objective_run(kofm_ctx* ctx, int32_t nsteps, int32_t checkpoint_interval,
                       const kofm_objective* obj, double* J);

/* Reverse sweep over the last kofm_objective_run: each checkpoint interval
 * is recomputed and the adjoint step runs backwards through it (with
 * k-omega: the transposed implicit solves and the adjoints of the row
 * assembly, the diffusivities and the k-omega collision; the transposed
 * solves use the context's relative tolerances).  dJ_dy[nN] receives dJ/dy
 * of every wall point of kofm_build_geometry (the x coordinates held fixed;
 * node types, link sets, and the wall links that the wall distances and
 * spacings were taken from, are held at the built geometry); dJ_dq[nIbb]
 * (may be NULL) receives the part through the boundary-link q of every link
 * in KOFM_GEOM_IBB order (0 for a link repeated later in the list). */
int kofm_adjoint(kofm_ctx* ctx, double* dJ_dy, double* dJ_dq);

/* dJ/dy of the wall points from given sensitivities of the k-omega wall
 * distance field (wall_d_bar[D], may be NULL) and of the axis wall spacings
 * (dx_bar[4*D], planes E, W, N, S, may be NULL), each differentiated through
 * the wall link (or link end point) the geometry build took it from; adds
 * into dJ_dy[nN].  Host only (works on a kofm_create_host context).  No
 * reference counterpart (the wall-field part of kofm_adjoint). */
int kofm_geometry_field_gradient(const kofm_ctx* ctx, const double* wall_d_bar,
                                 const double* dx_bar, double* dJ_dy);

/* --- ensembles ------------------------------------------------------------- */

/* Several independent channels (contexts with geometry and state, one
 * population storage type, k-omega all off or all on) stepped together: one
 * collide-stream launch per step covers every tile of a group of members;
 * without k-omega whole blocks of 32 steps replay a captured HIP graph, with
 * k-omega each step (the launch and every member's implicit solves) does (see
 * kofm_ensemble_set_cache_budget for the grouping).  Members keep their own arrays and are read
 * with kofm_download as usual; calls on a member are ordered after the
 * ensemble's steps.  Members may change between ensemble steps (new
 * parameters, geometry, launch shape or padding): the next step rebuilds the
 * ensemble's tables and graphs from them.  A member destroyed with
 * kofm_destroy makes every later kofm_ensemble_step fail with KOFM_ERR_STATE;
 * destroy the ensemble before or after its members.  The reference runs one
 * channel per process; this has no counterpart there. */
typedef struct kofm_ensemble kofm_ensemble;

int kofm_ensemble_create(kofm_ctx* const* members, int32_t n, kofm_ensemble** out);
int kofm_ensemble_step(kofm_ensemble* e, int32_t nsteps, uint32_t options);
/* Mean device time per ensemble step over nsteps (HIP events) and the total
 * number of fluid nodes updated per step. */
int kofm_ensemble_time_steps(kofm_ensemble* e, int32_t nsteps, uint32_t options,
                             double* ms_per_step, int64_t* fluid_nodes);
/* Work-group width of the ensemble launch (128, 256, 512 or 1024). */
int kofm_ensemble_set_block(kofm_ensemble* e, int32_t block_x);
/* Launch shape: steps_per_launch 1 with block_x 128, 256, 512 or 1024, or
 * 2 (the temporally blocked kernels) with block_x the tile's node count,
 * 832 (64 x 13), or with two nodes per lane 1664 (64 x 26), 1536 (128 x 12)
 * or, with 16- or 32-bit storage, 3584 (128 x 28); every member's nX a
 * multiple of the tile width.  Results are bit-identical for every shape. */
int kofm_ensemble_set_launch(kofm_ensemble* e, int32_t block_x, int32_t steps_per_launch);
/* Members are packed (largest first) into groups whose population arrays
 * (both buffers) total at most `bytes`; kofm_ensemble_step advances each
 * group through all its steps before the next, so a group that fits the
 * last-level cache (256 MB on MI355X) runs from it.  0 puts every member in
 * one launch per step.  The default is 160 MiB.  *groups (may be NULL)
 * receives the number of groups.  Results do not depend on the grouping. */
int kofm_ensemble_set_cache_budget(kofm_ensemble* e, int64_t bytes, int32_t* groups);
/* Times `steps` ensemble steps for each launch shape (one step per launch
 * with 256 or 512 wide groups; with k-omega off also the two-step 832 and
 * 1664 tiles when every nX is a multiple of 64, and 1536 and 3584 when it
 * is a multiple of 128) and each cache budget (0, 96
 * and 160 MiB), keeps the fastest and reports it.  Every member's state
 * (populations, and with k-omega k, omega, strain, eddy viscosity and
 * macros) is restored after each trial, so the run continues as if no
 * tuning had happened. */
int kofm_ensemble_autotune(kofm_ensemble* e, int32_t steps, int32_t* block_x,
                           int32_t* steps_per_launch, int64_t* cache_budget, double* ms_per_step);
const char* kofm_ensemble_last_error(const kofm_ensemble* e);
void kofm_ensemble_destroy(kofm_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* KOFM_H */
