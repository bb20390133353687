/*
 * pincflow_hip.h -- C ABI of the MI355X-native semi-implicit dynamical-core
 * step (libpincflow_hip.so).
 *
 * The library owns one HBM-resident model state per handle (`pcf_state`).
 * Every entry point below mirrors a function of the PinCFlow.jl reference
 * (paths relative to the reference repository root) and keeps its argument
 * meaning and error behaviour:
 *
 *   return 0              success
 *   return PCF_E_NOCONV   BiCGSTAB did not converge (reference errflag = true)
 *   return < 0            the reference would have raised `error(...)`, or a
 *                         HIP call failed; pcf_last_error() has the message.
 *
 * All arrays handed across the ABI are plain FP64 host or device pointers in
 * the reference's Julia column-major layout (x fastest).  Haloed fields have
 * (x_size + 2 nbx) x (y_size + 2 nby) x (z_size + 2 nbz) elements, Poisson
 * vectors have x_size x y_size x z_size elements.
 *
 * One host thread per handle.  Each handle owns one HIP stream; no call
 * returns while work it enqueued could still read a caller-owned host buffer.
 */
#ifndef PINCFLOW_HIP_H
#define PINCFLOW_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCF_ABI_VERSION 8

/* ---- return codes ------------------------------------------------------ */
#define PCF_OK 0
#define PCF_E_NOCONV 1     /* apply_bicgstab!.jl:146-149 (errflag = true)   */
#define PCF_E_DT_ZERO (-2) /* solve_poisson!.jl:52-54                        */
#define PCF_E_DIVNORM (-3) /* compute_lhs!.jl:149-157                        */
#define PCF_E_DTMIN (-4)   /* compute_time_step.jl:142-150                   */
#define PCF_E_OPTION (-5)  /* @dispatch.jl: invalid option                   */
#define PCF_E_CONFIG (-6)  /* Domain.jl:146-174 / Grid.jl:221 / Atmosphere   */
#define PCF_E_HIP (-7)     /* a HIP runtime call failed                      */
#define PCF_E_ARG (-8)     /* bad argument (null pointer, unknown field ...)  */

/* ---- namelist options (Types/NamelistTypes) ---------------------------- */
enum pcf_model {
  PCF_BOUSSINESQ = 0,
  PCF_PSEUDO_INCOMPRESSIBLE = 1,
  PCF_COMPRESSIBLE = 2 /* P predicted (set_p.jl), pbar follows it */
};

enum pcf_background {
  PCF_NEUTRAL_STRATIFICATION = 0, /* Atmosphere.jl (Boussinesq)            */
  PCF_STABLE_STRATIFICATION = 1,  /* Atmosphere.jl (Boussinesq)            */
  PCF_ISOTHERMAL = 2,             /* Atmosphere.jl (pseudo-incompressible) */
  PCF_ISENTROPIC = 3,             /* Atmosphere.jl (pseudo-incompressible) */
  PCF_REALISTIC = 4,              /* Atmosphere.jl (pseudo-incompressible) */
  PCF_LAPSE_RATES = 5             /* Atmosphere.jl (pseudo-incompressible) */
};

/* The reference takes Julia closures for topography, stretching, initial
 * fields and sponges (GridNamelist.jl, AtmosphereNamelist.jl,
 * SpongeNamelist.jl).  Closures cannot run on the device, so the ABI takes a
 * closed set of parametric forms that covers the reference examples. */
enum pcf_topography {
  PCF_TOPO_FLAT = 0,
  PCF_TOPO_COSINE_HILL = 1, /* h0/2 (1 + cos(pi x / l0))  periodic_hill.jl:29 */
  PCF_TOPO_AGNESI = 2,      /* h0 / (1 + (x^2+y^2)/l0^2)  mountain_wave.jl:34 */
  PCF_TOPO_GAUSSIANS = 3    /* sum_n h_n exp(-((x-x_n)^2+(y-y_n)^2)/w_n^2)   */
};

enum pcf_stretching {
  PCF_STRETCH_NONE = 0, /* zhat -> zhat (GridNamelist.jl default)            */
  PCF_STRETCH_SINH = 1  /* zhat -> lz sinh(b zhat/lz) / sinh(b), b = param   */
};

enum pcf_buoyancy_initialization {
  PCF_INIT_RHOP = 0,  /* Predictands.jl: buoyancy_initialization = :initial_rhop   */
  PCF_INIT_THETAP = 1 /* Predictands.jl: buoyancy_initialization = :initial_thetap */
};

enum pcf_perturbation {
  PCF_PERT_NONE = 0,
  PCF_PERT_COSINE_BUBBLE = 1, /* A (1 + cos(pi r)) for r <= 1, cold_bubble.jl:21-28 */
  PCF_PERT_VORTEX = 2         /* vortex.jl:23-39 winds, amplitude A             */
};

enum pcf_lhs_sponge {
  PCF_LHS_SPONGE_NONE = 0,
  PCF_LHS_SPONGE_SIN2 = 1 /* mountain_wave.jl:40-53 */
};

enum pcf_rhs_sponge {
  PCF_RHS_SPONGE_NONE = 0,
  PCF_RHS_SPONGE_SIN2_TOP = 1 /* periodic_hill.jl:33-36: A sin^2(...)/dt */
};

/* Arithmetic of the Poisson solve (an addition of this build; the reference
 * is FP64 throughout). */
/* Initial guess of the mixed-precision solve (the FP64 solve starts from
 * x = 0 as apply_bicgstab!.jl:44 does).  PREVIOUS: the solution of the same
 * implicit stage in the previous time step, rescaled by the ratio of the
 * time steps.  EXTRAPOLATED: the same stage's solutions of the previous two
 * time steps (per unit time step) extrapolated linearly in time to this
 * one (PREVIOUS until two exist). */
enum pcf_initial_guess {
  PCF_GUESS_ZERO = 0,
  PCF_GUESS_PREVIOUS = 1,
  PCF_GUESS_LEARNED = 2, /* the attached network (pcf_set_guess_network) */
  PCF_GUESS_EXTRAPOLATED = 3
};

/* Preconditioner of the mixed-precision solve.  LINE_RELAXATION: the
 * reference's (apply_preconditioner!.jl:40-124), evaluated in FP32.
 * MULTIGRID: V-cycles over Galerkin 2 x 2 horizontal aggregations of the
 * operator with the reference's z-line pseudo-time step as the smoother on
 * every level (FP32). */
enum pcf_preconditioner_type {
  PCF_PC_LINE_RELAXATION = 0,
  PCF_PC_MULTIGRID = 1
};

/* Storage of the finest level's operator coefficients read by the
 * mixed-precision preconditioner (its horizontal steps and the multigrid
 * residual): FP32, or BF16 with FP32 arithmetic.  The line-solve factors stay
 * FP32. */
enum pcf_preconditioner_storage {
  PCF_STORAGE_FP32 = 0,
  PCF_STORAGE_BF16 = 1
};

/* Tracer transport (TracerNamelist.jl: tracer_setup :NoTracer / :TracerOn). */
enum pcf_tracer_setup { PCF_NO_TRACER = 0, PCF_TRACER_ON = 1 };

/* Parametric forms of TracerNamelist.initial_tracer (a closure in the
 * reference).  DIPOLE: vortex.jl:50-58, sign(x) (1 + cos(pi r)) / 2 for
 * r = sqrt(((|x| - rx) / rx)^2 + (y / ry)^2) <= 1, else 0. */
enum pcf_initial_tracer { PCF_TRACER_ZERO = 0, PCF_TRACER_DIPOLE = 1 };

enum pcf_precision {
  PCF_PRECISION_FP64 = 0, /* reference arithmetic                            */
  PCF_PRECISION_MIXED = 1 /* FP64 BiCGSTAB with the preconditioner in FP32   */
};

#define PCF_MAX_HILLS 32

typedef struct pcf_hill {
  double height; /* m */
  double x0, y0; /* m, centre */
  double width;  /* m, e-folding half-width */
} pcf_hill;

/* Mirror of Namelists(...) (Types/NamelistTypes/Namelists.jl).  Field order
 * is part of the ABI; pcf_config_default() fills the reference defaults. */
typedef struct pcf_config {
  int32_t abi_version;
  /* DomainNamelist (DomainNamelist.jl:72-102) */
  int32_t x_size, y_size, z_size;
  int32_t nbx, nby, nbz;
  double lx, ly, lz;
  /* DiscretizationNamelist (DiscretizationNamelist.jl:44-60) */
  double cfl_number, dtmin, dtmax;
  int32_t adaptive_time_step;
  /* PoissonNamelist (PoissonNamelist.jl:48-66) */
  int32_t poisson_iterations;
  double tolerance;
  double dtau;
  int32_t preconditioner;
  int32_t preconditioner_iterations;
  int32_t initial_cleaning;
  int32_t tolerance_is_relative;
  /* AtmosphereNamelist (AtmosphereNamelist.jl:108-156) */
  int32_t model;
  int32_t background;
  int32_t specify_reynolds_number;
  int32_t buoyancy_initialization;
  double inverse_reynolds_number;
  double kinematic_viscosity;
  double buoyancy_frequency;
  double potential_temperature;
  double temperature;
  double ground_pressure;
  double coriolis_frequency;
  double tropopause_height;
  double troposphere_lapse_rate;
  double stratosphere_lapse_rate;
  double initial_u, initial_v, initial_w; /* m/s, uniform parts          */
  int32_t perturbation;                   /* enum pcf_perturbation        */
  int32_t pad0;
  double pert_amplitude;                  /* kg/m^3 (rhop), K (thetap), m/s */
  double pert_x0, pert_y0, pert_z0;       /* m */
  double pert_rx, pert_ry, pert_rz;       /* m; <= 0 drops that direction */
  /* GridNamelist (GridNamelist.jl:35-45) */
  int32_t stretching;                     /* enum pcf_stretching          */
  int32_t topography;                     /* enum pcf_topography          */
  double stretching_parameter;
  double topo_height, topo_width;         /* m (cosine hill, Agnesi)      */
  int32_t n_hills;
  int32_t pad1;
  pcf_hill hills[PCF_MAX_HILLS];
  /* SpongeNamelist (SpongeNamelist.jl:48-66) */
  int32_t damp_horizontal_wind_on_rhs;
  int32_t relax_to_mean;
  int32_t lhs_sponge;                     /* enum pcf_lhs_sponge          */
  int32_t rhs_sponge;                     /* enum pcf_rhs_sponge          */
  double lhs_alpharmax;                   /* 1/s */
  double lhs_dxr, lhs_dyr, lhs_dzr;       /* m; <= 0 disables a direction */
  double lhs_divisor;
  double rhs_zr;                          /* m */
  double rhs_amplitude;
  double relaxed_u, relaxed_v, relaxed_w; /* m/s */
  /* OutputNamelist (OutputNamelist.jl) -- time-loop control only */
  int32_t output_steps;
  int32_t iterations;
  int32_t nout;
  int32_t precision;                      /* enum pcf_precision           */
  double output_interval;                 /* s */
  double tmax;                            /* s */
  /* Additions of this library (no reference counterpart), mixed precision
   * only: the initial guess and the preconditioner of the solve. */
  int32_t initial_guess;                  /* enum pcf_initial_guess       */
  int32_t preconditioner_type;            /* enum pcf_preconditioner_type */
  int32_t mg_smoothing_steps;             /* multigrid pre- and post-smoothing steps (2) */
  int32_t mg_coarse_steps;                /* multigrid steps on the coarsest level (20)  */
  int32_t preconditioner_storage;         /* enum pcf_preconditioner_storage            */
  /* TracerNamelist.jl (leading_order_impact needs MS-GWaM: not supported) */
  int32_t tracer_setup;                   /* enum pcf_tracer_setup        */
  int32_t initial_tracer;                 /* enum pcf_initial_tracer      */
  double tracer_rx, tracer_ry;            /* m, DIPOLE radii              */
  /* AtmosphereNamelist.thermal_conductivity (AtmosphereNamelist.jl:113),
   * m^2/s at the surface; heat conduction of the compressible model */
  double thermal_conductivity;
  /* Addition of this library: arithmetic of the explicit stage's fluxes.
   * PCF_PRECISION_FP64 (reference) or PCF_PRECISION_MIXED: the
   * reconstruction inputs and stress components stored in FP32 and the MUSCL
   * reconstruction, fluxes and flux divergences of pcf_step's RK stages
   * evaluated in FP32; the updates, the other fields and pcf_compute_fluxes
   * stay FP64. */
  int32_t flux_precision;                 /* enum pcf_precision           */
  int32_t pad2;
} pcf_config;

/* Fields reachable through pcf_get_field / pcf_set_field (haloed, FP64). */
enum pcf_field {
  PCF_F_RHO = 0, PCF_F_RHOP, PCF_F_U, PCF_F_V, PCF_F_W, PCF_F_PIP, /* Predictands */
  PCF_F_PBAR, PCF_F_THETABAR, PCF_F_RHOBAR, PCF_F_N2,               /* Atmosphere  */
  PCF_F_JAC, PCF_F_MET13, PCF_F_MET23, PCF_F_MET33,                 /* Grid        */
  PCF_F_ZC, PCF_F_ZCTILDE,
  PCF_F_ALPHAR, PCF_F_BETAR,                                        /* Sponge      */
  PCF_F_DPIP, PCF_F_CORX, PCF_F_CORY,                               /* corrector   */
  PCF_F_DRHO, PCF_F_DRHOP, PCF_F_DU, PCF_F_DV, PCF_F_DW,            /* Increments  */
  PCF_F_RHOOLD, PCF_F_RHOPOLD, PCF_F_UOLD, PCF_F_VOLD, PCF_F_WOLD,  /* Backups     */
  PCF_F_P0_RHO, PCF_F_P0_RHOP, PCF_F_P0_U, PCF_F_P0_V, PCF_F_P0_W,  /* p0 copy     */
  PCF_F_PHIRHO_X, PCF_F_PHIRHO_Y, PCF_F_PHIRHO_Z,                   /* Fluxes      */
  PCF_F_PHIRHOP_X, PCF_F_PHIRHOP_Y, PCF_F_PHIRHOP_Z,
  PCF_F_PHIU_X, PCF_F_PHIU_Y, PCF_F_PHIU_Z,
  PCF_F_PHIV_X, PCF_F_PHIV_Y, PCF_F_PHIV_Z,
  PCF_F_PHIW_X, PCF_F_PHIW_Y, PCF_F_PHIW_Z,
  /* tracer (TracerPredictands, TracerIncrements, TracerAuxiliaries, the
   * chi0 backup of integrate.jl:276 and TracerFluxes) */
  PCF_F_CHI, PCF_F_DCHI, PCF_F_INITIALTRACER, PCF_F_CHI0,
  PCF_F_PHICHI_X, PCF_F_PHICHI_Y, PCF_F_PHICHI_Z,
  /* compressible model: P (Predictands.jl:67), its increment, its p0 copy,
   * the P fluxes and the heat-conduction fluxes phitheta (Fluxes.jl:58-66) */
  PCF_F_P, PCF_F_DP, PCF_F_P0_P,
  PCF_F_PHIP_X, PCF_F_PHIP_Y, PCF_F_PHIP_Z,
  PCF_F_PHITHETA_X, PCF_F_PHITHETA_Y, PCF_F_PHITHETA_Z,
  PCF_F_COUNT
};

/* Poisson vectors (interior-sized, FP64) reachable through pcf_get_vector. */
enum pcf_vector {
  PCF_V_RHS = 0, PCF_V_SOLUTION,
  PCF_V_AC, PCF_V_AL, PCF_V_AR, PCF_V_AB, PCF_V_AF, PCF_V_AD, PCF_V_AU,
  PCF_V_ALD, PCF_V_ALU, PCF_V_ARD, PCF_V_ARU, PCF_V_ABD, PCF_V_ABU,
  PCF_V_AFD, PCF_V_AFU, PCF_V_ADD, PCF_V_AUU, PCF_V_ALDD, PCF_V_ALUU,
  PCF_V_ARDD, PCF_V_ARUU, PCF_V_ABDD, PCF_V_ABUU, PCF_V_AFDD, PCF_V_AFUU,
  PCF_V_COUNT
};

/* Boundary sets of set_boundaries!(state, ::Boundary...) (set_boundaries!.jl:43-68).
 * RECONSTRUCTIONS has nothing to fill here (pcf_reconstruct) and returns 0. */
enum pcf_boundaries {
  PCF_BOUNDARY_PREDICTANDS = 0,
  PCF_BOUNDARY_RECONSTRUCTIONS = 1,
  PCF_BOUNDARY_FLUXES = 2
};

/* Variables of update! / apply_lhs_sponge! (Types.jl singletons). */
enum pcf_variable { PCF_VAR_RHO = 0, PCF_VAR_RHOP, PCF_VAR_U, PCF_VAR_V, PCF_VAR_W, PCF_VAR_PIP, PCF_VAR_CHI, PCF_VAR_P };

/* Operator kinds of apply_operator! (apply_operator!.jl:48 / :275). */
enum pcf_operator { PCF_OP_TOTAL = 0, PCF_OP_HORIZONTAL = 1 };

typedef struct pcf_state pcf_state;

/* Per-step record of pcf_step (the values integrate prints). */
typedef struct pcf_step_info {
  double time;      /* s, after the step          integrate.jl:241        */
  double dt;        /* s                          integrate.jl:230        */
  int32_t niter[2]; /* BiCGSTAB iterations of the two implicit stages     */
  int32_t output;   /* 1 if this step lands on an output time integrate.jl:233-239 */
  int32_t pad;
} pcf_step_info;

/* ---- library ------------------------------------------------------------ */
int pcf_abi_version(void);
const char* pcf_last_error(void);
/* Fills the reference namelist defaults (each Namelist constructor's keyword defaults). */
void pcf_config_default(pcf_config* cfg);
/* sizeof(pcf_config), for bindings that mirror the structure. */
int pcf_config_size(void);

/* ---- state (State.jl:106-131) ------------------------------------------- */
/* Builds the grid, background, sponge arrays and initial predictands on the
 * host in FP64 and uploads them; all later work stays on the device. */
int pcf_create(const pcf_config* cfg, pcf_state** out);
int pcf_destroy(pcf_state* st);
/* Interior and haloed extents: dims[0..5] = nx, ny, nz, nxx, nyy, nzz. */
int pcf_dims(const pcf_state* st, int32_t dims[6]);
/* Model time in seconds and the number of steps taken (integrate.jl:217-241). */
int pcf_get_time(const pcf_state* st, double* time_s, int32_t* itime);
/* Restart support (integrate.jl:168-196): set the model time to time_s
 * seconds and the next output time to time_s + output_interval.  Fails with
 * "Restart error: tmax < time!" (PCF_E_CONFIG) past tmax. */
int pcf_set_time(pcf_state* st, double time_s);

/* ---- time integration (Integration/) ------------------------------------ */
/* Initial divergence cleaning, integrate.jl:136-156. */
int pcf_initial_cleaning(pcf_state* st, int32_t* niter);
/* One iteration of the time loop of integrate (integrate.jl:217-378):
 * compute_time_step, output-time correction, compute_sponges!, the five
 * semi-implicit stages.  Returns PCF_E_NOCONV / PCF_E_DTMIN like the reference. */
int pcf_step(pcf_state* st, pcf_step_info* info);
/* compute_time_step (compute_time_step.jl:46-153); dt is non-dimensional. */
int pcf_compute_time_step(pcf_state* st, double* dt);
/* compute_sponges! (compute_sponges!.jl:26-65), dt and time non-dimensional. */
int pcf_compute_sponges(pcf_state* st, double dt, double time);
/* explicit_integration!(state, p0, dt, time, LHS()) (explicit_integration!.jl:42-101)
 * with p0 = the device copy made by pcf_backup_predictands. */
int pcf_explicit_integration_lhs(pcf_state* st, double dt, double time);
/* explicit_integration!(state, p0, dt, time, RHS()) (explicit_integration!.jl:103-131). */
int pcf_explicit_integration_rhs(pcf_state* st, double dt, double time);
/* implicit_integration! (implicit_integration!.jl:37-79); *niter gets the iterations. */
int pcf_implicit_integration(pcf_state* st, double dt, double time, double rayleigh_factor,
                             int32_t* niter);
/* backup_predictands (backup_predictands.jl): p0 := predictands. */
int pcf_backup_predictands(pcf_state* st);

/* ---- boundaries (Boundaries/) ------------------------------------------- */
int pcf_set_boundaries(pcf_state* st, int which);

/* ---- advection (FluxCalculator/) ---------------------------------------- */
/* reconstruct!(state) (reconstruct!.jl:102-315): the reconstruction inputs
 * (rho / P, rho' / P, the edge-weighted momenta over P, the Cartesian
 * vertical wind) from the current predictands.  The MUSCL limiter
 * (apply_3d_muscl!.jl:35-75) is evaluated from them inside the flux kernels,
 * so the reconstructed face values are never stored and
 * set_boundaries!(state, BoundaryReconstructions()) has nothing to fill. */
int pcf_reconstruct(pcf_state* st);
/* compute_fluxes!(state, p0) (compute_fluxes!.jl:326-1503) followed by
 * set_boundaries!(state, BoundaryFluxes()) (explicit_integration!.jl:53-58),
 * with p0 = the copy made by pcf_backup_predictands.  It evaluates the
 * reconstruction itself first, so the reference's sequence pcf_reconstruct,
 * pcf_set_boundaries(PCF_BOUNDARY_RECONSTRUCTIONS), pcf_compute_fluxes and
 * this call alone give the same fluxes. */
int pcf_compute_fluxes(pcf_state* st);
/* compute_fluxes!(state, p0, Theta()) (compute_fluxes!.jl:1592-1756): the
 * heat-conduction fluxes of the compressible model from the p0 copy; nothing
 * in the other models or with thermal_conductivity = 0. */
int pcf_compute_theta_fluxes(pcf_state* st);
/* modify_compressible_wind!(state, *) (multiply = 1) or (state, /) (0),
 * modify_compressible_wind!.jl:52-80; nothing outside the compressible model. */
int pcf_modify_compressible_wind(pcf_state* st, int multiply);
/* synchronize_compressible_atmosphere!(state, state.variables.predictands)
 * (synchronize_compressible_atmosphere!.jl:61-72): pbar := P. */
int pcf_synchronize_compressible_atmosphere(pcf_state* st);

/* ---- updates (Update/) --------------------------------------------------- */
/* update!(state, dt, m, var[, LHS()]) (update!.jl:505-1192), m in 1..3. */
int pcf_update_lhs(pcf_state* st, double dt, int32_t m, int var);
/* update!(state, dt, var, RHS(), Explicit() | Implicit(), rf)
 * (update!.jl:600-727, 809-881, 960-1032, 1195-1342). */
int pcf_update_rhs(pcf_state* st, double dt, int var, int implicit, double rayleigh_factor);
/* apply_lhs_sponge! (apply_lhs_sponge!.jl:255-636). */
int pcf_apply_lhs_sponge(pcf_state* st, double dt, double time, int var);

/* ---- pressure solve (PoissonSolver/) ------------------------------------ */
/* apply_corrector! (apply_corrector!.jl:38-66). */
int pcf_apply_corrector(pcf_state* st, double dt, double rayleigh_factor, int32_t* niter);
/* compute_lhs! (compute_lhs!.jl:68-160): fills the RHS, returns tolref. */
int pcf_compute_lhs(pcf_state* st, double* tolref);
/* compute_operator! (compute_operator!.jl:40-1606). */
int pcf_compute_operator(pcf_state* st, double dt, double rayleigh_factor);
/* apply_operator!(sin, ls, Total()|Horizontal(), state) (apply_operator!.jl:48-479);
 * d_in/d_out are device pointers to interior-sized FP64 vectors. */
int pcf_apply_operator(pcf_state* st, const double* d_in, double* d_out, int kind);
/* d_out = A^T d_in for the total operator of the last compute_operator
 * (no reference counterpart: the gradient of the residual loss that trains
 * the initial-guess network).  Device pointers, interior-sized FP64. */
int pcf_apply_operator_transpose(pcf_state* st, const double* d_in, double* d_out);
/* apply_preconditioner!(sin, sout, state) (apply_preconditioner!.jl:40-124);
 * device pointers, interior-sized. */
int pcf_apply_preconditioner(pcf_state* st, const double* d_in, double* d_out);
/* apply_bicgstab!(state, tolref) (apply_bicgstab!.jl:24-150) on the current
 * RHS and operator; solution left in the PCF_V_SOLUTION vector. */
int pcf_apply_bicgstab(pcf_state* st, double tolref, int32_t* niter);
/* compute_global_dot_product (compute_global_dot_product.jl:20-36), device pointers. */
int pcf_dot(pcf_state* st, const double* d_a, const double* d_b, double* result);
/* pcf_dot over n elements without a state, allocating nothing: no reference
 * counterpart, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The result (the bits pcf_dot gives
 * for the same n) is written to the device pointer d_result on `stream` (a
 * hipStream_t; null = the default stream); nothing is synchronized.  The
 * block partials live in a 16 KiB array the library holds on each device
 * until the process ends; calls that overlap on one device race on it.
 * 0 <= n <= PCF_DOT_NOALLOC_MAX_N, else PCF_E_ARG with a message. */
#define PCF_DOT_NOALLOC_MAX_N ((int64_t)1 << 34)
int pcf_dot_noalloc(const double* d_a, const double* d_b, int64_t n, double* d_result, void* stream);

/* ---- host <-> device copies (test and output support) ------------------- */
int pcf_get_field(pcf_state* st, int field, double* host_out);
int pcf_set_field(pcf_state* st, int field, const double* host_in);
int pcf_get_vector(pcf_state* st, int vec, double* host_out);
int pcf_set_vector(pcf_state* st, int vec, const double* host_in);
/* Residual record of the last BiCGSTAB solve: final ||r||/||b|| (RMS norms). */
int pcf_last_residual(const pcf_state* st, double* rel_residual);


/* ---- kernel timing (instrumentation of this build; the reference has no
 * counterpart: it prints wall-clock times only) ---------------------------- */
enum pcf_kernel_group {
  PCF_K_APPLY_TOTAL_DOT = 0, /* 25-point apply with a fused dot product     */
  PCF_K_PC_THOMAS,           /* preconditioner line solves in z            */
  PCF_K_PC_RHS,              /* preconditioner horizontal apply + update   */
  PCF_K_COMPUTE_OPERATOR,    /* assembly of the 25 coefficients           */
  PCF_K_BICG_VECTOR,         /* BiCGSTAB vector updates and reductions     */
  PCF_K_FLUXES,              /* reconstruction, stress and flux kernels    */
  PCF_K_UPDATES,             /* RK3 and RHS updates, sponges, boundaries   */
  PCF_K_NET,                 /* the learned guess: input, network, output  */
  PCF_K_COUNT
};
/* Record HIP events around every launch of the groups above (on = 1) or stop
 * (on = 0).  Turning it on clears the totals. */
int pcf_profile_enable(pcf_state* st, int on);
/* The same for the groups whose bit (1 << group) is set in mask only, so
 * that the events of the other groups do not delay their launches; mask 0
 * stops. */
int pcf_profile_groups(pcf_state* st, uint32_t mask);
/* Launch count and summed device time of one group since it was enabled. */
int pcf_profile_query(pcf_state* st, int group, int64_t* launches, double* total_ms);
/* Launch shapes the autotuner chose for the fused RK-stage kernels (no
 * reference counterpart): shape[2 v + p] for variable v (0 rho, 1 rho', 2 u,
 * 3 v, 4 w, 5 chi, 6 unused) and flux precision p (0 FP64, 1 FP32), -1 where
 * not tuned yet; shape s means 1, 2 or 4 rows (s / 3) and 4, 8 or 16 levels
 * (s % 3) per thread.  ms (may be null) gets the 9 timed shapes per slot in
 * milliseconds per launch (14 x 9 floats). */
int pcf_stage_tuning(const pcf_state* st, int32_t* shape, float* ms);
/* Set the launch shape of one slot of pcf_stage_tuning (0..8), or -1 to
 * have it tuned again on its next launch. */
int pcf_set_stage_shape(pcf_state* st, int32_t slot, int32_t shape);
/* BiCGSTAB iterations replayed from a hipGraph captured on the first
 * iteration (on = 1) or launched kernel by kernel (on = 0, the default);
 * the kernels, and so the results, are the same either way.  While a
 * capture is in progress, HIP refuses stream synchronisations and
 * allocations of other host threads, so turn it on only for a state whose
 * thread is the only one using HIP (one state per process, or several
 * stepped from one thread).  No reference counterpart: the reference has no
 * launches to batch. */
int pcf_set_solver_graphs(pcf_state* st, int32_t on);

/* ---- ensembles (src/ensemble.jl:20-74) ------------------------------------
 * The reference runs each member of ensemble(simulation, parameters,
 * keywords) in an MPI communicator of its own.  Here the members share the
 * GPU, each a state stepped by a host thread of its own (one thread per
 * state, as for every state).  Attached to a pcf_ensemble, a member's
 * BiCGSTAB solves run in the ensemble's batch: a service thread admits a
 * solve at the next iteration boundary, runs one iteration of every solve
 * in progress as one launch per kernel whose grid covers these solves (a
 * list of their members in a device table of the members' vectors and
 * operators; a solve's convergence flag masks its blocks in the iteration
 * after its last), and releases a solve as soon as it has converged or
 * reached its poisson_iterations.  Members step independently; every member
 * gets the bits of its own solve.  The batch
 * covers the line preconditioner (FP64, or FP32 in the mixed solve) on grids
 * with an even x_size (a multiple of 4 for the mixed solve); other solves of
 * a member run on their own.  A member's thread detaches when it stops
 * stepping. */
typedef struct pcf_ensemble pcf_ensemble;
/* no device work until the first attach */
int pcf_ensemble_create(pcf_ensemble** out);
/* PCF_E_ARG while members are attached */
int pcf_ensemble_destroy(pcf_ensemble* e);
/* at most 128 members and 8 solve kinds, one device; PCF_E_ARG for a state
 * attached already.  Call from the thread that steps st, between its calls. */
int pcf_ensemble_attach(pcf_ensemble* e, pcf_state* st);
/* also done by pcf_destroy */
int pcf_ensemble_detach(pcf_state* st);
/* out: batched iterations (passes of the service), member solves run in the
 * batch, solves run on their own, member-table builds */
int pcf_ensemble_stats(pcf_ensemble* e, int64_t out[4]);

/* ---- initial-guess network of the learned solve (no reference
 * counterpart) -------------------------------------------------------------
 * Six 3x3x3 convolutions, 3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1 channels,
 * GELU after the first five; periodic in x and y, zero beyond the walls in
 * z.  Parameters: one FP32 vector, layer by layer W[27][cout][cin] then
 * b[cout] (tap t = (dx+1) + 3 (dy+1) + 9 (dz+1)).  Device vectors are
 * channels-last FP32: input n x 3, output and target n, with voxel
 * i + nx (j + ny k).  A handle serves one grid size (nx a multiple of 32,
 * at most 2^31 - 1 voxels); training = 1 keeps what the backward pass
 * needs. */
typedef struct pcf_net pcf_net;
int pcf_net_param_count(void);
int pcf_net_create(int32_t nx, int32_t ny, int32_t nz, int32_t training, pcf_net** out);
int pcf_net_destroy(pcf_net* net);
int pcf_net_set_params(pcf_net* net, const float* host_params);
int pcf_net_get_params(pcf_net* net, float* host_params);
/* output = network(input) */
int pcf_net_forward(pcf_net* net, const float* d_input, float* d_output);
/* mean-squared-error loss against d_target and its gradient (host FP32,
 * pcf_net_param_count() values) */
int pcf_net_gradients(pcf_net* net, const float* d_input, const float* d_target, float* host_grads, double* loss);
/* one Adam step (learning rate lr) on the loss against d_target */
int pcf_net_train_step(pcf_net* net, const float* d_input, const float* d_target, float lr, double* loss);
/* one Adam step (learning rate lr) for a loss whose gradient with respect to
 * the network output is d_dout (device, FP32, one value per voxel), e.g. the
 * Poisson residual of the guess it proposes */
int pcf_net_train_step_dout(pcf_net* net, const float* d_input, const float* d_dout, float lr);
/* Input channels of the learned initial guess for the right-hand side d_b
 * and the previous solution d_prev (interior FP64 device vectors; d_prev may
 * be null = zero): b / |b|, prev / |b| and jac - 1 (RMS norm |b|), written
 * to d_input (n x 3 FP32); *b_norm gets |b|.  The guess is |b| times the
 * network's output. */
int pcf_guess_input(pcf_state* st, const double* d_b, const double* d_prev, float* d_input, double* b_norm);
/* Attach the network used by initial_guess = PCF_GUESS_LEARNED (not owned;
 * its grid must be the state's). */
int pcf_set_guess_network(pcf_state* st, pcf_net* net);
/* Number of solves whose initial guess was replaced by the zero start
 * because it left a residual no smaller than |b|. */
int pcf_guess_fallbacks(const pcf_state* st, int32_t* count);
/* Attach the network of the learned flexible-BiCGSTAB preconditioning stage
 * (no reference counterpart; not owned, grid of the state; null detaches):
 * in the first `iterations` BiCGSTAB iterations of each mixed solve, every
 * preconditioner application y = M^-1 v is followed by
 * y += |v| net(v / |v|, y / |v|, J - 1) (|v| the RMS norm).  The solution is
 * accumulated in the flexible (textbook) form; a solve that does not
 * converge with the stage is rerun without it from x0 = 0 (the plain
 * preconditioned BiCGSTAB is the fallback).  Needs precision = mixed and a
 * preconditioner (PCF_E_CONFIG otherwise). */
int pcf_set_stage_network(pcf_state* st, pcf_net* net, int32_t iterations);
/* Number of solves rerun without the learned stage. */
int pcf_stage_fallbacks(const pcf_state* st, int32_t* count);
/* Objectives of pcf_net_train_residual. */
enum pcf_objective { PCF_OBJECTIVE_RESIDUAL = 0, PCF_OBJECTIVE_ENERGY = 1 };
/* Train a network on what it proposes for A x = v (no reference
 * counterpart): for the interior-sized FP64 device vectors v and base (base
 * may be null = 0), x = base + |v| net(v / |v|, base / |v|, J - 1), A the
 * state's current total operator.  Takes `inner` Adam steps (learning rate
 * lr) on the objective:
 *   PCF_OBJECTIVE_RESIDUAL  L = |A x - v|^2 / |v|^2, output gradient
 *                           2 |v| A^T (A x - v) / |v|^2;
 *   PCF_OBJECTIVE_ENERGY    the energy norm of the error under -A (whose
 *                           symmetric part is positive semidefinite), output
 *                           gradient -2 |v| (A x - v) / |v|^2: weights the
 *                           smooth error components that set BiCGSTAB's
 *                           iteration count more than the residual does.
 * inner = 0 only evaluates.  *loss_base = L of base alone, *loss_net = L with
 * the network after the steps (the residual loss for either objective;
 * either pointer may be null).  v = b, base = the warm start trains the
 * initial guess; v = a vector being preconditioned, base = M^-1 v trains the
 * preconditioning stage. */
int pcf_net_train_residual(pcf_state* st, pcf_net* net, const double* d_v, const double* d_base, int32_t objective,
                           int32_t inner, float lr, double* loss_base, double* loss_net);

#ifdef __cplusplus
}
#endif

#endif /* PINCFLOW_HIP_H */
