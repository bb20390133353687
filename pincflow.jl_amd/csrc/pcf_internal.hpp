// Internal declarations of libpincflow_hip.so.
//
// Device layout: every field is one FP64 array in the reference's Julia
// column-major order, x fastest.  Haloed fields keep the reference's three
// halo layers in x, y and z (nxx x nyy x nzz) so every kernel runs over the
// reference's own index ranges (i0 - 1 : i1 and so on); Poisson vectors are
// interior-sized (nx x ny x nz).  Indices in kernels are 1-based like the
// reference loops they implement.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pincflow_hip.h"

namespace pcf {

// ---------------------------------------------------------------------------
// grid geometry passed by value to kernels
// ---------------------------------------------------------------------------
struct Geo {
  int nx, ny, nz;        // interior
  int nxx, nyy, nzz;     // haloed
  int nbx, nby, nbz;
  int i0, i1, j0, j1, k0, k1;
  int z_size;            // global z size (= nz on one GPU)
  double dx, dy, dz;
  __host__ __device__ size_t h(int i, int j, int k) const {  // haloed, 1-based
    return (size_t)(i - 1) + (size_t)nxx * ((size_t)(j - 1) + (size_t)nyy * (size_t)(k - 1));
  }
  __host__ __device__ size_t v(int i, int j, int k) const {  // interior vector, 1-based
    return (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
  }
  __host__ __device__ size_t nh() const { return (size_t)nxx * nyy * nzz; }
  __host__ __device__ size_t nv() const { return (size_t)nx * ny * nz; }
};

// Physical constants used on the device (Constants.jl:108-182).
struct Phys {
  double kappa, kappainv, mainv2, ma2, g_ndim, re, fc;  // fc = f tref
  double tref, lref, uref;
  int model;        // enum pcf_model
  int damp;         // damp_horizontal_wind_on_rhs
  int viscous;      // 1 / re > eps
  // compressible model: gamma and dP/dpi = dpdpi_c p^(2 - gamma) with
  // dpdpi_c = 1 / (gamma - 1) (rsp / pref)^(1 - gamma) (update!.jl:1385-1386)
  double gamma, rsp, dpdpi_c;
};

// Index range of a kernel launch (1-based, inclusive start, counts).
struct Range3 {
  int i0, ni, j0, nj, k0, nk;
  __host__ __device__ long long count() const { return (long long)ni * nj * nk; }
};

// ---------------------------------------------------------------------------
// the 25 operator coefficients (order of enum pcf_vector from PCF_V_AC)
// ---------------------------------------------------------------------------
enum {
  T_AC = 0, T_AL, T_AR, T_AB, T_AF, T_AD, T_AU,
  T_ALD, T_ALU, T_ARD, T_ARU, T_ABD, T_ABU, T_AFD, T_AFU,
  T_ADD, T_AUU, T_ALDD, T_ALUU, T_ARDD, T_ARUU, T_ABDD, T_ABUU, T_AFDD, T_AFUU,
  T_COUNT
};

struct Tensor {
  double* a[T_COUNT];
};

// One level of the multigrid preconditioner (FP32): operator, line-solve
// factors and vectors; fx, fy: aggregation factors to the next level.
struct MgLevel {
  int nx = 0, ny = 0, nz = 0, fx = 1, fy = 1;
  bool zf = false;  // arrays stored z fastest (else x fastest)
  float deta = 0.0f;
  float* a[T_COUNT] = {};
  uint16_t* a16[T_COUNT] = {};  // bf16 copies read by the row kernels (storage BF16)
  float *P = nullptr, *E = nullptr, *Q = nullptr;
  float *b = nullptr, *s = nullptr, *rhs = nullptr, *col = nullptr;
};

// BiCGSTAB scalars kept on the device between kernels (apply_bicgstab!.jl:24-150).
struct SolverScalars {
  double rr0;      // <r, r0> of the current iterate
  double vr0;      // <v, r0>
  double alpha;
  double ts, tt;   // <t, s>, <t, t>
  double omega;
  double rr;       // <r, r>
  double rr0_new;  // <r_new, r0>
  double beta;
  double res;      // RMS residual of the last update
  double b_norm;
  double tol;
  double tolref;
  double divl2, divl2_norm;
  int32_t done;    // 1 once res / b_norm <= tol; 2: stopped at the ensemble solve's maxit
  int32_t niter;
  int32_t pad[2];
};

constexpr int kReduceBlock = 256;     // threads per reduction block
// Partial-sum slots per dot product in pcf_state::partials: one per block of
// the dot-fused kernels, ceil(nx / 64) ceil(ny / 4) nz at most; pcf_create
// refuses grids that need more (check_reduction_sizes)
constexpr int kPartialStride = 1 << 20;
constexpr int kMaxPartials = 4096;    // partial sums per reduction
constexpr int kMaxDots = 4;

// ---------------------------------------------------------------------------
// ensemble solve: one launch per BiCGSTAB kernel over the members of an
// ensemble whose solves are in progress (pcf_ensemble); the member table
// lives in device memory
// ---------------------------------------------------------------------------
constexpr int kEnsMax = 128;   // members of an ensemble
constexpr int kEnsGroups = 8;  // solve kinds of an ensemble
// the solver vectors of a member, by index
enum { EV_X = 0, EV_R0, EV_R, EV_P, EV_V, EV_S, EV_T, EV_Y, EV_Z, EV_PCTMP, EV_PCRHS, EV_COUNT };
struct EnsMember {
  Tensor T;                         // FP64 operator
  double* v[EV_COUNT];
  const double *P, *E, *Q;          // FP64 line-solve factors
  const float *P32, *E32, *Q32;     // mixed: FP32 factors, coefficients and work
  const float* T32[T_COUNT];
  const uint16_t* T16[T_COUNT];     // bf16 storage (else null)
  float *col32, *cur32, *rhs32;
  double* partials;
  SolverScalars* sc;
  size_t n;                         // cells
  double nxyz;                      // x_size y_size z_size
  double deta;
  int nx, ny, nz;
  int aparts;                       // partials of the dot-fused apply
  int maxit;                        // poisson_iterations
  int pad;
};
struct EnsTable {
  int count;
  EnsMember m[kEnsMax];
};
// The solves in one batched iteration: their members' slots in the table
// and the first block of each in the launch grids of the iteration's
// kernels, which cover these members only.
struct EnsList {
  int count;
  int member[kEnsMax];
  int pair_start[kEnsMax + 1];      // apply and pc-rhs grids
  int thomas_start[kEnsMax + 1];    // line-solve grids
  int rhs32_start[kEnsMax + 1];     // FP32 pc-rhs grid (mixed)
};

// ---------------------------------------------------------------------------
// host-side state
// ---------------------------------------------------------------------------
struct HostSetup;  // setup.cpp

}  // namespace pcf

struct pcf_net;

// The members of an ensemble (pcf_ensemble_attach), each stepped by a host
// thread of its own, and the service thread that runs their solves: a
// member's solve is admitted at the next iteration boundary, runs with the
// solves already in progress as one launch per kernel over the solves of
// its kind, and is released as soon as the service sees its convergence
// flag.
struct pcf_ensemble {
  std::mutex mu;
  std::condition_variable done_cv;       // member threads wait here for their solve
  std::condition_variable wake;          // the service waits here for work
  std::thread service;
  bool stop = false;
  int err = 0;                           // the service's first error (PCF_E_*), sticky
  std::vector<pcf_state*> pending;       // solves waiting for admission
  hipStream_t stream = nullptr;
  int device = 0;
  int attached = 0;
  // per solve kind (FP64 or mixed, preconditioner iterations, coefficient
  // storage): the members, the device member table and its host copy, each
  // slot's blocks in the grids, and the solves of the current iteration
  // (host, two pinned slots, device)
  struct Group {
    int key = -1;
    std::vector<pcf_state*> slots;
    bool stale = true;                   // the table does not match slots yet
    std::vector<pcf_state*> active;      // solves in the batch
    pcf::EnsTable* table = nullptr;
    std::unique_ptr<pcf::EnsTable> host;
    std::vector<int> pair_blocks, thomas_blocks, rhs32_blocks;
    pcf::EnsList* list_host = nullptr;   // pinned, two slots
    pcf::EnsList* list = nullptr;
  };
  std::vector<std::unique_ptr<Group>> groups;
  int32_t* flags_dev = nullptr;          // done flags, kEnsMax per group
  int32_t* flags_host = nullptr;         // pinned, two slots of kEnsGroups x kEnsMax
  hipEvent_t ev[2] = {nullptr, nullptr};
  int64_t iterations = 0, member_solves = 0, solves_alone = 0, table_builds = 0;
};

struct pcf_state {
  pcf_config cfg;
  pcf::Geo g;
  pcf::Phys ph;
  double tref, lref, uref, rhoref, thetaref;
  double alphark[3], betark[3], stepfrac[3];
  double dzcmin;
  double deta;            // preconditioner pseudo-time step
  hipStream_t stream = nullptr;
  int device = 0;

  // predictands and copies (haloed); P and its copies only in the
  // compressible model (set_p.jl), else null
  double *rho, *rhop, *u, *v, *w, *pip;
  double* p = nullptr;
  double *p0[7] = {}, *p1[7] = {};  // rho, rhop, u, v, w, pip, p
  bool compressible = false;
  double *dp = nullptr;
  double* phitheta[3] = {};           // heat-conduction fluxes (compressible)
  double *drho, *drhop, *du, *dv, *dw;
  double *rhoold, *rhopold, *uold, *vold, *wold;
  // tracer (cfg.tracer_setup == PCF_TRACER_ON): rho chi, its increment, its
  // initial state (the sponge's background) and the chi0 backup of a step
  bool tracer = false;
  double *chi = nullptr, *dchi = nullptr, *chi_init = nullptr, *chi0 = nullptr;
  // background and grid (haloed, static)
  double *pbar, *thetabar, *rhobar, *n2, *jac, *met13, *met23, *met33, *zc, *zctilde, *fcscal;
  // sponges
  double *alphar, *betar, *betar_s;  // betar_s: dt-independent part of betar
  // fluxes (haloed): [var][dir]
  double* flux[7][3] = {};            // rho, rhop, u, v, w, chi, p
  double* phi[6] = {};                // reconstruction inputs
  double* wcart;                      // Cartesian vertical wind
  double* stress[6];                  // 11, 12, 13, 22, 23, 33
  // flux_precision MIXED: FP32 copies of phi and the stress components that
  // the RK stages' flux evaluation reads (null otherwise)
  float* phi32[6] = {};
  float* stress32[6] = {};
  // the face carriers of the p0 winds (k_carriers), FP64 and, with
  // flux_precision MIXED, FP32
  double* carrier[3] = {};
  float* carrier32[3] = {};
  // t13, t23 of the stress kernel (the U and V vertical viscous fluxes' cell
  // values), FP64 and FP32
  double* tstress[2] = {};
  float* tstress32[2] = {};
  // with FP32 fluxes, the FP64 background part of the Rho flux at the x, y
  // and z faces (k_rho_background; pseudo-incompressible and compressible)
  double* rho_bg[3] = {};
  // between flux computations stress[0] and stress[1] hold the flux
  // divergences of the stage's U and V updates, which the W update reuses
  bool lhs_div_valid[2] = {false, false};
  // the same divergences from the fused stage kernels (stage_update), which
  // read the stress arrays while they write these
  double *div_u = nullptr, *div_v = nullptr;
  double* fdiv_w = nullptr;  // the W flux divergence of a fused stage
  // launch-shape autotuner of the fused stage kernels: the chosen shape per
  // (variable, FP64/FP32) (-1: not tuned yet), the timings, and copies of a
  // variable, its increment and its flux divergence the timing runs write
  int stage_shape[14];
  float stage_shape_ms[14][9] = {};
  double *tune_q = nullptr, *tune_dq = nullptr, *tune_div = nullptr;
  // corrector
  double *dpip, *corx, *cory;
  // Poisson (interior)
  pcf::Tensor T;
  double *rhs, *solution;
  double *bp, *br0, *br, *bs, *bt, *bv, *bvpc, *bmat;
  double *pc_rhs, *pc_q, *pc_tmp;     // preconditioner work (pc_q: also scratch of the sponge)
  double *pc_P, *pc_E, *pc_Q;         // line-solve factors of the current operator
  bool pc_factors_valid = false;
  // mixed precision (allocated when cfg.precision == PCF_PRECISION_MIXED):
  // FP32 copies of the 22 horizontal operator coefficients and of the
  // line-solve factors, FP32 preconditioner work, and z = M^-1 s
  float* T32[pcf::T_COUNT] = {};
  uint16_t* T16[pcf::T_COUNT] = {};   // bf16 bits of T32 (preconditioner_storage BF16)
  float *pc32_P = nullptr, *pc32_E = nullptr, *pc32_Q = nullptr;
  float *m_pc_tmp = nullptr, *m_pc_cur = nullptr, *m_pc_rhs = nullptr;
  double* bzpc = nullptr;
  bool t32_valid = false;
  std::vector<pcf::MgLevel> mg;       // multigrid levels (level 0 aliases the buffers above)
  // warm start of the mixed solve: solution / dt of each implicit stage of
  // the previous time step (slot 0: first stage, 1: second stage)
  double* m_prev[2] = {nullptr, nullptr};
  bool m_prev_valid[2] = {false, false};
  // initial_guess EXTRAPOLATED: the same per slot one step older, the time
  // step of the solve m_prev holds, and the extrapolated guess
  double* m_prev2[2] = {nullptr, nullptr};
  bool m_prev2_valid[2] = {false, false};
  double m_prev_dt[2] = {0.0, 0.0};
  double* m_guess = nullptr;
  int solve_slot = -1;                // stage of the solve in progress, -1 outside pcf_step
  // learned initial guess: the attached network and its FP32 input / output
  pcf_net* guess_net = nullptr;
  float *guess_in = nullptr, *guess_out = nullptr;
  double* guess_x = nullptr;
  int guess_fallbacks = 0;            // solves whose guess was dropped for a zero start
  // learned flexible preconditioning stage of the mixed solve: the attached
  // network, the number of leading BiCGSTAB iterations whose preconditioner
  // applications it follows, its FP32 input and the device RMS norm of the
  // vector being preconditioned
  pcf_net* stage_net = nullptr;
  int stage_iterations = 0;
  float* stage_in = nullptr;
  double* stage_norm = nullptr;
  int stage_fallbacks = 0;            // solves rerun without the stage after it failed to converge
  double* train_buf[3] = {};          // network training work: x, A x - v, A^T (A x - v)
  // the BiCGSTAB iteration captured as a hipGraph (FP64 / mixed solve),
  // instantiated on first use and replayed once per iteration (when
  // solver_graphs is set: pcf_set_solver_graphs)
  hipGraphExec_t iter_graph[2] = {nullptr, nullptr};
  bool solver_graphs = false;
  // ensemble membership (pcf_ensemble_attach): the ensemble, the member's
  // group (-1: its solves run on their own) and slot in the group's table,
  // the events that order the service's work after this state's stream and
  // this state's after the service's, and the state of a solve in the batch
  pcf_ensemble* ens = nullptr;
  int ens_group = -1, ens_slot = -1;
  hipEvent_t ens_ready = nullptr, ens_done = nullptr;
  bool ens_active = false;
  long long ens_since = 0;            // the service pass that admitted the solve
  int ens_rc = -1;                    // -1 while the solve is in the batch
  double* partials;                   // reduction partials
  pcf::SolverScalars* sc;             // device scalars
  pcf::SolverScalars* sc_host;        // pinned copy
  double* cfl_host;                   // pinned CFL maxima
  double* cfl_dev;

  // time loop
  double time = 0.0, nextoutputtime = 0.0;
  int itime = 0;
  double last_residual = 0.0;

  // kernel timing (profile.cpp)
  struct ProfRec {
    int group;
    hipEvent_t a, b;
  };
  bool prof_on = false;
  uint32_t prof_mask = 0;             // groups recorded while prof_on
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  double prof_ms[PCF_K_COUNT] = {};
  long long prof_n[PCF_K_COUNT] = {};

  std::vector<void*> allocations;  // hipMalloc bases, freed by pcf_destroy
  unsigned alloc_count = 0;         // staggers successive buffers (device_alloc)
};

namespace pcf {
// Bracket a launch with events when profiling is on (profile.cpp).
void prof_begin(pcf_state* st, int group, hipEvent_t* a);
void prof_end(pcf_state* st, int group, hipEvent_t a);
void prof_flush(pcf_state* st);
void prof_release(pcf_state* st);
struct ProfScope {
  pcf_state* st;
  int group;
  hipEvent_t a = nullptr;
  ProfScope(pcf_state* s, int g) : st(s), group(g) { prof_begin(st, group, &a); }
  ~ProfScope() { prof_end(st, group, a); }
};
}  // namespace pcf

namespace pcf {

// ---- setup.cpp ----------------------------------------------------------
int host_setup(pcf_state* st, std::string& err);

// ---- errors -------------------------------------------------------------
void set_error(const std::string& msg);
#define PCF_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::pcf::set_error(std::string(#call) + ": " + hipGetErrorString(e_));    \
      return PCF_E_HIP;                                                       \
    }                                                                         \
  } while (0)

// ---- boundaries.hip -------------------------------------------------------
int bc_xy(pcf_state* st, double* const* fields, int n, int nbx = -1, int nby = -1, int nbz = -1);
int bc_z(pcf_state* st, double* f, int mode, bool staggered);
int set_boundaries_predictands(pcf_state* st);
int set_boundaries_fluxes(pcf_state* st);

// ---- poisson.hip ----------------------------------------------------------
int compute_lhs(pcf_state* st, double* tolref_out);
int compute_operator(pcf_state* st, double dt, double rf);
int apply_operator(pcf_state* st, const double* sin, double* ls, bool total);
int apply_operator_transpose(pcf_state* st, const double* r, double* out);
int apply_preconditioner(pcf_state* st, const double* sin, double* sout);
int apply_bicgstab(pcf_state* st, double tolref, int* niter);
int bicgstab_core(pcf_state* st, double tolref, bool mixed, const double* guess, double guess_scale, int* niter);
int solve(pcf_state* st, double tolref, double dt, int* niter);  // apply_bicgstab, configured precision
int pc_factors(pcf_state* st);
// ensemble (poisson.hip): the solve kind of a member (-1: the member table
// does not cover it), whether this solve joins the batch, the table upload,
// one batched iteration of a group's active solves (list slot `parity`,
// done flags to `flags`); ensemble.cpp: the member's side
int ensemble_key(const pcf_state* st);
bool ensemble_batchable(const pcf_state* st, bool mixed, bool stage);
int ensemble_upload(pcf_ensemble* e, pcf_ensemble::Group& grp);
int ensemble_iterate(pcf_ensemble* e, pcf_ensemble::Group& grp, int parity, int32_t* flags);
int ensemble_solve(pcf_state* st);
int check_reduction_sizes(const Geo& g, std::string& err);
int apply_corrector(pcf_state* st, double dt, double rf, int* niter);
int correct(pcf_state* st, double dt, double rf);
int dot(pcf_state* st, const double* a, const double* b, double* out);

// ---- mixed.hip -------------------------------------------------------------
int apply_bicgstab_mixed(pcf_state* st, double tolref, double dt, int* niter);
int precondition_mixed(pcf_state* st, const double* sin, double* sout, const SolverScalars* sc);
// the ensemble batch's FP32 line preconditioner and its pc-rhs grid per member
int ensemble_rhs32_blocks(const pcf_state* st);
void ensemble_precondition_mixed(hipStream_t s, const EnsTable* dT, const EnsList* dL, const EnsList& L, int iters,
                                 bool bf16, int in, int out);
int alloc_mixed(pcf_state* st);
// Device buffer of `bytes` owned by the state (freed by pcf_destroy).  The
// start of the n-th buffer is offset by (n mod 28) x 2304 B within its
// allocation: buffers that kernels stream side by side (the 25 operator
// coefficients, the solver vectors) then start on different HBM channels
// instead of all at the same offset from a large power of two, which made
// the streams of one wavefront contend for the same channels.
int device_alloc(pcf_state* st, void** p, size_t bytes);
template <class T>
int device_alloc(pcf_state* st, T** p, size_t count) {
  return device_alloc(st, (void**)p, count * sizeof(T));
}

// ---- net.hip -----------------------------------------------------------------
int learned_guess(pcf_state* st, double dt, const double* prev, const double** x0);
// y += |v| net(v / |v|, y / |v|, J - 1) on the state's stream, |v| the RMS
// norm computed on the device; kernels after convergence (sc->done) leave y.
int learned_stage(pcf_state* st, const double* v, double* y, const SolverScalars* sc);

// ---- explicit.hip --------------------------------------------------------
int reconstruct(pcf_state* st);
int compute_fluxes(pcf_state* st);
int stage_inputs(pcf_state* st, int m);  // m: the RK stage (1: also the face carriers)
// the fused fluxes + update of one variable in RK stage m, with the LHS
// sponge of time step sdt applied in the same kernel when sponge_fused
int stage_update(pcf_state* st, double dt, int m, int var, double sdt);
bool sponge_fused(const pcf_state* st, int var);
int update_lhs(pcf_state* st, double dt, int m, int var);
int update_rhs(pcf_state* st, double dt, int var, bool implicit, double rf);
int apply_lhs_sponge(pcf_state* st, double dt, double time, int var);
int compute_sponges(pcf_state* st, double dt, double time);
int cfl_maxima(pcf_state* st, double out[3]);
int copy_field(pcf_state* st, double* dst, const double* src);
int compute_theta_fluxes(pcf_state* st);
int modify_compressible_wind(pcf_state* st, bool multiply);
int synchronize_compressible_atmosphere(pcf_state* st, const double* p);
int synchronize_density_fluctuations(pcf_state* st);
int refresh_fcscal(pcf_state* st);

// ---- launch helpers -------------------------------------------------------
inline unsigned blocks_for(long long n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace pcf
