// C ABI of libpincflow_hip.so (include/pincflow_hip.h) and the host side
// of the time loop (src/Integration).  The host only sequences kernels on
// the state's stream; fields never leave HBM during a step.  Host and device
// synchronise once per step for the CFL time step and a few times per
// BiCGSTAB solve (see poisson.hip).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>

#include "pcf_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace pcf {
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace pcf

namespace pcf {

constexpr size_t kStaggerStep = 2304;  // 9 x 256 B
constexpr unsigned kStaggerSlots = 28;

int device_alloc(pcf_state* st, void** p, size_t bytes) {
  const size_t off = (size_t)(st->alloc_count++ % kStaggerSlots) * kStaggerStep;
  char* base = nullptr;
  PCF_HIP(hipMalloc((void**)&base, bytes + off));
  st->allocations.push_back(base);
  *p = base + off;
  return PCF_OK;
}

}  // namespace pcf

using namespace pcf;

namespace {

int alloc(pcf_state* st, double** p, size_t n) {
  int rc = device_alloc(st, p, n);
  if (rc) return rc;
  PCF_HIP(hipMemsetAsync(*p, 0, n * sizeof(double), st->stream));
  return PCF_OK;
}

int alloc32(pcf_state* st, float** p, size_t n) {
  int rc = device_alloc(st, p, n);
  if (rc) return rc;
  PCF_HIP(hipMemsetAsync(*p, 0, n * sizeof(float), st->stream));
  return PCF_OK;
}

double* field_ptr(pcf_state* st, int f) {
  switch (f) {
    case PCF_F_RHO: return st->rho;
    case PCF_F_RHOP: return st->rhop;
    case PCF_F_U: return st->u;
    case PCF_F_V: return st->v;
    case PCF_F_W: return st->w;
    case PCF_F_PIP: return st->pip;
    case PCF_F_PBAR: return st->pbar;
    case PCF_F_THETABAR: return st->thetabar;
    case PCF_F_RHOBAR: return st->rhobar;
    case PCF_F_N2: return st->n2;
    case PCF_F_JAC: return st->jac;
    case PCF_F_MET13: return st->met13;
    case PCF_F_MET23: return st->met23;
    case PCF_F_MET33: return st->met33;
    case PCF_F_ZC: return st->zc;
    case PCF_F_ZCTILDE: return st->zctilde;
    case PCF_F_ALPHAR: return st->alphar;
    case PCF_F_BETAR: return st->betar;
    case PCF_F_DPIP: return st->dpip;
    case PCF_F_CORX: return st->corx;
    case PCF_F_CORY: return st->cory;
    case PCF_F_DRHO: return st->drho;
    case PCF_F_DRHOP: return st->drhop;
    case PCF_F_DU: return st->du;
    case PCF_F_DV: return st->dv;
    case PCF_F_DW: return st->dw;
    case PCF_F_RHOOLD: return st->rhoold;
    case PCF_F_RHOPOLD: return st->rhopold;
    case PCF_F_UOLD: return st->uold;
    case PCF_F_VOLD: return st->vold;
    case PCF_F_WOLD: return st->wold;
    case PCF_F_P0_RHO: return st->p0[0];
    case PCF_F_P0_RHOP: return st->p0[1];
    case PCF_F_P0_U: return st->p0[2];
    case PCF_F_P0_V: return st->p0[3];
    case PCF_F_P0_W: return st->p0[4];
    case PCF_F_CHI: return st->chi;
    case PCF_F_DCHI: return st->dchi;
    case PCF_F_INITIALTRACER: return st->chi_init;
    case PCF_F_CHI0: return st->chi0;
    case PCF_F_PHICHI_X: case PCF_F_PHICHI_Y: case PCF_F_PHICHI_Z:
      return st->flux[5][f - PCF_F_PHICHI_X];
    case PCF_F_P: return st->p;
    case PCF_F_DP: return st->dp;
    case PCF_F_P0_P: return st->p0[6];
    case PCF_F_PHIP_X: case PCF_F_PHIP_Y: case PCF_F_PHIP_Z:
      return st->flux[6][f - PCF_F_PHIP_X];
    case PCF_F_PHITHETA_X: case PCF_F_PHITHETA_Y: case PCF_F_PHITHETA_Z:
      return st->phitheta[f - PCF_F_PHITHETA_X];
    default:
      if (f >= PCF_F_PHIRHO_X && f <= PCF_F_PHIW_Z) {
        int q = (f - PCF_F_PHIRHO_X) / 3, d = (f - PCF_F_PHIRHO_X) % 3;
        return st->flux[q][d];
      }
      return nullptr;
  }
}

double* vector_ptr(pcf_state* st, int v) {
  if (v == PCF_V_RHS) return st->rhs;
  if (v == PCF_V_SOLUTION) return st->solution;
  if (v >= PCF_V_AC && v < PCF_V_COUNT) return st->T.a[v - PCF_V_AC];
  return nullptr;
}

// the predictands in the order of p0 / p1; returns how many there are (P
// only in the compressible model)
int predictand_slots(pcf_state* st, double* out[7]) {
  out[0] = st->rho; out[1] = st->rhop; out[2] = st->u; out[3] = st->v; out[4] = st->w; out[5] = st->pip;
  out[6] = st->p;
  return st->compressible ? 7 : 6;
}

int save_backups(pcf_state* st, bool rho, bool rhop, bool u, bool v, bool w) {  // save_backups!.jl
  int rc = PCF_OK;
  if (rho && (rc = copy_field(st, st->rhoold, st->rho))) return rc;
  if (rhop && (rc = copy_field(st, st->rhopold, st->rhop))) return rc;
  if (u && (rc = copy_field(st, st->uold, st->u))) return rc;
  if (v && (rc = copy_field(st, st->vold, st->v))) return rc;
  if (w && (rc = copy_field(st, st->wold, st->w))) return rc;
  return rc;
}

#define TRY(x)            \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

// explicit_integration!.jl:42-101.  Each RK stage computes the flux inputs
// once (reconstruct!, stress tensor) and then, per variable, its fluxes and
// update in one fused kernel (stage_update); compute_fluxes! and update!
// stay available as separate calls through the ABI.
int explicit_lhs(pcf_state* st, double dt, double time) {
  for (int m = 1; m <= 3; ++m) {
    const double sdt = st->stepfrac[m - 1] * dt;
    TRY(stage_inputs(st, m));
    st->lhs_div_valid[0] = st->lhs_div_valid[1] = false;  // the stress arrays were rewritten
    TRY(save_backups(st, true, false, true, true, false));
    // the sponges that the fused kernels do not apply themselves follow
    // their variable's update (the wind's after all three updates), in the
    // reference's order
    for (int var : {PCF_VAR_RHO, PCF_VAR_RHOP, PCF_VAR_P, PCF_VAR_CHI}) {  // P: compressible; CHI: tracer
      TRY(stage_update(st, dt, m, var, sdt));
      if (!sponge_fused(st, var)) TRY(apply_lhs_sponge(st, sdt, time, var));
    }
    TRY(set_boundaries_predictands(st));
    TRY(stage_update(st, dt, m, PCF_VAR_U, sdt));
    TRY(stage_update(st, dt, m, PCF_VAR_V, sdt));
    TRY(stage_update(st, dt, m, PCF_VAR_W, sdt));
    for (int var : {PCF_VAR_U, PCF_VAR_V, PCF_VAR_W})
      if (!sponge_fused(st, var)) TRY(apply_lhs_sponge(st, sdt, time, var));
    TRY(set_boundaries_predictands(st));
  }
  TRY(synchronize_compressible_atmosphere(st, st->p));
  return apply_lhs_sponge(st, dt, time, PCF_VAR_PIP);
}

// explicit_integration!.jl:103-131
int explicit_rhs(pcf_state* st, double dt, double time) {
  (void)time;
  TRY(synchronize_compressible_atmosphere(st, st->p));
  TRY(modify_compressible_wind(st, true));
  TRY(set_boundaries_predictands(st));
  TRY(save_backups(st, false, true, true, true, true));
  TRY(update_rhs(st, dt, PCF_VAR_RHOP, false, 0.0));
  TRY(update_rhs(st, dt, PCF_VAR_U, false, 0.0));
  TRY(update_rhs(st, dt, PCF_VAR_V, false, 0.0));
  TRY(update_rhs(st, dt, PCF_VAR_W, false, 0.0));
  TRY(update_rhs(st, dt, PCF_VAR_PIP, false, 0.0));
  TRY(modify_compressible_wind(st, false));
  return set_boundaries_predictands(st);
}

// implicit_integration!.jl:37-79
int implicit(pcf_state* st, double dt, double time, double rf, int* niter) {
  (void)time;
  TRY(modify_compressible_wind(st, true));
  TRY(set_boundaries_predictands(st));
  TRY(save_backups(st, false, false, false, false, true));
  TRY(update_rhs(st, dt, PCF_VAR_U, true, rf));
  TRY(update_rhs(st, dt, PCF_VAR_V, true, rf));
  TRY(update_rhs(st, dt, PCF_VAR_W, true, rf));
  TRY(set_boundaries_predictands(st));
  TRY(update_rhs(st, dt, PCF_VAR_RHOP, true, rf));
  TRY(set_boundaries_predictands(st));
  TRY(apply_corrector(st, dt, rf, niter));
  TRY(modify_compressible_wind(st, false));
  return set_boundaries_predictands(st);
}

int copy_set(pcf_state* st, double* const* dst, double* const* src, int n) {
  for (int q = 0; q < n; ++q) TRY(copy_field(st, dst[q], src[q]));
  return PCF_OK;
}

// compute_time_step.jl:46-153 (NoWKB)
int time_step(pcf_state* st, double* dt_out) {
  const pcf_config& n = st->cfg;
  double dt;
  if (!n.adaptive_time_step) {
    dt = n.dtmax / st->tref;
  } else {
    double m[3];
    TRY(cfl_maxima(st, m));
    double umax = m[0] + DBL_EPSILON, vmax = m[1] + DBL_EPSILON, wmax = m[2] + DBL_EPSILON;
    const Geo& g = st->g;
    double dtconv = n.cfl_number * std::fmin(std::fmin(g.dx / umax, g.dy / vmax), g.dz / wmax);
    double a = g.dx * g.dx, b = g.dy * g.dy, c = st->dzcmin * st->dzcmin;
    double dtvisc = 0.5 * (a < b ? (a < c ? a : c) : (b < c ? b : c)) * st->ph.re;
    double dm = n.dtmax / st->tref;
    dt = dtvisc < dtconv ? dtvisc : dtconv;
    dt = dt < dm ? dt : dm;
  }
  if (dt * st->tref < n.dtmin) {
    set_error("Error in compute_time_step: dt < dtmin");
    return PCF_E_DTMIN;
  }
  *dt_out = dt;
  return PCF_OK;
}

}  // namespace

extern "C" {

int pcf_abi_version(void) { return PCF_ABI_VERSION; }
const char* pcf_last_error(void) { return g_last_error.c_str(); }
int pcf_config_size(void) { return (int)sizeof(pcf_config); }

void pcf_config_default(pcf_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->abi_version = PCF_ABI_VERSION;
  c->x_size = c->y_size = c->z_size = 1;
  c->nbx = c->nby = c->nbz = 3;
  c->lx = c->ly = c->lz = 1.0e3;
  c->cfl_number = 0.5;
  c->dtmin = 1.0e-6;
  c->dtmax = 1.0e3;
  c->adaptive_time_step = 1;
  c->tolerance = 1.0e-8;
  c->poisson_iterations = 1000;
  c->preconditioner = 1;
  c->dtau = 1.0;
  c->preconditioner_iterations = 2;
  c->initial_cleaning = 1;
  c->tolerance_is_relative = 0;
  c->model = PCF_PSEUDO_INCOMPRESSIBLE;
  c->background = PCF_ISOTHERMAL;
  c->kinematic_viscosity = 1.5e-5;
  c->thermal_conductivity = 3.0e-5;
  c->buoyancy_frequency = 1.0e-2;
  c->potential_temperature = 3.0e2;
  c->temperature = 3.0e2;
  c->ground_pressure = 1.0e5;
  c->coriolis_frequency = 1.0e-4;
  c->tropopause_height = 1.0e4;
  c->troposphere_lapse_rate = 6.5e-3;
  c->stratosphere_lapse_rate = -5.0e-3;
  c->lhs_divisor = 3.0;
  c->rhs_amplitude = 1.0;
  c->nout = 1;
  c->iterations = 1;
  c->output_interval = 3.6e3;
  c->tmax = 3.6e3;
  c->mg_smoothing_steps = 2;
  c->mg_coarse_steps = 20;
}

int pcf_create(const pcf_config* cfg, pcf_state** out) {
  if (!cfg || !out) { set_error("null argument"); return PCF_E_ARG; }
  if (cfg->abi_version != PCF_ABI_VERSION) { set_error("pcf_config.abi_version mismatch"); return PCF_E_ARG; }
  *out = nullptr;
  pcf_state* st = new pcf_state();
  st->cfg = *cfg;
  auto fail = [&](int rc) {
    pcf_destroy(st);
    return rc;
  };
  if (hipGetDevice(&st->device) != hipSuccess) { set_error("no HIP device"); delete st; return PCF_E_HIP; }
  if (hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete st;
    return PCF_E_HIP;
  }
  std::string err;
  int rc = host_setup(st, err);
  if (rc) {
    if (!err.empty()) set_error(err);
    return fail(rc);
  }
  const Geo& g = st->g;
  if ((rc = check_reduction_sizes(g, err))) {
    set_error(err);
    return fail(rc);
  }
  const size_t NH = g.nh(), NV = g.nv();
  double** haloed[] = {&st->drho, &st->drhop, &st->du, &st->dv, &st->dw, &st->rhoold, &st->rhopold, &st->uold,
                       &st->vold, &st->wold, &st->betar, &st->dpip, &st->corx, &st->cory, &st->wcart,
                       &st->div_u, &st->div_v, &st->fdiv_w, &st->tune_q, &st->tune_dq, &st->tune_div};
  for (double** p : haloed)
    if ((rc = alloc(st, p, NH))) return fail(rc);
  for (int& s : st->stage_shape) s = -1;
  for (int q = 0; q < (st->compressible ? 7 : 6); ++q) {
    if ((rc = alloc(st, &st->p0[q], NH))) return fail(rc);
    if ((rc = alloc(st, &st->p1[q], NH))) return fail(rc);
  }
  if (st->compressible) {  // Increments.jl:91-94, Fluxes.jl:89-92
    if ((rc = alloc(st, &st->dp, NH))) return fail(rc);
    for (int d = 0; d < 3; ++d) {
      if ((rc = alloc(st, &st->flux[6][d], NH))) return fail(rc);
      if ((rc = alloc(st, &st->phitheta[d], NH))) return fail(rc);
    }
  }
  for (int q = 0; q < (st->tracer ? 6 : 5); ++q) {
    if ((rc = alloc(st, &st->phi[q], NH))) return fail(rc);
    for (int d = 0; d < 3; ++d)
      if ((rc = alloc(st, &st->flux[q][d], NH))) return fail(rc);
  }
  if (st->tracer)
    for (double** p : {&st->dchi, &st->chi0})
      if ((rc = alloc(st, p, NH))) return fail(rc);
  for (int q = 0; q < 6; ++q)
    if ((rc = alloc(st, &st->stress[q], NH))) return fail(rc);
  for (int d = 0; d < 3; ++d)
    if ((rc = alloc(st, &st->carrier[d], NH))) return fail(rc);
  for (int d = 0; d < 2; ++d)
    if ((rc = alloc(st, &st->tstress[d], NH))) return fail(rc);
  if (cfg->flux_precision == PCF_PRECISION_MIXED) {
    // zeroed like the FP64 copies: the flux stencils read entries outside
    // the range the input kernels write (phi two levels below the ground),
    // which hold zeros there as in the reference's freshly allocated arrays
    for (int q = 0; q < 6; ++q) {
      if ((q < 5 || st->tracer) && (rc = alloc32(st, &st->phi32[q], NH))) return fail(rc);
      if ((rc = alloc32(st, &st->stress32[q], NH))) return fail(rc);
    }
    for (int d = 0; d < 3; ++d)
      if ((rc = alloc32(st, &st->carrier32[d], NH))) return fail(rc);
    for (int d = 0; d < 2; ++d)
      if ((rc = alloc32(st, &st->tstress32[d], NH))) return fail(rc);
    if (cfg->model != PCF_BOUSSINESQ)
      for (int d = 0; d < 3; ++d)
        if ((rc = alloc(st, &st->rho_bg[d], NH))) return fail(rc);
  }
  for (int q = 0; q < T_COUNT; ++q)
    if ((rc = alloc(st, &st->T.a[q], NV))) return fail(rc);
  double** vecs[] = {&st->rhs, &st->solution, &st->bp, &st->br0, &st->br, &st->bs, &st->bt, &st->bv,
                     &st->bvpc, &st->bmat, &st->pc_rhs, &st->pc_q, &st->pc_tmp, &st->pc_P, &st->pc_E, &st->pc_Q};
  for (double** p : vecs)
    if ((rc = alloc(st, p, NV > (size_t)g.nz ? NV : (size_t)g.nz))) return fail(rc);
  if ((rc = alloc(st, &st->partials, (size_t)kMaxDots * kPartialStride))) return fail(rc);
  if (cfg->precision == PCF_PRECISION_MIXED && (rc = alloc_mixed(st))) return fail(rc);
  if ((rc = alloc(st, &st->cfl_dev, 8))) return fail(rc);
  {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(SolverScalars)) != hipSuccess) { set_error("hipMalloc failed"); return fail(PCF_E_HIP); }
    st->allocations.push_back(p);
    st->sc = (SolverScalars*)p;
    if (hipMemsetAsync(p, 0, sizeof(SolverScalars), st->stream) != hipSuccess) return fail(PCF_E_HIP);
    if (hipHostMalloc(&p, 2 * sizeof(SolverScalars), hipHostMallocDefault) != hipSuccess) {
      set_error("hipHostMalloc failed");
      return fail(PCF_E_HIP);
    }
    st->sc_host = (SolverScalars*)p;
    std::memset(p, 0, 2 * sizeof(SolverScalars));
    if (hipHostMalloc(&p, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      set_error("hipHostMalloc failed");
      return fail(PCF_E_HIP);
    }
    st->cfl_host = (double*)p;
  }
  // the stored sponge: alphar is static (uploaded), betar = 0 until the first
  // compute_sponges (Sponge.jl:29-42 allocates zeros)
  if (hipStreamSynchronize(st->stream) != hipSuccess) { set_error("hipStreamSynchronize failed"); return fail(PCF_E_HIP); }
  st->time = 0.0;
  st->nextoutputtime = cfg->output_interval;
  *out = st;
  return PCF_OK;
}

int pcf_destroy(pcf_state* st) {
  if (!st) return PCF_OK;
  if (st->ens) (void)pcf_ensemble_detach(st);
  if (st->stream) (void)hipStreamSynchronize(st->stream);
  if (st->ens_ready) (void)hipEventDestroy(st->ens_ready);
  prof_release(st);
  for (hipGraphExec_t& e : st->iter_graph)
    if (e) (void)hipGraphExecDestroy(e);
  for (void* p : st->allocations) (void)hipFree(p);
  if (st->sc_host) (void)hipHostFree(st->sc_host);
  if (st->cfl_host) (void)hipHostFree(st->cfl_host);
  if (st->stream) (void)hipStreamDestroy(st->stream);
  delete st;
  return PCF_OK;
}

int pcf_dims(const pcf_state* st, int32_t d[6]) {
  if (!st || !d) return PCF_E_ARG;
  d[0] = st->g.nx; d[1] = st->g.ny; d[2] = st->g.nz;
  d[3] = st->g.nxx; d[4] = st->g.nyy; d[5] = st->g.nzz;
  return PCF_OK;
}

int pcf_get_time(const pcf_state* st, double* time_s, int32_t* itime) {
  if (!st || !time_s) return PCF_E_ARG;
  *time_s = st->time * st->tref;
  if (itime) *itime = st->itime;
  return PCF_OK;
}

int pcf_set_time(pcf_state* st, double time_s) {  // integrate.jl:168-196
  if (!st) return PCF_E_ARG;
  if (st->cfg.tmax < time_s) {
    set_error("Restart error: tmax < time!");
    return PCF_E_CONFIG;
  }
  st->time = time_s / st->tref;
  st->nextoutputtime = time_s + st->cfg.output_interval;
  return PCF_OK;
}

int pcf_initial_cleaning(pcf_state* st, int32_t* niter) {  // integrate.jl:136-156
  if (!st) return PCF_E_ARG;
  TRY(modify_compressible_wind(st, true));
  TRY(set_boundaries_predictands(st));
  int it = 0;
  int rc = apply_corrector(st, 1.0, 1.0, &it);
  if (niter) *niter = it;
  if (rc) return rc;
  TRY(modify_compressible_wind(st, false));
  TRY(set_boundaries_predictands(st));
  if (hipStreamSynchronize(st->stream) != hipSuccess) return PCF_E_HIP;
  return PCF_OK;
}

int pcf_compute_time_step(pcf_state* st, double* dt) {
  if (!st || !dt) return PCF_E_ARG;
  return time_step(st, dt);
}

int pcf_compute_sponges(pcf_state* st, double dt, double time) {
  if (!st) return PCF_E_ARG;
  return compute_sponges(st, dt, time);
}

int pcf_backup_predictands(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  double* cur[7];
  const int np = predictand_slots(st, cur);
  TRY(copy_set(st, st->p0, cur, np));
  if (st->tracer) TRY(copy_field(st, st->chi0, st->chi));  // backup_predictands.jl:18
  return PCF_OK;
}

int pcf_explicit_integration_lhs(pcf_state* st, double dt, double time) {
  if (!st) return PCF_E_ARG;
  return explicit_lhs(st, dt, time);
}

int pcf_explicit_integration_rhs(pcf_state* st, double dt, double time) {
  if (!st) return PCF_E_ARG;
  return explicit_rhs(st, dt, time);
}

int pcf_implicit_integration(pcf_state* st, double dt, double time, double rf, int32_t* niter) {
  if (!st) return PCF_E_ARG;
  int it = 0;
  int rc = implicit(st, dt, time, rf, &it);
  if (niter) *niter = it;
  return rc;
}

// One pass of the time loop of integrate.jl:217-378.
int pcf_step(pcf_state* st, pcf_step_info* info) {
  if (!st) return PCF_E_ARG;
  const pcf_config& n = st->cfg;
  pcf_step_info loc;
  if (!info) info = &loc;
  std::memset(info, 0, sizeof(*info));
  double dt;
  TRY(time_step(st, &dt));
  int output = 0;
  if (!n.output_steps && (st->time + dt) * st->tref + n.dtmin > st->nextoutputtime) {
    dt = st->nextoutputtime / st->tref - st->time;
    output = 1;
  }
  st->time += dt;
  const double time = st->time;
  TRY(compute_sponges(st, dt, time));
  TRY(synchronize_density_fluctuations(st));
  TRY(set_boundaries_predictands(st));
  double* cur[7];
  const int np = predictand_slots(st, cur);
  TRY(copy_set(st, st->p0, cur, np));  // backup_predictands
  if (st->tracer) TRY(copy_field(st, st->chi0, st->chi));
  TRY(compute_theta_fluxes(st));                              // integrate.jl:277
  TRY(explicit_lhs(st, 0.5 * dt, time));                      // (1)
  int it1 = 0, it2 = 0;
  st->solve_slot = 0;
  int rc = implicit(st, 0.5 * dt, time, 1.0, &it1);          // (2)
  st->solve_slot = -1;
  info->niter[0] = it1;
  if (rc) return rc;
  TRY(copy_set(st, st->p1, cur, np));                         // p1 = predictands
  // reset_predictands! (reset_predictands!.jl:71-118): rho (PI and
  // compressible), rhop, u, v, w; pip and p too in the compressible model,
  // else pip keeps its new value
  if (n.model != PCF_BOUSSINESQ) TRY(copy_field(st, st->rho, st->p0[0]));
  TRY(copy_field(st, st->rhop, st->p0[1]));
  TRY(copy_field(st, st->u, st->p0[2]));
  TRY(copy_field(st, st->v, st->p0[3]));
  TRY(copy_field(st, st->w, st->p0[4]));
  if (st->compressible) {
    TRY(copy_field(st, st->pip, st->p0[5]));
    TRY(copy_field(st, st->p, st->p0[6]));
  }
  if (st->tracer) TRY(copy_field(st, st->chi, st->chi0));     // reset_predictands!.jl:63-67
  TRY(explicit_rhs(st, 0.5 * dt, time));                      // (3)
  TRY(copy_set(st, st->p0, st->p1, np));                      // p0 = p1
  TRY(synchronize_compressible_atmosphere(st, st->p0[6]));   // pbar = p1.p
  TRY(explicit_lhs(st, dt, time));                            // (4)
  st->solve_slot = 1;
  rc = implicit(st, 0.5 * dt, time, 2.0, &it2);               // (5)
  st->solve_slot = -1;
  info->niter[1] = it2;
  if (rc) return rc;
  if (hipStreamSynchronize(st->stream) != hipSuccess) return PCF_E_HIP;
  prof_flush(st);
  st->itime += 1;
  info->time = time * st->tref;
  info->dt = dt * st->tref;
  info->output = output;
  if (output) {
    st->nextoutputtime += n.output_interval;
    if (st->nextoutputtime >= n.tmax) st->nextoutputtime = n.tmax;
  }
  return PCF_OK;
}

int pcf_set_boundaries(pcf_state* st, int which) {
  if (!st) return PCF_E_ARG;
  if (which == PCF_BOUNDARY_PREDICTANDS) return set_boundaries_predictands(st);
  // the reconstructions are evaluated inside the flux kernels from inputs
  // whose halo entries are already the periodic images: nothing to fill
  if (which == PCF_BOUNDARY_RECONSTRUCTIONS) return PCF_OK;
  if (which == PCF_BOUNDARY_FLUXES) return set_boundaries_fluxes(st);
  set_error("unknown boundary set");
  return PCF_E_OPTION;
}

int pcf_reconstruct(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  return reconstruct(st);
}

int pcf_compute_fluxes(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  return compute_fluxes(st);
}

int pcf_compute_theta_fluxes(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  return compute_theta_fluxes(st);
}

int pcf_modify_compressible_wind(pcf_state* st, int multiply) {
  if (!st) return PCF_E_ARG;
  return modify_compressible_wind(st, multiply != 0);
}

int pcf_synchronize_compressible_atmosphere(pcf_state* st) {
  if (!st) return PCF_E_ARG;
  return synchronize_compressible_atmosphere(st, st->p);
}

int pcf_update_lhs(pcf_state* st, double dt, int32_t m, int var) {
  if (!st || m < 1 || m > 3) return PCF_E_ARG;
  return update_lhs(st, dt, m, var);
}

int pcf_update_rhs(pcf_state* st, double dt, int var, int implicit_, double rf) {
  if (!st) return PCF_E_ARG;
  return update_rhs(st, dt, var, implicit_ != 0, rf);
}

int pcf_apply_lhs_sponge(pcf_state* st, double dt, double time, int var) {
  if (!st) return PCF_E_ARG;
  return apply_lhs_sponge(st, dt, time, var);
}

int pcf_apply_corrector(pcf_state* st, double dt, double rf, int32_t* niter) {
  if (!st) return PCF_E_ARG;
  int it = 0;
  int rc = apply_corrector(st, dt, rf, &it);
  if (niter) *niter = it;
  return rc;
}

int pcf_compute_lhs(pcf_state* st, double* tolref) {
  if (!st || !tolref) return PCF_E_ARG;
  return compute_lhs(st, tolref);
}

int pcf_compute_operator(pcf_state* st, double dt, double rf) {
  if (!st) return PCF_E_ARG;
  return compute_operator(st, dt, rf);
}

int pcf_apply_operator(pcf_state* st, const double* d_in, double* d_out, int kind) {
  if (!st || !d_in || !d_out) return PCF_E_ARG;
  if (kind != PCF_OP_TOTAL && kind != PCF_OP_HORIZONTAL) return PCF_E_OPTION;
  TRY(apply_operator(st, d_in, d_out, kind == PCF_OP_TOTAL));
  return hipStreamSynchronize(st->stream) == hipSuccess ? PCF_OK : PCF_E_HIP;
}

int pcf_apply_operator_transpose(pcf_state* st, const double* d_in, double* d_out) {
  if (!st || !d_in || !d_out || d_in == d_out) return PCF_E_ARG;
  TRY(apply_operator_transpose(st, d_in, d_out));
  return hipStreamSynchronize(st->stream) == hipSuccess ? PCF_OK : PCF_E_HIP;
}

int pcf_apply_preconditioner(pcf_state* st, const double* d_in, double* d_out) {
  if (!st || !d_in || !d_out || d_in == d_out) return PCF_E_ARG;
  TRY(apply_preconditioner(st, d_in, d_out));
  return hipStreamSynchronize(st->stream) == hipSuccess ? PCF_OK : PCF_E_HIP;
}

int pcf_apply_bicgstab(pcf_state* st, double tolref, int32_t* niter) {
  if (!st) return PCF_E_ARG;
  int it = 0;
  int rc = solve(st, tolref, 0.0, &it);
  if (niter) *niter = it;
  return rc;
}

int pcf_stage_tuning(const pcf_state* st, int32_t* shape, float* ms) {
  if (!st || !shape) return PCF_E_ARG;
  for (int q = 0; q < 14; ++q) {
    shape[q] = st->stage_shape[q];
    if (ms)
      for (int c = 0; c < 9; ++c) ms[9 * q + c] = st->stage_shape_ms[q][c];
  }
  return PCF_OK;
}

int pcf_set_solver_graphs(pcf_state* st, int32_t on) {
  if (!st) return PCF_E_ARG;
  st->solver_graphs = on != 0;
  return PCF_OK;
}

int pcf_set_stage_shape(pcf_state* st, int32_t slot, int32_t shape) {
  if (!st || slot < 0 || slot >= 14 || shape < -1 || shape >= 9) return PCF_E_ARG;
  st->stage_shape[slot] = shape;
  return PCF_OK;
}

int pcf_dot(pcf_state* st, const double* a, const double* b, double* result) {
  if (!st || !a || !b || !result) return PCF_E_ARG;
  return dot(st, a, b, result);
}

int pcf_get_field(pcf_state* st, int field, double* host_out) {
  if (!st || !host_out) return PCF_E_ARG;
  double* p = field_ptr(st, field);
  if (!p) { set_error("unknown field"); return PCF_E_ARG; }
  PCF_HIP(hipMemcpyAsync(host_out, p, st->g.nh() * sizeof(double), hipMemcpyDeviceToHost, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  return PCF_OK;
}

int pcf_set_field(pcf_state* st, int field, const double* host_in) {
  if (!st || !host_in) return PCF_E_ARG;
  double* p = field_ptr(st, field);
  if (!p) { set_error("unknown field"); return PCF_E_ARG; }
  PCF_HIP(hipMemcpyAsync(p, host_in, st->g.nh() * sizeof(double), hipMemcpyHostToDevice, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  st->lhs_div_valid[0] = st->lhs_div_valid[1] = false;  // fluxes or the Jacobian may have changed
  if (field == PCF_F_PBAR || field == PCF_F_RHOBAR) TRY(refresh_fcscal(st));
  return PCF_OK;
}

int pcf_get_vector(pcf_state* st, int vec, double* host_out) {
  if (!st || !host_out) return PCF_E_ARG;
  double* p = vector_ptr(st, vec);
  if (!p) { set_error("unknown vector"); return PCF_E_ARG; }
  PCF_HIP(hipMemcpyAsync(host_out, p, st->g.nv() * sizeof(double), hipMemcpyDeviceToHost, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  return PCF_OK;
}

int pcf_set_vector(pcf_state* st, int vec, const double* host_in) {
  if (!st || !host_in) return PCF_E_ARG;
  double* p = vector_ptr(st, vec);
  if (!p) { set_error("unknown vector"); return PCF_E_ARG; }
  PCF_HIP(hipMemcpyAsync(p, host_in, st->g.nv() * sizeof(double), hipMemcpyHostToDevice, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  if (vec == PCF_V_AC || vec == PCF_V_AD || vec == PCF_V_AU) st->pc_factors_valid = false;
  if (vec >= PCF_V_AC) st->t32_valid = false;
  return PCF_OK;
}

int pcf_last_residual(const pcf_state* st, double* rel) {
  if (!st || !rel) return PCF_E_ARG;
  *rel = st->last_residual;
  return PCF_OK;
}

}  // extern "C"
