// ORACLE -- test infrastructure only (see orc.hpp).
// src/Integration/*.jl and a C API (orc_*) that tests/ and bench.py load
// with ctypes.  The oracle never touches a GPU.
#include <cfloat>
#include <cstring>
#include <omp.h>

#include "orc.hpp"

namespace orc {

void copy_predictands(Predictands& d, const Predictands& s) {
  d.rho.a = s.rho.a;
  d.rhop.a = s.rhop.a;
  d.u.a = s.u.a;
  d.v.a = s.v.a;
  d.w.a = s.w.a;
  d.pip.a = s.pip.a;
  d.p.a = s.p.a;
}

// modify_compressible_wind!.jl:52-80: the winds times (multiply) or divided
// by J P interpolated to their points; nothing outside the compressible model
void modify_compressible_wind(State& st, bool multiply) {
  if (!st.compressible) return;
  Arr3 &u = st.pr.u, &v = st.pr.v, &w = st.pr.w, &p = st.pr.p, &jac = st.jac;
  auto op = [multiply](double a, double b) { return multiply ? a * b : a / b; };
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        u(i, j, k) = op(u(i, j, k), (jac(i, j, k) * p(i, j, k) + jac(i + 1, j, k) * p(i + 1, j, k)) / 2);
        v(i, j, k) = op(v(i, j, k), (jac(i, j, k) * p(i, j, k) + jac(i, j + 1, k) * p(i, j + 1, k)) / 2);
        w(i, j, k) = op(w(i, j, k), jac(i, j, k) * jac(i, j, k + 1) * (p(i, j, k) + p(i, j, k + 1)) /
                                        (jac(i, j, k) + jac(i, j, k + 1)));
      }
}

// synchronize_compressible_atmosphere!.jl:61-72: pbar .= p (compressible only)
void synchronize_compressible_atmosphere(State& st, const Predictands& p) {
  if (st.compressible) st.pbar.a = p.p.a;
}

// compute_time_step.jl:46-153 (NoWKB)
int compute_time_step(State& st, double& dt) {
  const pcf_config& n = st.cfg;
  const double tref = st.c.tref;
  if (!n.adaptive_time_step) {
    dt = n.dtmax / tref;
  } else {
    double umax = 0.0, vmax = 0.0, wmax = 0.0;
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          umax = std::max(umax, std::fabs(st.pr.u(i, j, k)));
          vmax = std::max(vmax, std::fabs(st.pr.v(i, j, k)));
          wmax = std::max(wmax, std::fabs(st.pr.w(i, j, k)));
        }
    umax += DBL_EPSILON;
    vmax += DBL_EPSILON;
    wmax += DBL_EPSILON;
    double dtconv = n.cfl_number * std::min(std::min(st.dx / umax, st.dy / vmax), st.dz / wmax);
    double dtvisc = 0.5 * std::min(std::min(st.dx * st.dx, st.dy * st.dy), st.dzcmin * st.dzcmin) * st.c.re;
    dt = std::min(std::min(dtvisc, dtconv), n.dtmax / tref);
  }
  if (dt * tref < n.dtmin) {
    st.err = "Error in compute_time_step: dt < dtmin";
    return PCF_E_DTMIN;
  }
  return PCF_OK;
}

static void save_backups(State& st, bool rho, bool rhop, bool u, bool v, bool w) {
  if (rho) st.rhoold.a = st.pr.rho.a;
  if (rhop) st.rhopold.a = st.pr.rhop.a;
  if (u) st.uold.a = st.pr.u.a;
  if (v) st.vold.a = st.pr.v.a;
  if (w) st.wold.a = st.pr.w.a;
}

// explicit_integration!.jl:42-101
int explicit_integration_lhs(State& st, Predictands& p0, double dt, double time) {
  for (int m = 1; m <= st.nstages; ++m) {
    const double sdt = st.stepfrac[m - 1] * dt;
    reconstruct(st);
    set_boundaries_reconstructions(st);
    compute_fluxes(st, p0);
    set_boundaries_fluxes(st);
    save_backups(st, true, false, true, true, false);
    update_lhs(st, dt, m, PCF_VAR_RHO);
    apply_lhs_sponge(st, sdt, time, PCF_VAR_RHO);
    update_lhs(st, dt, m, PCF_VAR_RHOP);
    apply_lhs_sponge(st, sdt, time, PCF_VAR_RHOP);
    update_lhs(st, dt, m, PCF_VAR_P);  // no-ops outside the compressible model
    apply_lhs_sponge(st, sdt, time, PCF_VAR_P);
    update_lhs(st, dt, m, PCF_VAR_CHI);  // no-ops without a tracer
    apply_lhs_sponge(st, sdt, time, PCF_VAR_CHI);
    set_boundaries_predictands(st);
    update_lhs(st, dt, m, PCF_VAR_U);
    update_lhs(st, dt, m, PCF_VAR_V);
    update_lhs(st, dt, m, PCF_VAR_W);
    apply_lhs_sponge(st, sdt, time, PCF_VAR_U);
    apply_lhs_sponge(st, sdt, time, PCF_VAR_V);
    apply_lhs_sponge(st, sdt, time, PCF_VAR_W);
    set_boundaries_predictands(st);
  }
  synchronize_compressible_atmosphere(st, st.pr);
  apply_lhs_sponge(st, dt, time, PCF_VAR_PIP);
  return PCF_OK;
}

// explicit_integration!.jl:103-131
int explicit_integration_rhs(State& st, double dt, double time) {
  (void)time;
  synchronize_compressible_atmosphere(st, st.pr);
  modify_compressible_wind(st, true);
  set_boundaries_predictands(st);
  save_backups(st, false, true, true, true, true);
  update_rhs(st, dt, PCF_VAR_RHOP, false, 0.0);
  update_rhs(st, dt, PCF_VAR_U, false, 0.0);
  update_rhs(st, dt, PCF_VAR_V, false, 0.0);
  update_rhs(st, dt, PCF_VAR_W, false, 0.0);
  update_rhs(st, dt, PCF_VAR_PIP, false, 0.0);
  modify_compressible_wind(st, false);
  set_boundaries_predictands(st);
  return PCF_OK;
}

// implicit_integration!.jl:37-79
int implicit_integration(State& st, double dt, double time, double rf, int& niter) {
  (void)time;
  modify_compressible_wind(st, true);
  set_boundaries_predictands(st);
  save_backups(st, false, false, false, false, true);
  update_rhs(st, dt, PCF_VAR_U, true, rf);
  update_rhs(st, dt, PCF_VAR_V, true, rf);
  update_rhs(st, dt, PCF_VAR_W, true, rf);
  set_boundaries_predictands(st);
  update_rhs(st, dt, PCF_VAR_RHOP, true, rf);
  set_boundaries_predictands(st);
  int rc = apply_corrector(st, dt, rf, niter);
  if (rc) return rc;
  modify_compressible_wind(st, false);
  set_boundaries_predictands(st);
  return PCF_OK;
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// C API for ctypes (tests/, bench.py cpu_baseline)
// ---------------------------------------------------------------------------
extern "C" {

struct orc_step_info {
  double time, dt;
  int32_t niter[2];
  int32_t output;
  int32_t pad;
};

const char* orc_last_error(State* st) { return st ? st->err.c_str() : ""; }

State* orc_create(const pcf_config* cfg, int* rc) {
  State* st = new State();
  *rc = setup(*st, *cfg);
  // integrate.jl:113-115 (time = 0, nextoutputtime = output_interval)
  st->time = 0.0;
  st->nextoutputtime = cfg->output_interval;
  return st;
}

void orc_destroy(State* st) { delete st; }

void orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

void orc_dims(State* st, int32_t* d) {
  d[0] = st->nx; d[1] = st->ny; d[2] = st->nz;
  d[3] = st->nxx; d[4] = st->nyy; d[5] = st->nzz;
}

// integrate.jl:136-156
int orc_initial_cleaning(State* st, int32_t* niter) {
  modify_compressible_wind(*st, true);
  set_boundaries_predictands(*st);
  int it = 0;
  int rc = apply_corrector(*st, 1.0, 1.0, it);
  *niter = it;
  if (rc) return rc;
  modify_compressible_wind(*st, false);
  set_boundaries_predictands(*st);
  return PCF_OK;
}

// One pass of the time loop, integrate.jl:217-378 (output handled by the caller).
int orc_step(State* st, orc_step_info* info) {
  const pcf_config& n = st->cfg;
  const double tref = st->c.tref;
  double dt;
  int rc = compute_time_step(*st, dt);
  if (rc) return rc;
  int output = 0;
  if (!n.output_steps) {
    if ((st->time + dt) * tref + n.dtmin > st->nextoutputtime) {
      dt = st->nextoutputtime / tref - st->time;
      output = 1;
    }
  }
  st->time += dt;
  const double time = st->time;
  compute_sponges(*st, dt, time);
  // synchronize_density_fluctuations! (synchronize_density_fluctuations!.jl:65-86)
  if (n.model == PCF_PSEUDO_INCOMPRESSIBLE) st->pr.rhop.a = st->pr.rho.a;
  if (n.model == PCF_COMPRESSIBLE)
    for (size_t q = 0; q < st->pr.rhop.a.size(); ++q)
      st->pr.rhop.a[q] = st->pr.rho.a[q] + st->rhobar.a[q] - st->pbar.a[q] / st->thetabar.a[q];
  set_boundaries_predictands(*st);
  copy_predictands(st->p0, st->pr);
  if (st->tracer) st->chi0.a = st->chi.a;  // backup_predictands.jl:18
  compute_theta_fluxes(*st, st->p0);  // integrate.jl:277
  // (1)
  explicit_integration_lhs(*st, st->p0, 0.5 * dt, time);
  // (2)
  int it1 = 0, it2 = 0;
  rc = implicit_integration(*st, 0.5 * dt, time, 1.0, it1);
  info->niter[0] = it1;
  if (rc) return rc;
  copy_predictands(st->p1, st->pr);
  // (3) reset_predictands! (reset_predictands!.jl:71-118: rho, rhop, u, v, w;
  // pip and p too in the compressible model, else pip keeps its new value)
  if (n.model != PCF_BOUSSINESQ) st->pr.rho.a = st->p0.rho.a;
  st->pr.rhop.a = st->p0.rhop.a;
  st->pr.u.a = st->p0.u.a;
  st->pr.v.a = st->p0.v.a;
  st->pr.w.a = st->p0.w.a;
  if (n.model == PCF_COMPRESSIBLE) {
    st->pr.pip.a = st->p0.pip.a;
    st->pr.p.a = st->p0.p.a;
  }
  if (st->tracer) st->chi.a = st->chi0.a;  // reset_predictands!.jl:63-67
  explicit_integration_rhs(*st, 0.5 * dt, time);
  // (4)
  copy_predictands(st->p0, st->p1);
  synchronize_compressible_atmosphere(*st, st->p0);
  explicit_integration_lhs(*st, st->p0, dt, time);
  // (5)
  rc = implicit_integration(*st, 0.5 * dt, time, 2.0, it2);
  info->niter[1] = it2;
  if (rc) return rc;
  st->itime += 1;
  info->time = time * tref;
  info->dt = dt * tref;
  info->output = output;
  if (output) {  // integrate.jl:354-361
    st->nextoutputtime = st->nextoutputtime + n.output_interval;
    if (st->nextoutputtime >= n.tmax) st->nextoutputtime = n.tmax;
  }
  return PCF_OK;
}

double orc_time(State* st) { return st->time * st->c.tref; }

// Restart (integrate.jl:168-200): the model time read from the record and
// the next output time one output_interval later; tmax < time is an error.
int orc_set_time(State* st, double time_s) {
  if (st->cfg.tmax < time_s) {
    st->err = "Restart error: tmax < time!";
    return PCF_E_CONFIG;
  }
  st->time = time_s / st->c.tref;
  st->nextoutputtime = time_s + st->cfg.output_interval;
  return PCF_OK;
}

static Arr3* field_ptr(State* st, int f) {
  Predictands& p = st->pr;
  switch (f) {
    case PCF_F_RHO: return &p.rho;
    case PCF_F_RHOP: return &p.rhop;
    case PCF_F_U: return &p.u;
    case PCF_F_V: return &p.v;
    case PCF_F_W: return &p.w;
    case PCF_F_PIP: return &p.pip;
    case PCF_F_PBAR: return &st->pbar;
    case PCF_F_THETABAR: return &st->thetabar;
    case PCF_F_RHOBAR: return &st->rhobar;
    case PCF_F_N2: return &st->n2;
    case PCF_F_JAC: return &st->jac;
    case PCF_F_MET13: return &st->met13;
    case PCF_F_MET23: return &st->met23;
    case PCF_F_MET33: return &st->met33;
    case PCF_F_ZC: return &st->zc;
    case PCF_F_ZCTILDE: return &st->zctilde;
    case PCF_F_ALPHAR: return &st->alphar;
    case PCF_F_BETAR: return &st->betar;
    case PCF_F_DPIP: return &st->dpip;
    case PCF_F_CORX: return &st->corx;
    case PCF_F_CORY: return &st->cory;
    case PCF_F_DRHO: return &st->drho;
    case PCF_F_DRHOP: return &st->drhop;
    case PCF_F_DU: return &st->du;
    case PCF_F_DV: return &st->dv;
    case PCF_F_DW: return &st->dw;
    case PCF_F_RHOOLD: return &st->rhoold;
    case PCF_F_RHOPOLD: return &st->rhopold;
    case PCF_F_UOLD: return &st->uold;
    case PCF_F_VOLD: return &st->vold;
    case PCF_F_WOLD: return &st->wold;
    case PCF_F_P0_RHO: return &st->p0.rho;
    case PCF_F_P0_RHOP: return &st->p0.rhop;
    case PCF_F_P0_U: return &st->p0.u;
    case PCF_F_P0_V: return &st->p0.v;
    case PCF_F_P0_W: return &st->p0.w;
    case PCF_F_PHIRHO_X: case PCF_F_PHIRHO_Y: case PCF_F_PHIRHO_Z:
      return &st->phirho.c[f - PCF_F_PHIRHO_X];
    case PCF_F_PHIRHOP_X: case PCF_F_PHIRHOP_Y: case PCF_F_PHIRHOP_Z:
      return &st->phirhop.c[f - PCF_F_PHIRHOP_X];
    case PCF_F_PHIU_X: case PCF_F_PHIU_Y: case PCF_F_PHIU_Z:
      return &st->phiu.c[f - PCF_F_PHIU_X];
    case PCF_F_PHIV_X: case PCF_F_PHIV_Y: case PCF_F_PHIV_Z:
      return &st->phiv.c[f - PCF_F_PHIV_X];
    case PCF_F_PHIW_X: case PCF_F_PHIW_Y: case PCF_F_PHIW_Z:
      return &st->phiw.c[f - PCF_F_PHIW_X];
    case PCF_F_CHI: return st->tracer ? &st->chi : nullptr;
    case PCF_F_DCHI: return st->tracer ? &st->dchi : nullptr;
    case PCF_F_INITIALTRACER: return st->tracer ? &st->initialtracer : nullptr;
    case PCF_F_CHI0: return st->tracer ? &st->chi0 : nullptr;
    case PCF_F_PHICHI_X: case PCF_F_PHICHI_Y: case PCF_F_PHICHI_Z:
      return st->tracer ? &st->phichi.c[f - PCF_F_PHICHI_X] : nullptr;
    case PCF_F_P: return st->compressible ? &p.p : nullptr;
    case PCF_F_DP: return st->compressible ? &st->dp : nullptr;
    case PCF_F_P0_P: return st->compressible ? &st->p0.p : nullptr;
    case PCF_F_PHIP_X: case PCF_F_PHIP_Y: case PCF_F_PHIP_Z:
      return st->compressible ? &st->phip.c[f - PCF_F_PHIP_X] : nullptr;
    case PCF_F_PHITHETA_X: case PCF_F_PHITHETA_Y: case PCF_F_PHITHETA_Z:
      return st->compressible ? &st->phitheta.c[f - PCF_F_PHITHETA_X] : nullptr;
    default: return nullptr;
  }
}

static Arr3* vector_ptr(State* st, int v) {
  if (v == PCF_V_RHS) return &st->rhs;
  if (v == PCF_V_SOLUTION) return &st->solution;
  if (v >= PCF_V_AC && v < PCF_V_COUNT) return &st->tensor[v - PCF_V_AC];
  return nullptr;
}

int orc_get_field(State* st, int f, double* out) {
  Arr3* a = field_ptr(st, f);
  if (!a) return PCF_E_ARG;
  std::memcpy(out, a->a.data(), a->a.size() * sizeof(double));
  return PCF_OK;
}

int orc_set_field(State* st, int f, const double* in) {
  Arr3* a = field_ptr(st, f);
  if (!a) return PCF_E_ARG;
  std::memcpy(a->a.data(), in, a->a.size() * sizeof(double));
  return PCF_OK;
}

int orc_get_vector(State* st, int v, double* out) {
  Arr3* a = vector_ptr(st, v);
  if (!a) return PCF_E_ARG;
  std::memcpy(out, a->a.data(), a->a.size() * sizeof(double));
  return PCF_OK;
}

int orc_set_vector(State* st, int v, const double* in) {
  Arr3* a = vector_ptr(st, v);
  if (!a) return PCF_E_ARG;
  std::memcpy(a->a.data(), in, a->a.size() * sizeof(double));
  return PCF_OK;
}

// Single-function entry points used by the parity tests.
int orc_compute_lhs(State* st, double* tolref) { return compute_lhs(*st, *tolref); }
void orc_compute_operator(State* st, double dt, double rf) { compute_operator(*st, dt, rf); }

void orc_apply_operator(State* st, const double* in, double* out, int kind) {
  Arr3 a, b;
  a.init(st->nx, st->ny, st->nz);
  b.init(st->nx, st->ny, st->nz);
  std::memcpy(a.a.data(), in, a.a.size() * sizeof(double));
  apply_operator(*st, a, b, kind == PCF_OP_TOTAL);
  std::memcpy(out, b.a.data(), b.a.size() * sizeof(double));
}

void orc_apply_preconditioner(State* st, const double* in, double* out) {
  Arr3 a, b;
  a.init(st->nx, st->ny, st->nz);
  b.init(st->nx, st->ny, st->nz);
  std::memcpy(a.a.data(), in, a.a.size() * sizeof(double));
  apply_preconditioner(*st, a, b);
  std::memcpy(out, b.a.data(), b.a.size() * sizeof(double));
}

int orc_apply_bicgstab(State* st, double tolref, int32_t* niter) {
  int it = 0;
  int rc = apply_bicgstab(*st, tolref, it);
  *niter = it;
  return rc;
}

double orc_last_residual(State* st) { return st->last_residual; }

int orc_apply_corrector(State* st, double dt, double rf, int32_t* niter) {
  int it = 0;
  int rc = apply_corrector(*st, dt, rf, it);
  *niter = it;
  return rc;
}

void orc_correct(State* st, double dt, double rf) { correct(*st, dt, rf); }
void orc_solution_to_dpip(State* st, double dt) { solution_to_dpip(*st, 1.0 / dt); }
void orc_compute_sponges(State* st, double dt, double time) { compute_sponges(*st, dt, time); }
void orc_set_boundaries(State* st, int which) {
  if (which == PCF_BOUNDARY_PREDICTANDS) set_boundaries_predictands(*st);
  else if (which == PCF_BOUNDARY_FLUXES) set_boundaries_fluxes(*st);
}
void orc_backup_predictands(State* st) {
  copy_predictands(st->p0, st->pr);
  if (st->tracer) st->chi0.a = st->chi.a;
}
// reconstruct! + set_boundaries!(Reconstructions) + compute_fluxes!(p0) + set_boundaries!(Fluxes)
void orc_compute_fluxes(State* st) {
  reconstruct(*st);
  set_boundaries_reconstructions(*st);
  compute_fluxes(*st, st->p0);
  set_boundaries_fluxes(*st);
}
void orc_save_backups(State* st) { save_backups(*st, true, true, true, true, true); }
// compute_fluxes!(state, p0, Theta()) with the p0 copy
void orc_compute_theta_fluxes(State* st) { compute_theta_fluxes(*st, st->p0); }
void orc_modify_compressible_wind(State* st, int multiply) { modify_compressible_wind(*st, multiply != 0); }
void orc_synchronize_compressible_atmosphere(State* st) { synchronize_compressible_atmosphere(*st, st->pr); }
void orc_update_lhs(State* st, double dt, int m, int var) { update_lhs(*st, dt, m, var); }
void orc_update_rhs(State* st, double dt, int var, int implicit, double rf) {
  update_rhs(*st, dt, var, implicit != 0, rf);
}
void orc_apply_lhs_sponge(State* st, double dt, double time, int var) { apply_lhs_sponge(*st, dt, time, var); }
int orc_explicit_integration_lhs(State* st, double dt, double time) {
  return explicit_integration_lhs(*st, st->p0, dt, time);
}
int orc_explicit_integration_rhs(State* st, double dt, double time) { return explicit_integration_rhs(*st, dt, time); }
int orc_implicit_integration(State* st, double dt, double time, double rf, int32_t* niter) {
  int it = 0;
  int rc = implicit_integration(*st, dt, time, rf, it);
  *niter = it;
  return rc;
}
int orc_compute_time_step(State* st, double* dt) { return compute_time_step(*st, *dt); }
double orc_vertical_wind(State* st, int i, int j, int k) { return vertical_wind(*st, i, j, k); }
double orc_stress_tensor(State* st, int i, int j, int k, int mu, int nu) {
  return stress_tensor(*st, i, j, k, mu, nu);
}
void orc_constants(State* st, double* out) {
  // lref, tref, uref, rhoref, thetaref, pref, re, g_ndim, kappa, dx, dy, dz, dzcmin
  const Constants& c = st->c;
  double v[] = {c.lref, c.tref, c.uref, c.rhoref, c.thetaref, c.pref, c.re, c.g_ndim, c.kappa,
                st->dx, st->dy, st->dz, st->dzcmin};
  std::memcpy(out, v, sizeof v);
}
void orc_grid_1d(State* st, double* x, double* y) {
  std::memcpy(x, st->x.data() + 1, st->nxx * sizeof(double));
  std::memcpy(y, st->y.data() + 1, st->nyy * sizeof(double));
}

}  // extern "C"

extern "C" int orc_config_size(void) { return (int)sizeof(pcf_config); }
