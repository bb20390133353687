// Device-side pointwise functions shared by the kernels.  Each one evaluates
// the reference expression in the reference's operation order (the library
// is built with -ffp-contract=off, so a*b + c stays two rounded operations).
#pragma once

#include "pcf_internal.hpp"

namespace pcf {

// Read-only views of the fields the pointwise functions need.
struct Fields {
  const double *rho, *rhobar, *pbar, *jac, *m13, *m23, *m33, *n2;
  const double *u, *v, *w, *rhop;
  // compressible model (comp = 1): thetabar, P and the heat-conduction fluxes
  const double *thetabar, *p;
  const double* phith[3];
  int comp;
};

inline Fields make_fields(const pcf_state* st) {
  Fields F;
  F.rho = st->rho; F.rhobar = st->rhobar; F.pbar = st->pbar; F.jac = st->jac;
  F.m13 = st->met13; F.m23 = st->met23; F.m33 = st->met33; F.n2 = st->n2;
  F.u = st->u; F.v = st->v; F.w = st->w; F.rhop = st->rhop;
  F.thetabar = st->thetabar; F.p = st->p;
  for (int d = 0; d < 3; ++d) F.phith[d] = st->phitheta[d];
  F.comp = st->compressible ? 1 : 0;
  return F;
}

__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

// compute_compressible_wind_factor.jl:102-157: 1 outside the compressible
// model, else J P interpolated to the staggered wind points
__device__ __forceinline__ double wind_factor_u(const Fields& F, const Geo& g, int i, int j, int k) {
  if (!F.comp) return 1.0;
  size_t c = g.h(i, j, k), r = g.h(i + 1, j, k);
  return (F.jac[c] * F.p[c] + F.jac[r] * F.p[r]) / 2;
}
__device__ __forceinline__ double wind_factor_v(const Fields& F, const Geo& g, int i, int j, int k) {
  if (!F.comp) return 1.0;
  size_t c = g.h(i, j, k), f = g.h(i, j + 1, k);
  return (F.jac[c] * F.p[c] + F.jac[f] * F.p[f]) / 2;
}
__device__ __forceinline__ double wind_factor_w(const Fields& F, const Geo& g, int i, int j, int k) {
  if (!F.comp) return 1.0;
  size_t c = g.h(i, j, k), u = g.h(i, j, k + 1);
  return F.jac[c] * F.jac[u] * (F.p[c] + F.p[u]) / (F.jac[c] + F.jac[u]);
}

// conductive_heating.jl:67-117 (0 outside the compressible model)
__device__ __forceinline__ double conductive_heating(const Fields& F, const Geo& g, int i, int j, int k) {
  if (!F.comp) return 0.0;
  const size_t c = g.h(i, j, k);
  double rhotot = (F.rho[c] + F.rhobar[c]) / F.jac[c];
  return -rhotot * ((F.phith[0][c] - F.phith[0][g.h(i - 1, j, k)]) / g.dx +
                    (F.phith[1][c] - F.phith[1][g.h(i, j - 1, k)]) / g.dy +
                    (F.phith[2][c] - F.phith[2][g.h(i, j, k - 1)]) / g.dz);
}

// dP/dpi (update!.jl:1385-1386) with the host-evaluated constant factor
__device__ __forceinline__ double dpdpi(const Phys& ph, double p) {
  return ph.dpdpi_c * pow(p, 2 - ph.gamma);
}

// compute_buoyancy_factor.jl:108-133 (RhoP)
__device__ __forceinline__ double buoyancy_factor_rhop(const Fields& F, const Geo& g, int i, int j, int k) {
  size_t c = g.h(i, j, k);
  if (F.comp) return F.pbar[c] / F.thetabar[c] / (F.rho[c] + F.rhobar[c]);
  return F.rhobar[c] / (F.rho[c] + F.rhobar[c]);
}

// compute_buoyancy_factor.jl:135-172 (W)
__device__ __forceinline__ double buoyancy_factor_w(const Fields& F, const Geo& g, int i, int j, int k) {
  size_t c = g.h(i, j, k), u = g.h(i, j, k + 1);
  if (F.comp)
    return (F.jac[u] * F.pbar[c] / F.thetabar[c] + F.jac[c] * F.pbar[u] / F.thetabar[u]) /
           (F.jac[u] * (F.rho[c] + F.rhobar[c]) + F.jac[c] * (F.rho[u] + F.rhobar[u]));
  return (F.jac[u] * F.rhobar[c] + F.jac[c] * F.rhobar[u]) /
         (F.jac[u] * (F.rho[c] + F.rhobar[c]) + F.jac[c] * (F.rho[u] + F.rhobar[u]));
}

// compute_pressure_gradient.jl:106-156 (U)
__device__ __forceinline__ double pressure_gradient_u(const Fields& F, const Phys& ph, const Geo& g, const double* pip,
                                                      int i, int j, int k) {
  size_t c = g.h(i, j, k), r = g.h(i + 1, j, k);
  double rhoedger = 0.5 * (F.rho[c] + F.rho[r]);
  double rhobaredger = 0.5 * (F.rhobar[c] + F.rhobar[r]);
  rhoedger += rhobaredger;
  double pedger = 0.5 * (F.pbar[c] + F.pbar[r]);
  double met13edger = 0.5 * (F.m13[c] + F.m13[r]);
  double pre = ph.kappainv * ph.mainv2 * pedger / rhoedger;
  if (k == g.k0) {
    double pipuuedger = 0.5 * (pip[g.h(i, j, k + 2)] + pip[g.h(i + 1, j, k + 2)]);
    double pipuedger = 0.5 * (pip[g.h(i, j, k + 1)] + pip[g.h(i + 1, j, k + 1)]);
    double pipedger = 0.5 * (pip[c] + pip[r]);
    return pre * ((pip[r] - pip[c]) / g.dx + met13edger * (-pipuuedger + 4.0 * pipuedger - 3.0 * pipedger) * 0.5 / g.dz);
  } else if (k == g.z_size + g.nbz) {
    double pipddedger = 0.5 * (pip[g.h(i, j, k - 2)] + pip[g.h(i + 1, j, k - 2)]);
    double pipdedger = 0.5 * (pip[g.h(i, j, k - 1)] + pip[g.h(i + 1, j, k - 1)]);
    double pipedger = 0.5 * (pip[c] + pip[r]);
    return pre * ((pip[r] - pip[c]) / g.dx + met13edger * (pipddedger - 4.0 * pipdedger + 3.0 * pipedger) * 0.5 / g.dz);
  }
  double pipuedger = 0.5 * (pip[g.h(i, j, k + 1)] + pip[g.h(i + 1, j, k + 1)]);
  double pipdedger = 0.5 * (pip[g.h(i, j, k - 1)] + pip[g.h(i + 1, j, k - 1)]);
  return pre * ((pip[r] - pip[c]) / g.dx + met13edger * (pipuedger - pipdedger) * 0.5 / g.dz);
}

// compute_pressure_gradient.jl:158-209 (V)
__device__ __forceinline__ double pressure_gradient_v(const Fields& F, const Phys& ph, const Geo& g, const double* pip,
                                                      int i, int j, int k) {
  size_t c = g.h(i, j, k), f = g.h(i, j + 1, k);
  double rhoedgef = 0.5 * (F.rho[c] + F.rho[f]);
  double rhobaredgef = 0.5 * (F.rhobar[c] + F.rhobar[f]);
  rhoedgef += rhobaredgef;
  double pedgef = 0.5 * (F.pbar[c] + F.pbar[f]);
  double met23edgef = 0.5 * (F.m23[c] + F.m23[f]);
  double pre = ph.kappainv * ph.mainv2 * pedgef / rhoedgef;
  if (k == g.k0) {
    double pipuuedgef = 0.5 * (pip[g.h(i, j, k + 2)] + pip[g.h(i, j + 1, k + 2)]);
    double pipuedgef = 0.5 * (pip[g.h(i, j, k + 1)] + pip[g.h(i, j + 1, k + 1)]);
    double pipedgef = 0.5 * (pip[c] + pip[f]);
    return pre * ((pip[f] - pip[c]) / g.dy + met23edgef * (-pipuuedgef + 4.0 * pipuedgef - 3.0 * pipedgef) * 0.5 / g.dz);
  } else if (k == g.z_size + g.nbz) {
    double pipddedgef = 0.5 * (pip[g.h(i, j, k - 2)] + pip[g.h(i, j + 1, k - 2)]);
    double pipdedgef = 0.5 * (pip[g.h(i, j, k - 1)] + pip[g.h(i, j + 1, k - 1)]);
    double pipedgef = 0.5 * (pip[c] + pip[f]);
    return pre * ((pip[f] - pip[c]) / g.dy + met23edgef * (pipddedgef - 4.0 * pipdedgef + 3.0 * pipedgef) * 0.5 / g.dz);
  }
  double pipuedgef = 0.5 * (pip[g.h(i, j, k + 1)] + pip[g.h(i, j + 1, k + 1)]);
  double pipdedgef = 0.5 * (pip[g.h(i, j, k - 1)] + pip[g.h(i, j + 1, k - 1)]);
  return pre * ((pip[f] - pip[c]) / g.dy + met23edgef * (pipuedgef - pipdedgef) * 0.5 / g.dz);
}

// compute_pressure_gradient.jl:211-297 (W)
__device__ __forceinline__ double pressure_gradient_w(const Fields& F, const Phys& ph, const Geo& g, const double* pip,
                                                      int i, int j, int k) {
  size_t c = g.h(i, j, k), up = g.h(i, j, k + 1);
  const double jc = F.jac[c], ju = F.jac[up];
  double rhoedgeu = (ju * F.rho[c] + jc * F.rho[up]) / (jc + ju);
  rhoedgeu += (ju * F.rhobar[c] + jc * F.rhobar[up]) / (jc + ju);
  double pedgeu = (ju * F.pbar[c] + jc * F.pbar[up]) / (jc + ju);
  double met13edgeu = (ju * F.m13[c] + jc * F.m13[up]) / (jc + ju);
  double met23edgeu = (ju * F.m23[c] + jc * F.m23[up]) / (jc + ju);
  double met33edgeu = (ju * F.m33[c] + jc * F.m33[up]) / (jc + ju);
  size_t r = g.h(i + 1, j, k), ru = g.h(i + 1, j, k + 1), l = g.h(i - 1, j, k), lu = g.h(i - 1, j, k + 1);
  size_t f = g.h(i, j + 1, k), fu = g.h(i, j + 1, k + 1), b = g.h(i, j - 1, k), bu = g.h(i, j - 1, k + 1);
  double pipredgeu = (F.jac[ru] * pip[r] + F.jac[r] * pip[ru]) / (F.jac[r] + F.jac[ru]);
  double pipledgeu = (F.jac[lu] * pip[l] + F.jac[l] * pip[lu]) / (F.jac[l] + F.jac[lu]);
  double pipfedgeu = (F.jac[fu] * pip[f] + F.jac[f] * pip[fu]) / (F.jac[f] + F.jac[fu]);
  double pipbedgeu = (F.jac[bu] * pip[b] + F.jac[b] * pip[bu]) / (F.jac[b] + F.jac[bu]);
  return ph.kappainv * ph.mainv2 * pedgeu / rhoedgeu *
         (met13edgeu * (pipredgeu - pipledgeu) * 0.5 / g.dx + met23edgeu * (pipfedgeu - pipbedgeu) * 0.5 / g.dy +
          met33edgeu * (pip[up] - pip[c]) / g.dz);
}

// transform.jl:86-122 (Cartesian) inside compute_vertical_wind.jl:27-63
__device__ __forceinline__ double vertical_wind(const Fields& F, const Geo& g, int i, int j, int k) {
  size_t c = g.h(i, j, k), up = g.h(i, j, k + 1);
  const double jc = F.jac[c], ju = F.jac[up];
  double jacedgeu = 2.0 * jc * ju / (jc + ju);
  double uc = 0.5 * (F.u[c] + F.u[g.h(i - 1, j, k)]);
  double uu = 0.5 * (F.u[up] + F.u[g.h(i - 1, j, k + 1)]);
  double vc = 0.5 * (F.v[c] + F.v[g.h(i, j - 1, k)]);
  double vu = 0.5 * (F.v[up] + F.v[g.h(i, j - 1, k + 1)]);
  return jacedgeu * (-(ju * (F.m13[c] * uc + F.m23[c] * vc) + jc * (F.m13[up] * uu + F.m23[up] * vu)) / (jc + ju) +
                     F.w[c]);
}

// transform.jl:124-157 (Transformed)
__device__ __forceinline__ double transform_transformed(const Fields& F, const Geo& g, int i, int j, int k,
                                                        double uedger, double uuedger, double uedgel, double uuedgel,
                                                        double vedgef, double vuedgef, double vedgeb, double vuedgeb,
                                                        double wedgeu) {
  size_t c = g.h(i, j, k), up = g.h(i, j, k + 1);
  const double jc = F.jac[c], ju = F.jac[up];
  double jacedgeu = 2.0 * jc * ju / (jc + ju);
  double uc = 0.5 * (uedger + uedgel), uu = 0.5 * (uuedger + uuedgel);
  double vc = 0.5 * (vedgef + vedgeb), vu = 0.5 * (vuedgef + vuedgeb);
  return (ju * (F.m13[c] * uc + F.m23[c] * vc) + jc * (F.m13[up] * uu + F.m23[up] * vu)) / (jc + ju) +
         wedgeu / jacedgeu;
}

// Flat thread index -> (i, j, k) of a range.
__device__ __forceinline__ bool range_ijk(const Range3& R, long long t, int& i, int& j, int& k) {
  if (t >= R.count()) return false;
  long long plane = (long long)R.ni * R.nj;
  int kk = (int)(t / plane);
  long long r = t - kk * plane;
  int jj = (int)(r / R.ni);
  i = R.i0 + (int)(r - (long long)jj * R.ni);
  j = R.j0 + jj;
  k = R.k0 + kk;
  return true;
}

}  // namespace pcf
