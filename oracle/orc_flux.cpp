// ORACLE -- test infrastructure only (see orc.hpp).
// src/FluxCalculator/*.jl: MUSCL reconstruction (MC-variant limiter) and the
// advective plus viscous fluxes, kinematic_diffusivity = 0.
#include <algorithm>
#include <cfloat>

#include "orc.hpp"

namespace orc {

// apply_1d_muscl!.jl:51-93 on a strided line: phi(q) and tilde(q, s) for
// q = 1..n are accessed through element pointers and a stride.
static void muscl_line(const double* phi, double* t1, double* t2, size_t stride, int n) {
  for (int q = 0; q < n; ++q) {
    t1[q * stride] = 1000.0;
    t2[q * stride] = 1000.0;
  }
  for (int q = 1; q < n - 1; ++q) {
    double p = phi[q * stride];
    double deltal = p - phi[(q - 1) * stride];
    double deltar = phi[(q + 1) * stride] - p;
    if (deltar == 0.0) {
      t2[q * stride] = p;
      t1[q * stride] = p;
    } else if (deltal == 0.0) {
      double theta = deltal / deltar;
      double s = (2.0 + theta) / 3.0;
      double sigmal = std::max(0.0, std::min(std::min(2 * theta, s), 2.0));
      t2[q * stride] = p + 0.5 * sigmal * deltar;
      t1[q * stride] = p;
    } else {
      double theta = deltal / deltar;
      double s = (2.0 + theta) / 3.0;
      double sigmal = std::max(0.0, std::min(std::min(2 * theta, s), 2.0));
      s = (2.0 + 1.0 / theta) / 3.0;
      double sigmar = std::max(0.0, std::min(std::min(2 / theta, s), 2.0));
      t2[q * stride] = p + 0.5 * sigmal * deltar;
      t1[q * stride] = p - 0.5 * sigmar * deltal;
    }
  }
}

// apply_3d_muscl!.jl:35-75
static void muscl_3d(State& st, Arr3& phi, Arr5& tilde) {
  const int nxx = st.nxx, nyy = st.nyy, nzz = st.nzz;
  const size_t sx = 1, sy = (size_t)nxx, sz = (size_t)nxx * nyy;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nzz - 1; ++k)
    for (int j = 2; j <= nyy - 1; ++j)
      muscl_line(&phi(1, j, k), &tilde(1, j, k, 1, 1), &tilde(1, j, k, 1, 2), sx, nxx);
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nzz - 1; ++k)
    for (int i = 2; i <= nxx - 1; ++i)
      muscl_line(&phi(i, 1, k), &tilde(i, 1, k, 2, 1), &tilde(i, 1, k, 2, 2), sy, nyy);
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= nyy - 1; ++j)
    for (int i = 2; i <= nxx - 1; ++i)
      muscl_line(&phi(i, j, 1), &tilde(i, j, 1, 3, 1), &tilde(i, j, 1, 3, 2), sz, nzz);
}

// reconstruct!.jl:102-286.  `phi` is shared across variables and only its
// layers k0-1..k1+1 are written, exactly as in the reference.
void reconstruct(State& st) {
  Arr3 &phi = st.phi, &pbar = st.pbar, &rb = st.rhobar, &jac = st.jac;
  Predictands& p = st.pr;
  const int ka = st.k0 - 1, kb = st.k1 + 1;
  if (st.cfg.model != PCF_BOUSSINESQ) {  // reconstruct!.jl:126-152
    for (int k = ka; k <= kb; ++k)
      for (int j = 1; j <= st.nyy; ++j)
        for (int i = 1; i <= st.nxx; ++i) phi(i, j, k) = p.rho(i, j, k) / pbar(i, j, k);
    muscl_3d(st, phi, st.rhotilde);
  }
  for (int k = ka; k <= kb; ++k)  // reconstruct!.jl:154-176
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i) phi(i, j, k) = p.rhop(i, j, k) / pbar(i, j, k);
  muscl_3d(st, phi, st.rhoptilde);
  for (int k = ka; k <= kb; ++k)  // reconstruct!.jl:178-208
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx - 1; ++i) {
        double rhoedge = 0.5 * (p.rho(i, j, k) + p.rho(i + 1, j, k) + rb(i, j, k) + rb(i + 1, j, k));
        double pedge = 0.5 * (pbar(i, j, k) + pbar(i + 1, j, k));
        phi(i, j, k) = p.u(i, j, k) * rhoedge / pedge;
      }
  muscl_3d(st, phi, st.utilde);
  for (int k = ka; k <= kb; ++k)  // reconstruct!.jl:210-240
    for (int j = 1; j <= st.nyy - 1; ++j)
      for (int i = 1; i <= st.nxx; ++i) {
        double rhoedge = 0.5 * (p.rho(i, j, k) + p.rho(i, j + 1, k) + rb(i, j, k) + rb(i, j + 1, k));
        double pedge = 0.5 * (pbar(i, j, k) + pbar(i, j + 1, k));
        phi(i, j, k) = p.v(i, j, k) * rhoedge / pedge;
      }
  muscl_3d(st, phi, st.vtilde);
  // reconstruct!.jl:242-286
  for (int k = ka; k <= kb; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i) phi(i, j, k) = p.w(i, j, k);
  for (int k = ka; k <= kb; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) phi(i, j, k) = vertical_wind(st, i, j, k);
  zonal_bc(st, phi);
  meridional_bc(st, phi);
  for (int k = ka; k <= kb; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i) {
        double rhoedgeu = (jac(i, j, k + 1) * (p.rho(i, j, k) + rb(i, j, k)) +
                           jac(i, j, k) * (p.rho(i, j, k + 1) + rb(i, j, k + 1))) /
                          (jac(i, j, k) + jac(i, j, k + 1));
        double pedgeu = (jac(i, j, k + 1) * pbar(i, j, k) + jac(i, j, k) * pbar(i, j, k + 1)) /
                        (jac(i, j, k) + jac(i, j, k + 1));
        phi(i, j, k) *= rhoedgeu / pedgeu;
      }
  muscl_3d(st, phi, st.wtilde);
  if (st.tracer) {  // reconstruct!.jl:292-315
    for (int k = ka; k <= kb; ++k)
      for (int j = 1; j <= st.nyy; ++j)
        for (int i = 1; i <= st.nxx; ++i) phi(i, j, k) = st.chi(i, j, k) / pbar(i, j, k);
    muscl_3d(st, phi, st.chitilde);
  }
}

// compute_flux.jl:22-32
static inline double upwind(double usurf, double phiup, double phidown) {
  return usurf > 0.0 ? usurf * phiup : usurf * phidown;
}

// compute_fluxes!.jl:326-1503 (diffusion off)
void compute_fluxes(State& st, Predictands& p0) {
  const int i0 = st.i0, i1 = st.i1, j0 = st.j0, j1 = st.j1, k0 = st.k0, k1 = st.k1;
  Arr3 &jac = st.jac, &pbar = st.pbar, &rb = st.rhobar, &m13 = st.met13, &m23 = st.met23;
  Arr3 &u0 = p0.u, &v0 = p0.v, &w0 = p0.w;
  const bool pi = st.cfg.model != PCF_BOUSSINESQ;
  const double re = st.c.re;
  const bool viscous = !(1 / re <= DBL_EPSILON && true);

  // Rho (compute_fluxes!.jl:359-448), RhoP (:450-520) and the tracer
  // (:1513-1570, the carrier of RhoP); the tracer's come last in the
  // reference, which changes nothing as no pass reads another's output
  for (int pass = 0; pass < 3; ++pass) {
    if (pass == 0 && !pi) continue;
    if (pass == 2 && !st.tracer) continue;
    Arr5& t = pass == 0 ? st.rhotilde : pass == 1 ? st.rhoptilde : st.chitilde;
    Arr4& f = pass == 0 ? st.phirho : pass == 1 ? st.phirhop : st.phichi;
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0 - 1; i <= i1; ++i) {
          double r, l;
          if (pass == 0) {
            double rhobaredger = 0.5 * (rb(i, j, k) + rb(i + 1, j, k));
            double pe = 0.5 * (pbar(i, j, k) + pbar(i + 1, j, k));
            r = t(i + 1, j, k, 1, 1) + rhobaredger / pe;
            l = t(i, j, k, 1, 2) + rhobaredger / pe;
          } else {
            r = t(i + 1, j, k, 1, 1);
            l = t(i, j, k, 1, 2);
          }
          double pedger = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i + 1, j, k) * pbar(i + 1, j, k));
          f(i, j, k, 1) = upwind(pedger * u0(i, j, k), l, r);
        }
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; ++k)
      for (int j = j0 - 1; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          double fr, bk;
          if (pass == 0) {
            double rhobaredgef = 0.5 * (rb(i, j, k) + rb(i, j + 1, k));
            double pe = 0.5 * (pbar(i, j, k) + pbar(i, j + 1, k));
            fr = t(i, j + 1, k, 2, 1) + rhobaredgef / pe;
            bk = t(i, j, k, 2, 2) + rhobaredgef / pe;
          } else {
            fr = t(i, j + 1, k, 2, 1);
            bk = t(i, j, k, 2, 2);
          }
          double pedgef = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j + 1, k) * pbar(i, j + 1, k));
          f(i, j, k, 2) = upwind(pedgef * v0(i, j, k), bk, fr);
        }
#pragma omp parallel for schedule(static)
    for (int k = k0 - 1; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          double up, dn;
          const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
          if (pass == 0) {
            double rhobaredgeu = (ju * rb(i, j, k) + jc * rb(i, j, k + 1)) / (jc + ju);
            double pe = (ju * pbar(i, j, k) + jc * pbar(i, j, k + 1)) / (jc + ju);
            up = t(i, j, k + 1, 3, 1) + rhobaredgeu / pe;
            dn = t(i, j, k, 3, 2) + rhobaredgeu / pe;
          } else {
            up = t(i, j, k + 1, 3, 1);
            dn = t(i, j, k, 3, 2);
          }
          double pedgeu = jc * ju * (pbar(i, j, k) + pbar(i, j, k + 1)) / (jc + ju);
          f(i, j, k, 3) = upwind(pedgeu * w0(i, j, k), dn, up);
        }
  }

  // P (compute_fluxes!.jl:531-585, compressible only): the carrier flux J P u
  if (st.compressible) {
    Arr4& f = st.phip;
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0 - 1; i <= i1; ++i)
          f(i, j, k, 1) = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i + 1, j, k) * pbar(i + 1, j, k)) * u0(i, j, k);
#pragma omp parallel for schedule(static)
    for (int k = k0; k <= k1; ++k)
      for (int j = j0 - 1; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i)
          f(i, j, k, 2) = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j + 1, k) * pbar(i, j + 1, k)) * v0(i, j, k);
#pragma omp parallel for schedule(static)
    for (int k = k0 - 1; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i)
          f(i, j, k, 3) = jac(i, j, k) * jac(i, j, k + 1) * (pbar(i, j, k) + pbar(i, j, k + 1)) /
                          (jac(i, j, k) + jac(i, j, k + 1)) * w0(i, j, k);
  }

  const int kmin = k0, kmax = st.ko + st.nz == st.cfg.z_size ? k1 : k1 + 1;
  auto pe_x = [&](int i, int j, int k) { return 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i + 1, j, k) * pbar(i + 1, j, k)); };
  auto pe_y = [&](int i, int j, int k) { return 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j + 1, k) * pbar(i, j + 1, k)); };
  auto pe_z = [&](int i, int j, int k) {
    return jac(i, j, k) * jac(i, j, k + 1) * (pbar(i, j, k) + pbar(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
  };
  auto ST = [&](int i, int j, int k, int mu, int nu) { return stress_tensor(st, i, j, k, mu, nu); };

  // U (compute_fluxes!.jl:587-889)
  {
    Arr5& t = st.utilde;
    Arr4& f = st.phiu;
#pragma omp parallel for schedule(static)
    for (int k = kmin; k <= kmax; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0 - 2; i <= i1; ++i) {
          double usurf = 0.5 * (pe_x(i, j, k) * u0(i, j, k) + pe_x(i + 1, j, k) * u0(i + 1, j, k));
          f(i, j, k, 1) = upwind(usurf, t(i, j, k, 1, 2), t(i + 1, j, k, 1, 1));
        }
#pragma omp parallel for schedule(static)
    for (int k = kmin; k <= kmax; ++k)
      for (int j = j0 - 1; j <= j1; ++j)
        for (int i = i0 - 1; i <= i1; ++i) {
          double vsurf = 0.5 * (pe_y(i, j, k) * v0(i, j, k) + pe_y(i + 1, j, k) * v0(i + 1, j, k));
          f(i, j, k, 2) = upwind(vsurf, t(i, j, k, 2, 2), t(i, j + 1, k, 2, 1));
        }
#pragma omp parallel for schedule(static)
    for (int k = kmin - 1; k <= kmax; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0 - 1; i <= i1; ++i) {
          double wsurf = 0.5 * (pe_z(i, j, k) * w0(i, j, k) + pe_z(i + 1, j, k) * w0(i + 1, j, k));
          f(i, j, k, 3) = upwind(wsurf, t(i, j, k, 3, 2), t(i, j, k + 1, 3, 1));
        }
    if (viscous) {
#pragma omp parallel for schedule(static)
      for (int k = kmin; k <= kmax; ++k)
        for (int j = j0; j <= j1; ++j)
          for (int i = i0 - 2; i <= i1; ++i) {
            double coef_v = 1 / re * rb(i + 1, j, k0);
            f(i, j, k, 1) -= coef_v * jac(i + 1, j, k) * ST(i + 1, j, k, 1, 1);
          }
#pragma omp parallel for schedule(static)
      for (int k = kmin; k <= kmax; ++k)
        for (int j = j0 - 1; j <= j1; ++j)
          for (int i = i0 - 1; i <= i1; ++i) {
            double coef_v = 1 / re * 0.25 * (rb(i, j, k0) + rb(i + 1, j, k0) + rb(i, j + 1, k0) + rb(i + 1, j + 1, k0));
            f(i, j, k, 2) -= coef_v * 0.25 *
                             (jac(i, j, k) * ST(i, j, k, 1, 2) + jac(i + 1, j, k) * ST(i + 1, j, k, 1, 2) +
                              jac(i, j + 1, k) * ST(i, j + 1, k, 1, 2) + jac(i + 1, j + 1, k) * ST(i + 1, j + 1, k, 1, 2));
          }
#pragma omp parallel for schedule(static)
      for (int k = kmin - 1; k <= kmax; ++k)
        for (int j = j0; j <= j1; ++j)
          for (int i = i0 - 1; i <= i1; ++i) {
            double coef_v = 1 / re * 0.5 * (rb(i, j, k0) + rb(i + 1, j, k0));
            auto s13 = [&](int a, int b, int c) {
              return m13(a, b, c) * ST(a, b, c, 1, 1) + m23(a, b, c) * ST(a, b, c, 1, 2) + ST(a, b, c, 1, 3) / jac(a, b, c);
            };
            double t0 = s13(i, j, k), tr = s13(i + 1, j, k), tu = s13(i, j, k + 1), tru = s13(i + 1, j, k + 1);
            f(i, j, k, 3) -= coef_v * 0.5 *
                             (jac(i, j, k) * jac(i, j, k + 1) * (t0 + tu) / (jac(i, j, k) + jac(i, j, k + 1)) +
                              jac(i + 1, j, k) * jac(i + 1, j, k + 1) * (tr + tru) / (jac(i + 1, j, k) + jac(i + 1, j, k + 1)));
          }
    }
  }
  // V (compute_fluxes!.jl:891-1193)
  {
    Arr5& t = st.vtilde;
    Arr4& f = st.phiv;
#pragma omp parallel for schedule(static)
    for (int k = kmin; k <= kmax; ++k)
      for (int j = j0 - 1; j <= j1; ++j)
        for (int i = i0 - 1; i <= i1; ++i) {
          double usurf = 0.5 * (pe_x(i, j, k) * u0(i, j, k) + pe_x(i, j + 1, k) * u0(i, j + 1, k));
          f(i, j, k, 1) = upwind(usurf, t(i, j, k, 1, 2), t(i + 1, j, k, 1, 1));
        }
#pragma omp parallel for schedule(static)
    for (int k = kmin; k <= kmax; ++k)
      for (int j = j0 - 2; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          double vsurf = 0.5 * (pe_y(i, j, k) * v0(i, j, k) + pe_y(i, j + 1, k) * v0(i, j + 1, k));
          f(i, j, k, 2) = upwind(vsurf, t(i, j, k, 2, 2), t(i, j + 1, k, 2, 1));
        }
#pragma omp parallel for schedule(static)
    for (int k = kmin - 1; k <= kmax; ++k)
      for (int j = j0 - 1; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          double wsurf = 0.5 * (pe_z(i, j, k) * w0(i, j, k) + pe_z(i, j + 1, k) * w0(i, j + 1, k));
          f(i, j, k, 3) = upwind(wsurf, t(i, j, k, 3, 2), t(i, j, k + 1, 3, 1));
        }
    if (viscous) {
#pragma omp parallel for schedule(static)
      for (int k = kmin; k <= kmax; ++k)
        for (int j = j0 - 1; j <= j1; ++j)
          for (int i = i0 - 1; i <= i1; ++i) {
            double coef_v = 1 / re * 0.25 * (rb(i, j, k0) + rb(i + 1, j, k0) + rb(i, j + 1, k0) + rb(i + 1, j + 1, k0));
            f(i, j, k, 1) -= coef_v * 0.25 *
                             (jac(i, j, k) * ST(i, j, k, 2, 1) + jac(i + 1, j, k) * ST(i + 1, j, k, 2, 1) +
                              jac(i, j + 1, k) * ST(i, j + 1, k, 2, 1) + jac(i + 1, j + 1, k) * ST(i + 1, j + 1, k, 2, 1));
          }
#pragma omp parallel for schedule(static)
      for (int k = kmin; k <= kmax; ++k)
        for (int j = j0 - 2; j <= j1; ++j)
          for (int i = i0; i <= i1; ++i) {
            double coef_v = 1 / re * rb(i, j + 1, k0);
            f(i, j, k, 2) -= coef_v * jac(i, j + 1, k) * ST(i, j + 1, k, 2, 2);
          }
#pragma omp parallel for schedule(static)
      for (int k = kmin - 1; k <= kmax; ++k)
        for (int j = j0 - 1; j <= j1; ++j)
          for (int i = i0; i <= i1; ++i) {
            double coef_v = 1 / re * 0.5 * (rb(i, j, k0) + rb(i, j + 1, k0));
            auto s23 = [&](int a, int b, int c) {
              return m13(a, b, c) * ST(a, b, c, 2, 1) + m23(a, b, c) * ST(a, b, c, 2, 2) + ST(a, b, c, 2, 3) / jac(a, b, c);
            };
            double t0 = s23(i, j, k), tf = s23(i, j + 1, k), tu = s23(i, j, k + 1), tfu = s23(i, j + 1, k + 1);
            f(i, j, k, 3) -= coef_v * 0.5 *
                             (jac(i, j, k) * jac(i, j, k + 1) * (t0 + tu) / (jac(i, j, k) + jac(i, j, k + 1)) +
                              jac(i, j + 1, k) * jac(i, j + 1, k + 1) * (tf + tfu) / (jac(i, j + 1, k) + jac(i, j + 1, k + 1)));
          }
    }
  }
  // W (compute_fluxes!.jl:1195-1503)
  {
    Arr5& t = st.wtilde;
    Arr4& f = st.phiw;
#pragma omp parallel for schedule(static)
    for (int k = k0 - 1; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0 - 1; i <= i1; ++i) {
          double pedger = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i + 1, j, k) * pbar(i + 1, j, k));
          double puedger = 0.5 * (jac(i, j, k + 1) * pbar(i, j, k + 1) + jac(i + 1, j, k + 1) * pbar(i + 1, j, k + 1));
          double usurf = ((jac(i, j, k + 1) + jac(i + 1, j, k + 1)) * pedger * u0(i, j, k) +
                          (jac(i, j, k) + jac(i + 1, j, k)) * puedger * u0(i, j, k + 1)) /
                         (jac(i, j, k) + jac(i + 1, j, k) + jac(i, j, k + 1) + jac(i + 1, j, k + 1));
          f(i, j, k, 1) = upwind(usurf, t(i, j, k, 1, 2), t(i + 1, j, k, 1, 1));
        }
#pragma omp parallel for schedule(static)
    for (int k = k0 - 1; k <= k1; ++k)
      for (int j = j0 - 1; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          double pedgef = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j + 1, k) * pbar(i, j + 1, k));
          double puedgef = 0.5 * (jac(i, j, k + 1) * pbar(i, j, k + 1) + jac(i, j + 1, k + 1) * pbar(i, j + 1, k + 1));
          double vsurf = ((jac(i, j, k + 1) + jac(i, j + 1, k + 1)) * pedgef * v0(i, j, k) +
                          (jac(i, j, k) + jac(i, j + 1, k)) * puedgef * v0(i, j, k + 1)) /
                         (jac(i, j, k) + jac(i, j + 1, k) + jac(i, j, k + 1) + jac(i, j + 1, k + 1));
          f(i, j, k, 2) = upwind(vsurf, t(i, j, k, 2, 2), t(i, j + 1, k, 2, 1));
        }
#pragma omp parallel for schedule(static)
    for (int k = k0 - 2; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          double wsurf = 0.5 * (pe_z(i, j, k) * w0(i, j, k) + pe_z(i, j, k + 1) * w0(i, j, k + 1));
          f(i, j, k, 3) = upwind(wsurf, t(i, j, k, 3, 2), t(i, j, k + 1, 3, 1));
        }
    if (viscous) {
#pragma omp parallel for schedule(static)
      for (int k = k0 - 1; k <= k1; ++k)
        for (int j = j0; j <= j1; ++j)
          for (int i = i0 - 1; i <= i1; ++i) {
            double coef_v = 1 / re * 0.5 * (rb(i, j, k0) + rb(i + 1, j, k0));
            f(i, j, k, 1) -= coef_v * 0.5 *
                             (jac(i, j, k) * jac(i, j, k + 1) * (ST(i, j, k, 3, 1) + ST(i, j, k + 1, 3, 1)) /
                                  (jac(i, j, k) + jac(i, j, k + 1)) +
                              jac(i + 1, j, k) * jac(i + 1, j, k + 1) * (ST(i + 1, j, k, 3, 1) + ST(i + 1, j, k + 1, 3, 1)) /
                                  (jac(i + 1, j, k) + jac(i + 1, j, k + 1)));
          }
#pragma omp parallel for schedule(static)
      for (int k = k0 - 1; k <= k1; ++k)
        for (int j = j0 - 1; j <= j1; ++j)
          for (int i = i0; i <= i1; ++i) {
            double coef_v = 1 / re * 0.5 * (rb(i, j, k0) + rb(i, j + 1, k0));
            // compute_fluxes!.jl:1361-1373 uses the (3,1) component on both faces
            f(i, j, k, 2) -= coef_v * 0.5 *
                             (jac(i, j, k) * jac(i, j, k + 1) * (ST(i, j, k, 3, 1) + ST(i, j, k + 1, 3, 1)) /
                                  (jac(i, j, k) + jac(i, j, k + 1)) +
                              jac(i, j + 1, k) * jac(i, j + 1, k + 1) * (ST(i, j + 1, k, 3, 1) + ST(i, j + 1, k + 1, 3, 1)) /
                                  (jac(i, j + 1, k) + jac(i, j + 1, k + 1)));
          }
#pragma omp parallel for schedule(static)
      for (int k = k0 - 2; k <= k1; ++k)
        for (int j = j0; j <= j1; ++j)
          for (int i = i0; i <= i1; ++i) {
            double coef_v = 1 / re * rb(i, j, k0);
            f(i, j, k, 3) -= coef_v * (jac(i, j, k + 1) * m13(i, j, k + 1) * ST(i, j, k + 1, 3, 1) +
                                       jac(i, j, k + 1) * m23(i, j, k + 1) * ST(i, j, k + 1, 3, 2) +
                                       ST(i, j, k + 1, 3, 3));
          }
    }
  }
}

// compute_fluxes!(state, p0, Theta()) (compute_fluxes!.jl:1592-1756):
// heat conduction of the compressible model, theta = pbar / (rho0 + rhobar)
// with rho0 of the step's p0; nothing when thermal_conductivity = 0.
void compute_theta_fluxes(State& st, Predictands& p0) {
  if (!st.compressible || st.cfg.thermal_conductivity == 0.0) return;
  const int i0 = st.i0, i1 = st.i1, j0 = st.j0, j1 = st.j1, k0 = st.k0, k1 = st.k1;
  const double dx = st.dx, dy = st.dy, dz = st.dz;
  Arr3 &jac = st.jac, &pbar = st.pbar, &rb = st.rhobar, &rho = p0.rho;
  Arr3 &m13 = st.met13, &m23 = st.met23, &m33 = st.met33;
  Arr4& f = st.phitheta;
  const double mu_conduct = st.cfg.thermal_conductivity / st.c.uref / st.c.lref;
  auto th = [&](int i, int j, int k) { return pbar(i, j, k) / (rho(i, j, k) + rb(i, j, k)); };
#pragma omp parallel for schedule(static)
  for (int k = k0; k <= k1; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0 - 1; i <= i1; ++i) {
        double coef_t = mu_conduct * 0.5 * (rb(i, j, k0) / rb(i, j, k) + rb(i + 1, j, k0) / rb(i + 1, j, k));
        double thetal = th(i, j, k);
        double thetar = th(i + 1, j, k);
        double thetad = 0.5 * (th(i, j, k - 1) + th(i + 1, j, k - 1));
        double thetau = 0.5 * (th(i, j, k + 1) + th(i + 1, j, k + 1));
        double dtht_dxi = 0.5 * (jac(i, j, k) + jac(i + 1, j, k)) * (thetar - thetal) / dx +
                          0.5 * (jac(i, j, k) * m13(i, j, k) + jac(i + 1, j, k) * m13(i + 1, j, k)) *
                              (thetau - thetad) / (2.0 * dz);
        f(i, j, k, 1) = -coef_t * dtht_dxi;
      }
#pragma omp parallel for schedule(static)
  for (int k = k0; k <= k1; ++k)
    for (int j = j0 - 1; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) {
        double coef_t = mu_conduct * 0.5 * (rb(i, j, k0) / rb(i, j, k) + rb(i, j + 1, k0) / rb(i, j + 1, k));
        double thetab = th(i, j, k);
        double thetaf = th(i, j + 1, k);
        double thetad = 0.5 * (th(i, j, k - 1) + th(i, j + 1, k - 1));
        double thetau = 0.5 * (th(i, j, k + 1) + th(i, j + 1, k + 1));
        double dtht_dyi = 0.5 * (jac(i, j, k) + jac(i, j + 1, k)) * (thetaf - thetab) / dy +
                          0.5 * (jac(i, j, k) * m23(i, j, k) + jac(i, j + 1, k) * m23(i, j + 1, k)) *
                              (thetau - thetad) / (2.0 * dz);
        f(i, j, k, 2) = -coef_t * dtht_dyi;
      }
  // the vertical coefficient takes rhobar of the lowest halo layer (index 1),
  // the horizontal ones that of k0, as the reference does
#pragma omp parallel for schedule(static)
  for (int k = k0 - 1; k <= k1; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) {
        const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
        double coef_t = mu_conduct * (ju * rb(i, j, 1) / rb(i, j, k) + jc * rb(i, j, 1) / rb(i, j, k + 1)) / (ju + jc);
        // (J pbar) / (rho + rhobar) at k and k + 1, Jacobian-weighted to k + 1/2
        auto edge = [&](int ii, int jj) {
          return (jac(ii, jj, k + 1) * pbar(ii, jj, k) / (rho(ii, jj, k) + rb(ii, jj, k)) +
                  jac(ii, jj, k) * pbar(ii, jj, k + 1) / (rho(ii, jj, k + 1) + rb(ii, jj, k + 1))) /
                 (jac(ii, jj, k + 1) + jac(ii, jj, k));
        };
        double thetal = edge(i - 1, j);
        double thetar = edge(i + 1, j);
        double thetab = edge(i, j - 1);
        double thetaf = edge(i, j + 1);
        double thetad = th(i, j, k);
        double thetau = th(i, j, k + 1);
        double dtht_dzi = jc * ju *
                          ((m13(i, j, k) + m13(i, j, k + 1)) * (thetar - thetal) / (2.0 * dx) +
                           (m23(i, j, k) + m23(i, j, k + 1)) * (thetaf - thetab) / (2.0 * dy) +
                           (m33(i, j, k) + m33(i, j, k + 1)) * (thetau - thetad) / dz) /
                          (jc + ju);
        f(i, j, k, 3) = -coef_t * dtht_dzi;
      }
}

}  // namespace orc

// Test hook: apply_1d_muscl! on one contiguous line of n values; face_left[q]
// and face_right[q] are the limited states at the left and right faces of
// cell q (the two components of the reference's phitilde view).
extern "C" void orc_muscl_line(const double* phi, int n, double* face_left, double* face_right) {
  orc::muscl_line(phi, face_left, face_right, 1, n);
}
