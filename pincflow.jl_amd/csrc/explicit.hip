// Explicit part of the step on the device: src/FluxCalculator (reconstruct!,
// apply_3d_muscl!, compute_fluxes!) and src/Update (update!, apply_lhs_sponge!,
// compute_sponges!) for the Boussinesq, pseudo-incompressible and
// compressible models.
//
// The MUSCL reconstruction is evaluated inside the flux kernels from the
// reconstruction inputs phi (reconstruct!.jl:126-286), so the 30 haloed
// reconstruction arrays of the reference are never stored.  Halo values of a
// reconstruction are the periodic images of interior values in the
// reference; the inputs phi keep that property, so evaluating the limiter at a
// halo position gives the same bits.  Only fluxes the updates consume are
// computed; the wall faces that set_boundaries!(state, BoundaryFluxes()) zeroes
// are written as zeros.
#include <cfloat>
#include <cmath>
#include <type_traits>

#include "pcf_device.hpp"

namespace pcf {

namespace {

enum { V_RHO = 0, V_RHOP, V_U, V_V, V_W, V_CHI, V_P };

// Flux inputs in the arithmetic type T of the flux evaluation: double (the
// reference's arithmetic), or float for flux_precision = MIXED, where the
// reconstruction inputs and stress components are stored in FP32 and the
// MUSCL reconstruction, the fluxes and their divergence run in FP32 (the
// other fields are read in FP64 and rounded on load).
template <class T>
struct FluxArgsT {
  Fields F;
  const T* phi[6];
  // the carriers of the p0 winds at the faces, pe_x u0, pe_y v0 and pe_z w0
  // (J P at the face times the wind, k_carriers), computed once per stage
  const T* cx;
  const T* cy;
  const T* cz;
  const double* u0;
  const double* v0;
  const double* w0;
  const T* st[6];  // stress 11, 12, 13, 22, 23, 33
  const T* t13;    // met13 s11 + met23 s12 + s13 / J (k_stress)
  const T* t23;    // met13 s12 + met23 s22 + s23 / J
  // FP32 fluxes (T = float): the background part of the Rho flux at each
  // face, carrier x (rhobar / P at the face), in FP64 (k_rho_background);
  // the FP32 flux is then the fluctuation's alone
  const double* bg[3];
  double* flux[7][3];
  Geo g;
  T inv_re;        // 1 / re
  int viscous;
};
typedef FluxArgsT<double> FluxArgs;

// ---------------------------------------------------------------------------
// reconstruction inputs (reconstruct!.jl:126-286): phi of Rho, RhoP, U, V
// (reconstruct!.jl:154-240) and the tracer (reconstruct!.jl:292-315), the
// Cartesian vertical wind compute_vertical_wind (compute_vertical_wind.jl:
// 27-63) and phi of W, Cartesian w times the edge density over the edge P
// (reconstruct!.jl:242-286)
// ---------------------------------------------------------------------------
// k_phi, k_wcart and k_phi_w in one launch over i, j >= 1, k in 1..nzz-1:
// the reconstruction inputs on layers k0-1..k1+1, the Cartesian vertical
// wind where k_wcart computes it (i, j >= 2) and phi of W from it (the same
// expressions; the fields they share are read once)
template <class S>
__global__ void k_inputs(Fields F, Geo g, S* phr, S* phrp, S* phu, S* phv, int pi, const double* chi, S* phchi,
                         double* wc, S* phw) {
  // 256 threads over the (i, j) of one level, x fastest
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nxx * g.nyy) return;
  const int j = 1 + t / g.nxx;
  const int i = 1 + t - (j - 1) * g.nxx;
  const int k = 1 + blockIdx.y;
  const size_t c = g.h(i, j, k);
  const bool layer = k >= g.k0 - 1 && k <= g.k1 + 1;
  if (layer) {
    const double pb = F.pbar[c];
    if (pi) phr[c] = (S)(F.rho[c] / pb);
    phrp[c] = (S)(F.rhop[c] / pb);
    if (phchi) phchi[c] = (S)(chi[c] / pb);
    if (i < g.nxx) {
      const size_t r = g.h(i + 1, j, k);
      double rhoedge = 0.5 * (F.rho[c] + F.rho[r] + F.rhobar[c] + F.rhobar[r]);
      double pedge = 0.5 * (pb + F.pbar[r]);
      phu[c] = (S)(F.u[c] * rhoedge / pedge);
    }
    if (j < g.nyy) {
      const size_t f = g.h(i, j + 1, k);
      double rhoedge = 0.5 * (F.rho[c] + F.rho[f] + F.rhobar[c] + F.rhobar[f]);
      double pedge = 0.5 * (pb + F.pbar[f]);
      phv[c] = (S)(F.v[c] * rhoedge / pedge);
    }
  }
  if (i < 2 || j < 2) return;
  const double w = vertical_wind(F, g, i, j, k);
  wc[c] = w;
  if (layer) {
    const size_t up = g.h(i, j, k + 1);
    const double jc = F.jac[c], ju = F.jac[up];
    double rhoedgeu = (ju * (F.rho[c] + F.rhobar[c]) + jc * (F.rho[up] + F.rhobar[up])) / (jc + ju);
    double pedgeu = (ju * F.pbar[c] + jc * F.pbar[up]) / (jc + ju);
    phw[c] = (S)(w * (rhoedgeu / pedgeu));
  }
}

// compute_stress_tensor.jl:55-164: the six distinct components at one cell.
// Besides the six components: the Cartesian-to-transformed combinations the
// U and V vertical viscous fluxes take at each cell, t13 = met13 s11 +
// met23 s12 + s13 / J and t23 = met13 s12 + met23 s22 + s23 / J
// (compute_fluxes!.jl:743-765, 1047-1069), from the FP64 components.
template <class S>
__global__ void k_stress(Fields F, Geo g, const double* wc, S* s11, S* s12, S* s13, S* s22, S* s23, S* s33,
                         S* t13, S* t23, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const double *u = F.u, *v = F.v, *w = F.w, *jac = F.jac;
  const double dx = g.dx, dy = g.dy, dz = g.dz;
  const size_t c = g.h(i, j, k);
  double jacedger = 0.5 * (jac[c] + jac[g.h(i + 1, j, k)]);
  double jacedgel = 0.5 * (jac[c] + jac[g.h(i - 1, j, k)]);
  double jacedgef = 0.5 * (jac[c] + jac[g.h(i, j + 1, k)]);
  double jacedgeb = 0.5 * (jac[c] + jac[g.h(i, j - 1, k)]);
  double jacedgeu = 2.0 * jac[c] * jac[g.h(i, j, k + 1)] / (jac[c] + jac[g.h(i, j, k + 1)]);
  double jacedged = 2.0 * jac[c] * jac[g.h(i, j, k - 1)] / (jac[c] + jac[g.h(i, j, k - 1)]);
  double uf = 0.5 * (u[g.h(i, j + 1, k)] + u[g.h(i - 1, j + 1, k)]);
  double ub = 0.5 * (u[g.h(i, j - 1, k)] + u[g.h(i - 1, j - 1, k)]);
  double uu = 0.5 * (u[g.h(i, j, k + 1)] + u[g.h(i - 1, j, k + 1)]);
  double ud = 0.5 * (u[g.h(i, j, k - 1)] + u[g.h(i - 1, j, k - 1)]);
  double vr = 0.5 * (v[g.h(i + 1, j, k)] + v[g.h(i + 1, j - 1, k)]);
  double vl = 0.5 * (v[g.h(i - 1, j, k)] + v[g.h(i - 1, j - 1, k)]);
  double vu = 0.5 * (v[g.h(i, j, k + 1)] + v[g.h(i, j - 1, k + 1)]);
  double vd = 0.5 * (v[g.h(i, j, k - 1)] + v[g.h(i, j - 1, k - 1)]);
  double wr = 0.5 * (wc[g.h(i + 1, j, k)] + wc[g.h(i + 1, j, k - 1)]);
  double wl = 0.5 * (wc[g.h(i - 1, j, k)] + wc[g.h(i - 1, j, k - 1)]);
  double wf = 0.5 * (wc[g.h(i, j + 1, k)] + wc[g.h(i, j + 1, k - 1)]);
  double wb = 0.5 * (wc[g.h(i, j - 1, k)] + wc[g.h(i, j - 1, k - 1)]);
  // compute_stress_tensor.jl:112-118: 2/3 * (divergence) / jac, left to right
  double div = 2.0 / 3.0 *
               ((jacedger * u[c] - jacedgel * u[g.h(i - 1, j, k)]) / dx +
                (jacedgef * v[c] - jacedgeb * v[g.h(i, j - 1, k)]) / dy +
                (jacedgeu * w[c] - jacedged * w[g.h(i, j, k - 1)]) / dz) /
               jac[c];
  double dwc = wc[c] - wc[g.h(i, j, k - 1)];
  const double v11 = 2.0 * (u[c] - u[g.h(i - 1, j, k)]) / dx + F.m13[c] * (uu - ud) / dz - div;
  const double v12 = 0.5 * (uf - ub) / dy + 0.5 * F.m23[c] * (uu - ud) / dz + 0.5 * (vr - vl) / dx +
                     0.5 * F.m13[c] * (vu - vd) / dz;
  const double v13 = 0.5 * (uu - ud) / dz / jac[c] + 0.5 * (wr - wl) / dx + F.m13[c] * dwc / dz;
  const double v22 = 2.0 * (v[c] - v[g.h(i, j - 1, k)]) / dy + F.m23[c] * (vu - vd) / dz - div;
  const double v23 = 0.5 * (vu - vd) / dz / jac[c] + 0.5 * (wf - wb) / dy + F.m23[c] * dwc / dz;
  s11[c] = (S)v11;
  s12[c] = (S)v12;
  s13[c] = (S)v13;
  s22[c] = (S)v22;
  s23[c] = (S)v23;
  s33[c] = (S)(2.0 * dwc / dz / jac[c] - div);
  t13[c] = (S)(F.m13[c] * v11 + F.m23[c] * v12 + v13 / jac[c]);
  t23[c] = (S)(F.m13[c] * v12 + F.m23[c] * v22 + v23 / jac[c]);
}

// MUSCL face values from phi along a stride: right face of cell c and left
// face of cell c + 1.
template <class T>
__device__ __forceinline__ T tmin(T a, T b) { return b < a ? b : a; }
template <class T>
__device__ __forceinline__ T tmax(T a, T b) { return a < b ? b : a; }

// apply_1d_muscl!.jl:51-93 at one interior point of a line: given
// phi(q - 1), phi(q), phi(q + 1) return tilde(q, 1) (left face value) and
// tilde(q, 2) (right face value), in the arithmetic type T.
template <class T>
__device__ __forceinline__ void muscl_t(T pm, T p, T pp, T& t1, T& t2) {
  T deltal = p - pm;
  T deltar = pp - p;
  if (deltar == T(0)) {
    t2 = p;
    t1 = p;
  } else if (deltal == T(0)) {
    T theta = deltal / deltar;
    T s = (T(2) + theta) / T(3);
    T sigmal = tmax(T(0), tmin(tmin(T(2) * theta, s), T(2)));
    t2 = p + T(0.5) * sigmal * deltar;
    t1 = p;
  } else {
    T theta = deltal / deltar;
    T s = (T(2) + theta) / T(3);
    T sigmal = tmax(T(0), tmin(tmin(T(2) * theta, s), T(2)));
    s = (T(2) + T(1) / theta) / T(3);
    T sigmar = tmax(T(0), tmin(tmin(T(2) / theta, s), T(2)));
    t2 = p + T(0.5) * sigmal * deltar;
    t1 = p - T(0.5) * sigmar * deltal;
  }
}

template <class T>
__device__ __forceinline__ T upwind_t(T usurf, T phiup, T phidown) {  // compute_flux.jl:22-32
  return usurf > T(0) ? usurf * phiup : usurf * phidown;
}

template <class T>
__device__ __forceinline__ void faces(const T* phi, size_t c, size_t stride, T& right_of_c, T& left_of_next) {
  T pm = phi[c - stride], p0 = phi[c], p1 = phi[c + stride], p2 = phi[c + 2 * stride];
  T t1, t2;
  muscl_t<T>(pm, p0, p1, t1, t2);
  right_of_c = t2;
  muscl_t<T>(p0, p1, p2, t1, t2);
  left_of_next = t1;
}

// a field value in the arithmetic type T
#define LT(arr, idx) ((T)(arr)[idx])

template <class T>
__device__ __forceinline__ T pe_x(const Fields& F, const Geo& g, int i, int j, int k) {
  size_t c = g.h(i, j, k), r = g.h(i + 1, j, k);
  return T(0.5) * (LT(F.jac, c) * LT(F.pbar, c) + LT(F.jac, r) * LT(F.pbar, r));
}
template <class T>
__device__ __forceinline__ T pe_y(const Fields& F, const Geo& g, int i, int j, int k) {
  size_t c = g.h(i, j, k), f = g.h(i, j + 1, k);
  return T(0.5) * (LT(F.jac, c) * LT(F.pbar, c) + LT(F.jac, f) * LT(F.pbar, f));
}
template <class T>
__device__ __forceinline__ T pe_z(const Fields& F, const Geo& g, int i, int j, int k) {
  size_t c = g.h(i, j, k), u = g.h(i, j, k + 1);
  return LT(F.jac, c) * LT(F.jac, u) * (LT(F.pbar, c) + LT(F.pbar, u)) / (LT(F.jac, c) + LT(F.jac, u));
}

// The carrier of each face direction, pe u0 (compute_fluxes!.jl:359-520),
// where the face's far cell exists: x faces for i < nxx, y faces for
// j < nyy, z faces for k < nzz.
template <class T>
__global__ void k_carriers(Fields F, Geo g, const double* u0, const double* v0, const double* w0, T* cx, T* cy,
                           T* cz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nxx) return;
  const size_t c = g.h(i, j, k);
  if (i < g.nxx) cx[c] = pe_x<T>(F, g, i, j, k) * (T)u0[c];
  if (j < g.nyy) cy[c] = pe_y<T>(F, g, i, j, k) * (T)v0[c];
  if (k < g.nzz) cz[c] = pe_z<T>(F, g, i, j, k) * (T)w0[c];
}

// With FP32 fluxes, the part of the Rho flux that carries the background
// density, carrier x rhobar / P at each face (the term the FP64 flux adds
// to both reconstructed values, compute_fluxes!.jl:359-520), in FP64: the
// FP32 flux then carries only the small fluctuation rho / P, and the
// divergence of this large, nearly divergence-free part keeps FP64 rounding.
// Fixed through an LHS integration (the p0 winds and the background are).
__global__ void k_rho_background(Fields F, Geo g, const double* u0, const double* v0, const double* w0, double* bx,
                                 double* by, double* bz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nxx) return;
  const size_t c = g.h(i, j, k);
  if (i < g.nxx) {
    const size_t r = g.h(i + 1, j, k);
    const double rb = 0.5 * (F.rhobar[c] + F.rhobar[r]), pe = 0.5 * (F.pbar[c] + F.pbar[r]);
    bx[c] = pe_x<double>(F, g, i, j, k) * u0[c] * (rb / pe);
  }
  if (j < g.nyy) {
    const size_t f = g.h(i, j + 1, k);
    const double rb = 0.5 * (F.rhobar[c] + F.rhobar[f]), pe = 0.5 * (F.pbar[c] + F.pbar[f]);
    by[c] = pe_y<double>(F, g, i, j, k) * v0[c] * (rb / pe);
  }
  if (k < g.nzz) {
    const size_t u = g.h(i, j, k + 1);
    const double jc = F.jac[c], ju = F.jac[u];
    const double rb = (ju * F.rhobar[c] + jc * F.rhobar[u]) / (jc + ju);
    const double pe = (ju * F.pbar[c] + jc * F.pbar[u]) / (jc + ju);
    bz[c] = pe_z<double>(F, g, i, j, k) * w0[c] * (rb / pe);
  }
}

// One flux component of one variable at face (i, j, k)
// (compute_fluxes!.jl:359-520, 587-1503): a pure function of the
// reconstruction inputs, the p0 winds, the stress components and the static
// fields, so every kernel that evaluates it at a face gets the same bits.
template <int VAR, int DIR, class T = double>
__device__ __forceinline__ T flux_value(const FluxArgsT<T>& A, int i, int j, int k) {
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k);
  const size_t sx = 1, sy = (size_t)g.nxx, sz = (size_t)g.nxx * g.nyy;
  const size_t stride = DIR == 0 ? sx : DIR == 1 ? sy : sz;
  T right_of_c, left_of_next;
  faces(A.phi[VAR], c, stride, right_of_c, left_of_next);
  T val;
  if (VAR == V_RHO || VAR == V_RHOP || VAR == V_CHI) {  // the tracer: RhoP's carrier (compute_fluxes!.jl:1513-1570)
    T hi = left_of_next, lo = right_of_c;  // value at c + e, value at c
    // the background term rhobar / P enters FP64 fluxes here and FP32
    // fluxes as the separate FP64 part A.bg
    constexpr bool bg_here = VAR == V_RHO && sizeof(T) == 8;
    if (DIR == 0) {
      if (bg_here) {
        const size_t r = g.h(i + 1, j, k);
        T rhobaredger = T(0.5) * (LT(F.rhobar, c) + LT(F.rhobar, r));
        T pe = T(0.5) * (LT(F.pbar, c) + LT(F.pbar, r));
        hi = hi + rhobaredger / pe;
        lo = lo + rhobaredger / pe;
      }
      val = upwind_t<T>(A.cx[c], lo, hi);
    } else if (DIR == 1) {
      if (bg_here) {
        const size_t f = g.h(i, j + 1, k);
        T rhobaredgef = T(0.5) * (LT(F.rhobar, c) + LT(F.rhobar, f));
        T pe = T(0.5) * (LT(F.pbar, c) + LT(F.pbar, f));
        hi = hi + rhobaredgef / pe;
        lo = lo + rhobaredgef / pe;
      }
      val = upwind_t<T>(A.cy[c], lo, hi);
    } else {
      const size_t u = g.h(i, j, k + 1);
      const T jc = LT(F.jac, c), ju = LT(F.jac, u);
      if (bg_here) {
        T rhobaredgeu = (ju * LT(F.rhobar, c) + jc * LT(F.rhobar, u)) / (jc + ju);
        T pe = (ju * LT(F.pbar, c) + jc * LT(F.pbar, u)) / (jc + ju);
        hi = hi + rhobaredgeu / pe;
        lo = lo + rhobaredgeu / pe;
      }
      val = upwind_t<T>(A.cz[c], lo, hi);
    }
  } else if (VAR == V_U) {
    const T* S11 = A.st[0];
    const T* S12 = A.st[1];
    const T* S13 = A.st[2];
    if (DIR == 0) {
      T usurf = T(0.5) * (A.cx[c] + A.cx[g.h(i + 1, j, k)]);
      val = upwind_t<T>(usurf, right_of_c, left_of_next);
      if (A.viscous) {
        T coef_v = A.inv_re * LT(F.rhobar, g.h(i + 1, j, g.k0));
        val -= coef_v * LT(F.jac, g.h(i + 1, j, k)) * S11[g.h(i + 1, j, k)];
      }
    } else if (DIR == 1) {
      T vsurf = T(0.5) * (A.cy[c] + A.cy[g.h(i + 1, j, k)]);
      val = upwind_t<T>(vsurf, right_of_c, left_of_next);
      if (A.viscous) {
        const size_t r = g.h(i + 1, j, k), f = g.h(i, j + 1, k), rf = g.h(i + 1, j + 1, k);
        T coef_v = A.inv_re * T(0.25) *
                        (LT(F.rhobar, g.h(i, j, g.k0)) + LT(F.rhobar, g.h(i + 1, j, g.k0)) + LT(F.rhobar, g.h(i, j + 1, g.k0)) +
                         LT(F.rhobar, g.h(i + 1, j + 1, g.k0)));
        val -= coef_v * T(0.25) * (LT(F.jac, c) * S12[c] + LT(F.jac, r) * S12[r] + LT(F.jac, f) * S12[f] + LT(F.jac, rf) * S12[rf]);
      }
    } else {
      T wsurf = T(0.5) * (A.cz[c] + A.cz[g.h(i + 1, j, k)]);
      val = upwind_t<T>(wsurf, right_of_c, left_of_next);
      if (A.viscous) {
        T coef_v = A.inv_re * T(0.5) * (LT(F.rhobar, g.h(i, j, g.k0)) + LT(F.rhobar, g.h(i + 1, j, g.k0)));
        const T* T13 = A.t13;
        const size_t r = g.h(i + 1, j, k), u = g.h(i, j, k + 1), ru = g.h(i + 1, j, k + 1);
        T t0 = T13[c], tr = T13[r], tu = T13[u], tru = T13[ru];
        val -= coef_v * T(0.5) *
               (LT(F.jac, c) * LT(F.jac, u) * (t0 + tu) / (LT(F.jac, c) + LT(F.jac, u)) +
                LT(F.jac, r) * LT(F.jac, ru) * (tr + tru) / (LT(F.jac, r) + LT(F.jac, ru)));
      }
    }
  } else if (VAR == V_V) {
    const T* S12 = A.st[1];
    const T* S22 = A.st[3];
    const T* S23 = A.st[4];
    if (DIR == 0) {
      T usurf = T(0.5) * (A.cx[c] + A.cx[g.h(i, j + 1, k)]);
      val = upwind_t<T>(usurf, right_of_c, left_of_next);
      if (A.viscous) {
        const size_t r = g.h(i + 1, j, k), f = g.h(i, j + 1, k), rf = g.h(i + 1, j + 1, k);
        T coef_v = A.inv_re * T(0.25) *
                        (LT(F.rhobar, g.h(i, j, g.k0)) + LT(F.rhobar, g.h(i + 1, j, g.k0)) + LT(F.rhobar, g.h(i, j + 1, g.k0)) +
                         LT(F.rhobar, g.h(i + 1, j + 1, g.k0)));
        val -= coef_v * T(0.25) * (LT(F.jac, c) * S12[c] + LT(F.jac, r) * S12[r] + LT(F.jac, f) * S12[f] + LT(F.jac, rf) * S12[rf]);
      }
    } else if (DIR == 1) {
      T vsurf = T(0.5) * (A.cy[c] + A.cy[g.h(i, j + 1, k)]);
      val = upwind_t<T>(vsurf, right_of_c, left_of_next);
      if (A.viscous) {
        T coef_v = A.inv_re * LT(F.rhobar, g.h(i, j + 1, g.k0));
        val -= coef_v * LT(F.jac, g.h(i, j + 1, k)) * S22[g.h(i, j + 1, k)];
      }
    } else {
      T wsurf = T(0.5) * (A.cz[c] + A.cz[g.h(i, j + 1, k)]);
      val = upwind_t<T>(wsurf, right_of_c, left_of_next);
      if (A.viscous) {
        T coef_v = A.inv_re * T(0.5) * (LT(F.rhobar, g.h(i, j, g.k0)) + LT(F.rhobar, g.h(i, j + 1, g.k0)));
        const T* T23 = A.t23;
        const size_t f = g.h(i, j + 1, k), u = g.h(i, j, k + 1), fu = g.h(i, j + 1, k + 1);
        T t0 = T23[c], tf = T23[f], tu = T23[u], tfu = T23[fu];
        val -= coef_v * T(0.5) *
               (LT(F.jac, c) * LT(F.jac, u) * (t0 + tu) / (LT(F.jac, c) + LT(F.jac, u)) +
                LT(F.jac, f) * LT(F.jac, fu) * (tf + tfu) / (LT(F.jac, f) + LT(F.jac, fu)));
      }
    }
  } else {  // W
    const T* S13 = A.st[2];
    const T* S23 = A.st[4];
    const T* S33 = A.st[5];
    if (DIR == 0) {
      const size_t r = g.h(i + 1, j, k), u = g.h(i, j, k + 1), ru = g.h(i + 1, j, k + 1);
      T pedger = T(0.5) * (LT(F.jac, c) * LT(F.pbar, c) + LT(F.jac, r) * LT(F.pbar, r));
      T puedger = T(0.5) * (LT(F.jac, u) * LT(F.pbar, u) + LT(F.jac, ru) * LT(F.pbar, ru));
      T usurf = ((LT(F.jac, u) + LT(F.jac, ru)) * pedger * LT(A.u0, c) + (LT(F.jac, c) + LT(F.jac, r)) * puedger * LT(A.u0, u)) /
                     (LT(F.jac, c) + LT(F.jac, r) + LT(F.jac, u) + LT(F.jac, ru));
      val = upwind_t<T>(usurf, right_of_c, left_of_next);
      if (A.viscous) {
        T coef_v = A.inv_re * T(0.5) * (LT(F.rhobar, g.h(i, j, g.k0)) + LT(F.rhobar, g.h(i + 1, j, g.k0)));
        val -= coef_v * T(0.5) *
               (LT(F.jac, c) * LT(F.jac, u) * (S13[c] + S13[u]) / (LT(F.jac, c) + LT(F.jac, u)) +
                LT(F.jac, r) * LT(F.jac, ru) * (S13[r] + S13[ru]) / (LT(F.jac, r) + LT(F.jac, ru)));
      }
    } else if (DIR == 1) {
      const size_t f = g.h(i, j + 1, k), u = g.h(i, j, k + 1), fu = g.h(i, j + 1, k + 1);
      T pedgef = T(0.5) * (LT(F.jac, c) * LT(F.pbar, c) + LT(F.jac, f) * LT(F.pbar, f));
      T puedgef = T(0.5) * (LT(F.jac, u) * LT(F.pbar, u) + LT(F.jac, fu) * LT(F.pbar, fu));
      T vsurf = ((LT(F.jac, u) + LT(F.jac, fu)) * pedgef * LT(A.v0, c) + (LT(F.jac, c) + LT(F.jac, f)) * puedgef * LT(A.v0, u)) /
                     (LT(F.jac, c) + LT(F.jac, f) + LT(F.jac, u) + LT(F.jac, fu));
      val = upwind_t<T>(vsurf, right_of_c, left_of_next);
      if (A.viscous) {
        // compute_fluxes!.jl:1351-1367 takes the (3, 1) component on this face
        T coef_v = A.inv_re * T(0.5) * (LT(F.rhobar, g.h(i, j, g.k0)) + LT(F.rhobar, g.h(i, j + 1, g.k0)));
        val -= coef_v * T(0.5) *
               (LT(F.jac, c) * LT(F.jac, u) * (S13[c] + S13[u]) / (LT(F.jac, c) + LT(F.jac, u)) +
                LT(F.jac, f) * LT(F.jac, fu) * (S13[f] + S13[fu]) / (LT(F.jac, f) + LT(F.jac, fu)));
      }
    } else {
      T wsurf = T(0.5) * (A.cz[c] + A.cz[g.h(i, j, k + 1)]);
      val = upwind_t<T>(wsurf, right_of_c, left_of_next);
      if (A.viscous) {
        const size_t u = g.h(i, j, k + 1);
        T coef_v = A.inv_re * LT(F.rhobar, g.h(i, j, g.k0));
        val -= coef_v * (LT(F.jac, u) * LT(F.m13, u) * S13[u] + LT(F.jac, u) * LT(F.m23, u) * S23[u] + S33[u]);
      }
    }
  }
  return val;
}

template <int VAR, int DIR>
__global__ void k_flux(FluxArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  A.flux[VAR][DIR][A.g.h(i, j, k)] = flux_value<VAR, DIR>(A, i, j, k);
}

// compute_fluxes!(..., P()) (compute_fluxes!.jl:531-585): the carrier flux
// J P u of the compressible model
template <int DIR>
__global__ void k_flux_p(FluxArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  const size_t c = g.h(i, j, k);
  double val;
  if (DIR == 0) val = pe_x<double>(A.F, g, i, j, k) * A.u0[c];
  else if (DIR == 1) val = pe_y<double>(A.F, g, i, j, k) * A.v0[c];
  else val = pe_z<double>(A.F, g, i, j, k) * A.w0[c];
  A.flux[V_P][DIR][c] = val;
}

// compute_fluxes!(state, p0, Theta()) (compute_fluxes!.jl:1592-1756): heat
// conduction with theta = pbar / (rho0 + rhobar); the vertical wall faces,
// which set_boundaries!(state, BoundaryFluxes()) zeroes before any use, are
// written as zeros.
struct ThetaArgs {
  const double *rho0, *rhobar, *pbar, *jac, *m13, *m23, *m33;
  double* f[3];
  Geo g;
  double mu;  // thermal_conductivity / uref / lref
};

template <int DIR>
__global__ void k_flux_theta(ThetaArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  const double *rb = A.rhobar, *pb = A.pbar, *rho = A.rho0, *jac = A.jac;
  auto th = [&](int ii, int jj, int kk) {
    const size_t q = g.h(ii, jj, kk);
    return pb[q] / (rho[q] + rb[q]);
  };
  const size_t c = g.h(i, j, k);
  double val;
  if (DIR == 0) {
    const size_t r = g.h(i + 1, j, k);
    double coef_t = A.mu * 0.5 * (rb[g.h(i, j, g.k0)] / rb[c] + rb[g.h(i + 1, j, g.k0)] / rb[r]);
    double thetal = th(i, j, k);
    double thetar = th(i + 1, j, k);
    double thetad = 0.5 * (th(i, j, k - 1) + th(i + 1, j, k - 1));
    double thetau = 0.5 * (th(i, j, k + 1) + th(i + 1, j, k + 1));
    double d = 0.5 * (jac[c] + jac[r]) * (thetar - thetal) / g.dx +
               0.5 * (jac[c] * A.m13[c] + jac[r] * A.m13[r]) * (thetau - thetad) / (2.0 * g.dz);
    val = -coef_t * d;
  } else if (DIR == 1) {
    const size_t f = g.h(i, j + 1, k);
    double coef_t = A.mu * 0.5 * (rb[g.h(i, j, g.k0)] / rb[c] + rb[g.h(i, j + 1, g.k0)] / rb[f]);
    double thetab = th(i, j, k);
    double thetaf = th(i, j + 1, k);
    double thetad = 0.5 * (th(i, j, k - 1) + th(i, j + 1, k - 1));
    double thetau = 0.5 * (th(i, j, k + 1) + th(i, j + 1, k + 1));
    double d = 0.5 * (jac[c] + jac[f]) * (thetaf - thetab) / g.dy +
               0.5 * (jac[c] * A.m23[c] + jac[f] * A.m23[f]) * (thetau - thetad) / (2.0 * g.dz);
    val = -coef_t * d;
  } else if (k == g.k0 - 1 || k == g.k1) {
    val = 0.0;
  } else {
    const size_t u = g.h(i, j, k + 1);
    const double jc = jac[c], ju = jac[u];
    // rhobar of the lowest halo layer (index 1) as the reference takes it
    const double rb1 = rb[g.h(i, j, 1)];
    double coef_t = A.mu * (ju * rb1 / rb[c] + jc * rb1 / rb[u]) / (ju + jc);
    auto edge = [&](int ii, int jj) {  // (J pbar) / (rho + rhobar), weighted to k + 1/2
      const size_t q = g.h(ii, jj, k), qu = g.h(ii, jj, k + 1);
      return (jac[qu] * pb[q] / (rho[q] + rb[q]) + jac[q] * pb[qu] / (rho[qu] + rb[qu])) / (jac[qu] + jac[q]);
    };
    double thetal = edge(i - 1, j);
    double thetar = edge(i + 1, j);
    double thetab = edge(i, j - 1);
    double thetaf = edge(i, j + 1);
    double thetad = th(i, j, k);
    double thetau = th(i, j, k + 1);
    double d = jc * ju *
               ((A.m13[c] + A.m13[u]) * (thetar - thetal) / (2.0 * g.dx) +
                (A.m23[c] + A.m23[u]) * (thetaf - thetab) / (2.0 * g.dy) +
                (A.m33[c] + A.m33[u]) * (thetau - thetad) / g.dz) /
               (jc + ju);
    val = -coef_t * d;
  }
  A.f[DIR][c] = val;
}

// ---------------------------------------------------------------------------
// RK3 updates of the left-hand side (update!.jl:505-1192)
// ---------------------------------------------------------------------------
struct UpdArgs {
  Fields F;
  Geo g;
  Phys ph;
  double* q;            // updated predictand
  double* dq;           // increment
  const double* f[3];   // flux of q
  const double* fu[3];  // flux of u (W update)
  const double* fv[3];  // flux of v (W update)
  const double* rhoold;
  const double* uold;
  const double* vold;
  double dt, am, bm;
  int first;
  int force0;  // the tracer adds its (zero) volume force, update!.jl:1490-1492
  int heat;    // 1: RhoP, -heating / thetabar (update!.jl:586-594); 2: P, +heating (update!.jl:1436-1442)
  double* div_out;        // U / V updates: their flux divergence / J_edge, kept for the W update (or null)
  const double* div_u;    // W update: those of the U and V updates of this stage (or null: recompute)
  const double* div_v;
  double* fdiv;           // fused stage, W: its flux divergence, read by k_update_w_fdiv
  // fused stage: apply_lhs_sponge! (apply_lhs_sponge!.jl:278-530) on the
  // interior cells right after their update (sponge = 0: none), with the
  // sponge's time step and relaxation targets of k_lhs_sponge
  int sponge;
  const double* alphar;
  double sdt, bgu, bgv, bgw_u, bgw_v, bgw_w, uref;
};

// apply_lhs_sponge! at one cell of the interior: the new value of q there,
// given its updated value (the arithmetic of k_lhs_sponge)
template <int VAR>
__device__ __forceinline__ double sponged(const UpdArgs& A, int i, int j, int k, double q) {
  const Geo& g = A.g;
  if (!A.sponge || i < g.i0 || i > g.i1 || j < g.j0 || j > g.j1 || k < g.k0 || k > g.k1) return q;
  const size_t c = g.h(i, j, k);
  const double* alphar = A.alphar;
  double bg, alpha;
  if (VAR == 0 || VAR == 1) {  // V_RHO, V_RHOP
    bg = 0.0;
    alpha = alphar[c];
  } else if (VAR == 2) {  // V_U
    bg = A.bgu;
    alpha = 0.5 * (alphar[c] + alphar[g.h(i + 1, j, k)]);
  } else if (VAR == 3) {  // V_V
    bg = A.bgv;
    alpha = 0.5 * (alphar[c] + alphar[g.h(i, j + 1, k)]);
  } else {
    const size_t up = g.h(i, j, k + 1);
    const double *jac = A.F.jac, *m13 = A.F.m13, *m23 = A.F.m23;
    const double jc = jac[c], ju = jac[up];
    bg = (jc * (m13[up] * A.bgw_u + m23[up] * A.bgw_v + A.bgw_w / ju) +
          ju * (m13[c] * A.bgw_u + m23[c] * A.bgw_v + A.bgw_w / jc)) /
         (jc + ju) / A.uref;
    alpha = (ju * alphar[c] + jc * alphar[up]) / (jc + ju);
  }
  double beta = 1.0 / (1.0 + alpha * A.sdt);
  return (1.0 - beta) * bg + beta * q;
}

__device__ __forceinline__ double divergence(const double* const* f, const Geo& g, int i, int j, int k) {
  const size_t c = g.h(i, j, k);
  return (f[0][c] - f[0][g.h(i - 1, j, k)]) / g.dx + (f[1][c] - f[1][g.h(i, j - 1, k)]) / g.dy +
         (f[2][c] - f[2][g.h(i, j, k - 1)]) / g.dz;
}

// The updates at one cell from the flux divergence `fluxdiff` (the three
// differences over dx, dy, dz, summed in that order), shared by the update
// kernels that read stored fluxes and by the fused stage kernels.
__device__ __forceinline__ double update_scalar_at(const UpdArgs& A, int i, int j, int k, double fluxdiff) {
  // update!.jl:521-597
  const size_t c = A.g.h(i, j, k);
  fluxdiff /= A.F.jac[c];
  double fr = -fluxdiff;
  if (A.force0) fr = fr + 0.0;
  if (A.heat == 1) fr = fr - conductive_heating(A.F, A.g, i, j, k) / A.F.thetabar[c];
  if (A.heat == 2) fr = fr + conductive_heating(A.F, A.g, i, j, k);
  double dqv = A.first ? 0.0 : A.dq[c];
  dqv = A.dt * fr + A.am * dqv;
  A.dq[c] = dqv;
  return A.q[c] + A.bm * dqv;
}

__global__ void k_update_scalar(UpdArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  A.q[A.g.h(i, j, k)] = update_scalar_at(A, i, j, k, divergence(A.f, A.g, i, j, k));
}

__device__ __forceinline__ double update_u_at(const UpdArgs& A, int i, int j, int k, double fluxdiff) {
  // update!.jl:730-806
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k), r = g.h(i + 1, j, k);
  double jacedger = 0.5 * (F.jac[c] + F.jac[r]);
  fluxdiff /= jacedger;
  if (A.div_out) A.div_out[c] = fluxdiff;
  double vc = 0.5 * (F.v[c] + F.v[g.h(i, j - 1, k)]);
  double vr = 0.5 * (F.v[r] + F.v[g.h(i + 1, j - 1, k)]);
  double volforce = 0.5 * A.ph.fc * ((A.rhoold[c] + F.rhobar[c]) * vc + (A.rhoold[r] + F.rhobar[r]) * vr);
  double force = -fluxdiff + volforce;
  double rhom_1 = 0.5 * (A.rhoold[c] + A.rhoold[r]);
  double rhom = 0.5 * (F.rho[c] + F.rho[r]);
  double rhobaredger = 0.5 * (F.rhobar[c] + F.rhobar[r]);
  rhom_1 += rhobaredger;
  rhom += rhobaredger;
  double um_1 = A.q[c];
  double momm_1 = rhom_1 * um_1;
  double d = A.first ? 0.0 : A.dq[c];
  d = A.dt * force + A.am * d;
  A.dq[c] = d;
  double momm = momm_1 + A.bm * d;
  return momm / rhom;
}

__global__ void k_update_u(UpdArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  A.q[A.g.h(i, j, k)] = update_u_at(A, i, j, k, divergence(A.f, A.g, i, j, k));
}

__device__ __forceinline__ double update_v_at(const UpdArgs& A, int i, int j, int k, double fluxdiff) {
  // update!.jl:884-957
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k), f = g.h(i, j + 1, k);
  double jacedgef = 0.5 * (F.jac[c] + F.jac[f]);
  fluxdiff /= jacedgef;
  if (A.div_out) A.div_out[c] = fluxdiff;
  double uc = 0.5 * (A.uold[c] + A.uold[g.h(i - 1, j, k)]);
  double uf = 0.5 * (A.uold[f] + A.uold[g.h(i - 1, j + 1, k)]);
  double volforce = -0.5 * A.ph.fc * ((A.rhoold[c] + F.rhobar[c]) * uc + (A.rhoold[f] + F.rhobar[f]) * uf);
  double force = -fluxdiff + volforce;
  double rhom_1 = 0.5 * (A.rhoold[c] + A.rhoold[f]);
  double rhom = 0.5 * (F.rho[c] + F.rho[f]);
  double rhobaredgef = 0.5 * (F.rhobar[c] + F.rhobar[f]);
  rhom_1 += rhobaredgef;
  rhom += rhobaredgef;
  double vm_1 = A.q[c];
  double momm_1 = rhom_1 * vm_1;
  double d = A.first ? 0.0 : A.dq[c];
  d = A.dt * force + A.am * d;
  A.dq[c] = d;
  double momm = momm_1 + A.bm * d;
  return momm / rhom;
}

__global__ void k_update_v(UpdArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  A.q[A.g.h(i, j, k)] = update_v_at(A, i, j, k, divergence(A.f, A.g, i, j, k));
}

// REUSE: the divergences of this stage's U and V updates are known to be in
// A.div_u / A.div_v (the fused stage); else they are read from there when
// given and recomputed from the stored fluxes otherwise.
template <bool REUSE>
__device__ __forceinline__ double update_w_at(const UpdArgs& A, int i, int j, int k, double fluxdiff) {
  // update!.jl:1035-1192
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k), up = g.h(i, j, k + 1);
  const double jc = F.jac[c], ju = F.jac[up];
  double jacedgeu = 2.0 * jc * ju / (jc + ju);
  fluxdiff /= jacedgeu;
  double fdu[2][2], fdv[2][2];
  if (REUSE || A.div_u) {  // the same values, as the U and V updates of this stage computed them
    for (int ll = 0; ll <= 1; ++ll)
      for (int mm = 0; mm <= 1; ++mm) {
        fdu[ll][mm] = A.div_u[g.h(i - ll, j, k + mm)];
        fdv[ll][mm] = A.div_v[g.h(i, j - ll, k + mm)];
      }
  } else {
    for (int ll = 0; ll <= 1; ++ll)
      for (int mm = 0; mm <= 1; ++mm) {
        fdu[ll][mm] = divergence(A.fu, g, i - ll, j, k + mm);
        double jacedger = 0.5 * (F.jac[g.h(i - ll, j, k + mm)] + F.jac[g.h(i + 1 - ll, j, k + mm)]);
        fdu[ll][mm] /= jacedger;
      }
    for (int ll = 0; ll <= 1; ++ll)
      for (int mm = 0; mm <= 1; ++mm) {
        fdv[ll][mm] = divergence(A.fv, g, i, j - ll, k + mm);
        double jacedgef = 0.5 * (F.jac[g.h(i, j - ll, k + mm)] + F.jac[g.h(i, j + 1 - ll, k + mm)]);
        fdv[ll][mm] /= jacedgef;
      }
  }
  fluxdiff = transform_transformed(F, g, i, j, k, fdu[0][0], fdu[0][1], fdu[1][0], fdu[1][1], fdv[0][0], fdv[0][1],
                                   fdv[1][0], fdv[1][1], fluxdiff);
  const double fc = A.ph.fc;
  double vc = 0.5 * (A.vold[c] + A.vold[g.h(i, j - 1, k)]);
  double vu = 0.5 * (A.vold[up] + A.vold[g.h(i, j - 1, k + 1)]);
  double uc = 0.5 * (A.uold[c] + A.uold[g.h(i - 1, j, k)]);
  double uu = 0.5 * (A.uold[up] + A.uold[g.h(i - 1, j, k + 1)]);
  double volforce = fc * (ju * F.m13[c] * (A.rhoold[c] + F.rhobar[c]) * vc +
                          jc * F.m13[up] * (A.rhoold[up] + F.rhobar[up]) * vu) / (jc + ju) -
                    fc * (ju * F.m23[c] * (A.rhoold[c] + F.rhobar[c]) * uc +
                          jc * F.m23[up] * (A.rhoold[up] + F.rhobar[up]) * uu) / (jc + ju);
  double force = -fluxdiff + volforce;
  double rhom_1 = (ju * A.rhoold[c] + jc * A.rhoold[up]) / (jc + ju);
  double rhom = (ju * F.rho[c] + jc * F.rho[up]) / (jc + ju);
  double rhobaredgeu = (ju * F.rhobar[c] + jc * F.rhobar[up]) / (jc + ju);
  rhom_1 += rhobaredgeu;
  rhom += rhobaredgeu;
  double wm_1 = A.q[c];
  double momm_1 = rhom_1 * wm_1;
  double d = A.first ? 0.0 : A.dq[c];
  d = A.dt * force + A.am * d;
  A.dq[c] = d;
  double momm = momm_1 + A.bm * d;
  return momm / rhom;
}

__global__ void k_update_w(UpdArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  A.q[A.g.h(i, j, k)] = update_w_at<false>(A, i, j, k, divergence(A.f, A.g, i, j, k));
}

// the W update of the fused stage from its stored flux divergence, then
// the sponge; the top interior level, which the W update leaves, is in R
// for the sponge alone
__global__ void k_update_w_fdiv(UpdArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = A.g.h(i, j, k);
  const double q = k < A.g.k1 ? update_w_at<true>(A, i, j, k, A.fdiv[c]) : A.q[c];
  A.q[c] = sponged<V_W>(A, i, j, k, q);
}

// ---------------------------------------------------------------------------
// Fused LHS stage kernels (W: the W fluxes and their divergence, stored in
// A.fdiv, then k_update_w_fdiv, because the W update alone takes about as
// many registers as the rest of the fused kernel): the fluxes of one
// variable at the faces of a cell, their divergence, the RK3 update and the
// LHS sponge in one pass, for the stage loop of explicit_integration!
// (explicit_integration!.jl:42-101).  No flux array is stored: a thread
// evaluates the x faces on both sides of its cells (flux_value), walks JC
// rows in y carrying the y face between them, and marches up a chunk of kc
// levels carrying the z face of each row (JC and kc: launch_stage's
// autotuner); the vertical wall
// faces, which set_boundaries!(state, BoundaryFluxes()) zeroes, are zero
// here too.  The flux inputs (phi, the Cartesian w, the stress components)
// are computed at the start of the stage and the p0 winds are fixed, so
// evaluating a variable's fluxes just before its update gives the bits the
// stored fluxes of compute_fluxes! had.  The divergence is summed as
// divergence() sums it.  Threads map to the i of the update range R, x
// fastest.
// ---------------------------------------------------------------------------
template <int VAR, class T>
__device__ __forceinline__ T zface(const FluxArgsT<T>& FA, int i, int j, int k) {
  if (VAR != V_W && (k == FA.g.k0 - 1 || k == FA.g.k1)) return T(0);
  return flux_value<VAR, 2, T>(FA, i, j, k);
}

template <int VAR, class T, int JC>
__global__ void __launch_bounds__(256) k_stage_update(FluxArgsT<T> FA, UpdArgs A, Range3 R, int kc) {
  // a thread: one i of the range, JC consecutive j (the y face between two
  // of them is evaluated once and carried) and a chunk of kc levels
  const int njc = (R.nj + JC - 1) / JC;
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)R.ni * njc) return;
  const int jc = (int)(t / R.ni);
  const int i = R.i0 + (int)(t - (long long)jc * R.ni);
  const int jb = R.j0 + jc * JC;
  const int nj = min(JC, R.j0 + R.nj - jb);
  const int kb = R.k0 + (int)blockIdx.y * kc;
  const int ke = min(kb + kc, R.k0 + R.nk);
  const Geo& g = A.g;
  T fzd[JC];
#pragma unroll
  for (int jj = 0; jj < JC; ++jj)
    if (jj < nj) fzd[jj] = zface<VAR, T>(FA, i, jb + jj, kb - 1);
#pragma unroll 1
  for (int k = kb; k < ke; ++k) {
    T fyb = flux_value<VAR, 1, T>(FA, i, jb - 1, k);
#pragma unroll
    for (int jj = 0; jj < JC; ++jj) {
      if (jj >= nj) break;
      const int j = jb + jj;
      const T fzu = zface<VAR, T>(FA, i, j, k);
      const T fxr = flux_value<VAR, 0, T>(FA, i, j, k), fxl = flux_value<VAR, 0, T>(FA, i - 1, j, k);
      const T fyf = flux_value<VAR, 1, T>(FA, i, j, k);
      double fluxdiff = (double)((fxr - fxl) / T(g.dx) + (fyf - fyb) / T(g.dy) + (fzu - fzd[jj]) / T(g.dz));
      const size_t c = g.h(i, j, k);
      if constexpr (VAR == V_RHO && sizeof(T) == 4) {  // the FP64 background part of the Rho flux
        const size_t sy = (size_t)g.nxx, sz = sy * g.nyy;
        const double bzu = k == g.k1 ? 0.0 : FA.bg[2][c], bzd = k - 1 == g.k0 - 1 ? 0.0 : FA.bg[2][c - sz];
        fluxdiff += (FA.bg[0][c] - FA.bg[0][c - 1]) / g.dx + (FA.bg[1][c] - FA.bg[1][c - sy]) / g.dy +
                    (bzu - bzd) / g.dz;
      }
      if (VAR == V_U) A.q[c] = sponged<V_U>(A, i, j, k, update_u_at(A, i, j, k, fluxdiff));
      else if (VAR == V_V) A.q[c] = sponged<V_V>(A, i, j, k, update_v_at(A, i, j, k, fluxdiff));
      else if (VAR == V_W) A.fdiv[c] = fluxdiff;  // the update follows in k_update_w_fdiv
      else A.q[c] = sponged<VAR == V_CHI ? V_RHO : VAR>(A, i, j, k, update_scalar_at(A, i, j, k, fluxdiff));
      fyb = fyf;
      fzd[jj] = fzu;
    }
  }
}

// ---------------------------------------------------------------------------
// RHS updates (update!.jl:600-727, 809-881, 960-1032, 1195-1342)
// ---------------------------------------------------------------------------
struct RhsArgs {
  Fields F;
  Geo g;
  Phys ph;
  double *rhop, *u, *v, *w;
  const double* pip;
  const double* betar;
  const double* rhopold;
  const double* wold;
  double dt, rf;
  int implicit;
};

__global__ void k_rhs_u(RhsArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k), r = g.h(i + 1, j, k);
  double rhoedger = 0.5 * (F.rho[c] + F.rho[r]);
  double rhobaredger = 0.5 * (F.rhobar[c] + F.rhobar[r]);
  rhoedger += rhobaredger;
  double gradient = pressure_gradient_u(F, A.ph, g, A.pip, i, j, k);
  double force = 0.0;
  double jpedger = wind_factor_u(F, g, i, j, k);
  if (!A.implicit) {
    A.u[c] += A.dt * (-gradient + force / rhoedger) * jpedger;
  } else {
    double factor = 1.0;
    if (A.ph.damp) factor += A.dt * 0.5 * (A.betar[c] + A.betar[r]) * A.rf;
    A.u[c] = 1.0 / factor * (A.u[c] + A.dt * (-gradient + force / rhoedger) * jpedger);
  }
}

__global__ void k_rhs_v(RhsArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k), f = g.h(i, j + 1, k);
  double rhoedgef = 0.5 * (F.rho[c] + F.rho[f]);
  double rhobaredgef = 0.5 * (F.rhobar[c] + F.rhobar[f]);
  rhoedgef += rhobaredgef;
  double gradient = pressure_gradient_v(F, A.ph, g, A.pip, i, j, k);
  double force = 0.0;
  double jpedgef = wind_factor_v(F, g, i, j, k);
  if (!A.implicit) {
    A.v[c] += A.dt * (-gradient + force / rhoedgef) * jpedgef;
  } else {
    double factor = 1.0;
    if (A.ph.damp) factor += A.dt * 0.5 * (A.betar[c] + A.betar[f]) * A.rf;
    A.v[c] = 1.0 / factor * (A.v[c] + A.dt * (-gradient + force / rhoedgef) * jpedgef);
  }
}

// COMP: the model is compressible (folds the model's branches of the wind
// and buoyancy factors at compile time; the runtime F.comp is overridden).
template <bool COMP>
__global__ void k_rhs_w(RhsArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  Fields F = A.F;
  F.comp = COMP ? 1 : 0;
  const size_t c = g.h(i, j, k), up = g.h(i, j, k + 1);
  const double jc = F.jac[c], ju = F.jac[up], dt = A.dt, gnd = A.ph.g_ndim;
  double rhoc = F.rho[c];
  double rhou = F.rho[up];
  double rhoedgeu = (ju * F.rho[c] + jc * F.rho[up]) / (jc + ju);
  rhoc += F.rhobar[c];
  rhou += F.rhobar[up];
  rhoedgeu += (ju * F.rhobar[c] + jc * F.rhobar[up]) / (jc + ju);
  double gradient = pressure_gradient_w(F, A.ph, g, A.pip, i, j, k);
  double force = 0.0;
  if (!A.implicit) {
    double b = -gnd * (ju * A.rhopold[c] / rhoc / jc + jc * A.rhopold[up] / rhou / ju) / (jc + ju);
    double jpedgeu = wind_factor_w(F, g, i, j, k);
    A.w[c] += dt * (b - gradient + force / rhoedgeu) * jpedgeu;
  } else {
    const double dt2 = dt * dt;
    double n2edgeu = (ju * F.n2[c] + jc * F.n2[up]) / (jc + ju);
    double factor = 1.0;
    factor += dt * (ju * A.betar[c] + jc * A.betar[up]) / (jc + ju) * A.rf;
    double b = -gnd * (ju * A.rhop[c] / rhoc / jc + jc * A.rhop[up] / rhou / ju) / (jc + ju);
    double uc = 0.5 * (A.u[c] / wind_factor_u(F, g, i, j, k) + A.u[g.h(i - 1, j, k)] / wind_factor_u(F, g, i - 1, j, k));
    double uu = 0.5 * (A.u[up] / wind_factor_u(F, g, i, j, k + 1) +
                       A.u[g.h(i - 1, j, k + 1)] / wind_factor_u(F, g, i - 1, j, k + 1));
    double vc = 0.5 * (A.v[c] / wind_factor_v(F, g, i, j, k) + A.v[g.h(i, j - 1, k)] / wind_factor_v(F, g, i, j - 1, k));
    double vu = 0.5 * (A.v[up] / wind_factor_v(F, g, i, j, k + 1) +
                       A.v[g.h(i, j - 1, k + 1)] / wind_factor_v(F, g, i, j - 1, k + 1));
    double jpedgeu = wind_factor_w(F, g, i, j, k);
    double fw = buoyancy_factor_w(F, g, i, j, k);
    A.w[c] = 1.0 / (factor + fw * n2edgeu * dt2) *
             (A.w[c] - dt * gradient * jpedgeu + dt * b * jpedgeu + dt * force / rhoedgeu * jpedgeu +
              jpedgeu * fw * n2edgeu * dt2 * (ju * (F.m13[c] * uc + F.m23[c] * vc) + jc * (F.m13[up] * uu + F.m23[up] * vu)) /
                  (jc + ju));
  }
}

// update!(state, dt, RhoP(), RHS(), Explicit()) (update!.jl:600-629); the
// Cartesian vertical winds come from the current u, v, w.
__global__ void k_rhs_rhop_explicit(RhsArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  const Fields& F = A.F;
  const size_t c = g.h(i, j, k);
  const double gnd = A.ph.g_ndim;
  double jpu = wind_factor_w(F, g, i, j, k), jpd = wind_factor_w(F, g, i, j, k - 1);
  double wvrt = 0.5 * (vertical_wind(F, g, i, j, k) / jpu + vertical_wind(F, g, i, j, k - 1) / jpd);
  double buoy = -gnd * A.rhop[c] / (F.rho[c] + F.rhobar[c]);
  double fb = buoyancy_factor_rhop(F, g, i, j, k);
  buoy -= A.dt * fb * F.n2[c] * wvrt;
  A.rhop[c] = -buoy * (F.rho[c] + F.rhobar[c]) / gnd;
}

// update!(state, dt, RhoP(), RHS(), Implicit(), rf) (update!.jl:632-727)
template <bool COMP>
__global__ void k_rhs_rhop_implicit(RhsArgs A, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  Fields F = A.F;
  F.comp = COMP ? 1 : 0;
  const size_t c = g.h(i, j, k), up = g.h(i, j, k + 1), dn = g.h(i, j, k - 1);
  const double dt = A.dt, dt2 = dt * dt, gnd = A.ph.g_ndim;
  double rhoc = F.rho[c] + F.rhobar[c];
  const double jc = F.jac[c], ju = F.jac[up], jd = F.jac[dn];
  double rhoedgeu = (ju * F.rho[c] + jc * F.rho[up]) / (jc + ju);
  rhoedgeu += (ju * F.rhobar[c] + jc * F.rhobar[up]) / (jc + ju);
  double rhoedged = (jd * F.rho[c] + jc * F.rho[dn]) / (jc + jd);
  rhoedged += (jd * F.rhobar[c] + jc * F.rhobar[dn]) / (jc + jd);
  double jpedgeu = wind_factor_w(F, g, i, j, k), jpedged = wind_factor_w(F, g, i, j, k - 1);
  double wv = 0.5 * (A.wold[c] / jpedgeu + A.wold[dn] / jpedged);
  double lower_gradient = pressure_gradient_w(F, A.ph, g, A.pip, i, j, k - 1);
  double lower_force = 0.0;
  double upper_gradient = pressure_gradient_w(F, A.ph, g, A.pip, i, j, k);
  double upper_force = 0.0;
  if (k == g.k0) {
    lower_gradient = 0.0;
    lower_force = 0.0;
  } else if (k == g.z_size + g.nbz) {
    upper_gradient = 0.0;
    upper_force = 0.0;
  }
  double gradient = 0.5 * (lower_gradient + upper_gradient);
  double force = 0.5 * (lower_force / rhoedged + upper_force / rhoedgeu) * rhoc;
  double factor = 1.0;
  factor += dt * A.betar[c] * A.rf;
  double b = -gnd * A.rhop[c] / (F.rho[c] + F.rhobar[c]);
  double jpedger = wind_factor_u(F, g, i, j, k), jpedgel = wind_factor_u(F, g, i - 1, j, k);
  double jpedgef = wind_factor_v(F, g, i, j, k), jpedgeb = wind_factor_v(F, g, i, j - 1, k);
  double fb = buoyancy_factor_rhop(F, g, i, j, k);
  const double n2c = F.n2[c];
  b = 1.0 / (factor + fb * n2c * dt2) *
      (-fb * n2c * dt * jc * (wv + dt * (-gradient + force / rhoc)) + factor * b +
       fb * n2c * dt * jc * factor * 0.5 *
           (F.m13[c] * (A.u[c] / jpedger + A.u[g.h(i - 1, j, k)] / jpedgel) +
            F.m23[c] * (A.v[c] / jpedgef + A.v[g.h(i, j - 1, k)] / jpedgeb)));
  A.rhop[c] = -b * (F.rho[c] + F.rhobar[c]) / gnd;
}

// update!(state, dt, PiP()) (update!.jl:1360-1392, compressible): pi' follows
// the divergence of the J P-weighted winds saved before the explicit RHS step
__global__ void k_rhs_pip(RhsArgs A, double* pip, const double* uold, const double* vold, Range3 R) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const Geo& g = A.g;
  const size_t c = g.h(i, j, k);
  double fluxdiff = (uold[c] - uold[g.h(i - 1, j, k)]) / g.dx + (vold[c] - vold[g.h(i, j - 1, k)]) / g.dy +
                    (A.wold[c] - A.wold[g.h(i, j, k - 1)]) / g.dz;
  fluxdiff /= A.F.jac[c];
  double heating = conductive_heating(A.F, g, i, j, k);
  double d = dpdpi(A.ph, A.F.p[c]);
  pip[c] -= A.dt * (fluxdiff - heating) / d;
}

// modify_compressible_wind!.jl:52-80 over the interior: u, v, w times
// (multiply) or divided by J P at their points
__global__ void k_modify_wind(double* u, double* v, double* w, const double* p, const double* jac, Geo g, Range3 R,
                              int multiply) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k), r = g.h(i + 1, j, k), f = g.h(i, j + 1, k), up = g.h(i, j, k + 1);
  double fu = (jac[c] * p[c] + jac[r] * p[r]) / 2;
  double fv = (jac[c] * p[c] + jac[f] * p[f]) / 2;
  double fw = jac[c] * jac[up] * (p[c] + p[up]) / (jac[c] + jac[up]);
  if (multiply) {
    u[c] = u[c] * fu;
    v[c] = v[c] * fv;
    w[c] = w[c] * fw;
  } else {
    u[c] = u[c] / fu;
    v[c] = v[c] / fv;
    w[c] = w[c] / fw;
  }
}

// synchronize_compressible_atmosphere!.jl:61-72 (pbar .= p) over the whole
// haloed array, with the Poisson scaling sqrt(pbar^2 / rhobar) that follows pbar
__global__ void k_sync_pbar(double* pbar, double* fcscal, const double* p, const double* rhobar, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const double v = p[q];
    pbar[q] = v;
    fcscal[q] = sqrt(v * v / rhobar[q]);
  }
}

// synchronize_density_fluctuations!.jl:76-86 (compressible): rhop = rho + rhobar - pbar / thetabar
__global__ void k_sync_rhop(double* rhop, const double* rho, const double* rhobar, const double* pbar,
                            const double* thetabar, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    rhop[q] = rho[q] + rhobar[q] - pbar[q] / thetabar[q];
}

// apply_lhs_sponge! of the compressible model: RhoP towards
// rhobar (1 - P / thetabar / (rho + rhobar)) (apply_lhs_sponge!.jl:301-328),
// P towards rhobar P / (rho + rhobar) (:580-602) and pi' (:542-568)
__global__ void k_lhs_sponge_comp(Fields F, Phys ph, double* rhop, double* p, double* pip, const double* alphar, Geo g,
                                  Range3 R, double dt, int var) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k);
  const double rhoc = F.rho[c], rb = F.rhobar[c];
  if (var == PCF_VAR_RHOP) {
    double rhopbg = rb * (1.0 - p[c] / F.thetabar[c] / (rhoc + rb));
    double alpha = alphar[c];
    double old = rhop[c];
    double beta = 1.0 / (1.0 + alpha * dt);
    rhop[c] = (1.0 - beta) * rhopbg + beta * old;
  } else if (var == PCF_VAR_P) {
    double pb = rb * p[c] / (rhoc + rb);
    double alpha = alphar[c];
    double old = p[c];
    double beta = 1 / (1 + alpha * dt);
    p[c] = (1 - beta) * pb + beta * old;
  } else {
    double d = dpdpi(ph, p[c]);
    double pib = rb * p[c] / (rhoc + rb) / d;
    double alpha = alphar[c];
    double old = pip[c];
    pip[c] = old - alpha * dt * (p[c] / d - pib);
  }
}

// ---------------------------------------------------------------------------
// sponges
// ---------------------------------------------------------------------------
// compute_sponges!.jl:26-65: betar = rhs_sponge(...) tref with the 1/dt of
// the closure applied here; alphar does not change.
__global__ void k_betar(const double* num, double* betar, size_t n, double dtd, double tref) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    betar[q] = num[q] / dtd * tref;
}

// apply_lhs_sponge!.jl:255-530 (relax_to_mean = false)
__global__ void k_lhs_sponge(double* q, const double* alphar, const double* jac, const double* m13, const double* m23,
                             Geo g, Range3 R, double dt, int var, double bgu, double bgv, double bgw_u, double bgw_v,
                             double bgw_w, double uref) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k);
  double bg, alpha;
  if (var == PCF_VAR_RHO || var == PCF_VAR_RHOP) {
    bg = 0.0;
    alpha = alphar[c];
  } else if (var == PCF_VAR_U) {
    bg = bgu;
    alpha = 0.5 * (alphar[c] + alphar[g.h(i + 1, j, k)]);
  } else if (var == PCF_VAR_V) {
    bg = bgv;
    alpha = 0.5 * (alphar[c] + alphar[g.h(i, j + 1, k)]);
  } else {
    const size_t up = g.h(i, j, k + 1);
    const double jc = jac[c], ju = jac[up];
    bg = (jc * (m13[up] * bgw_u + m23[up] * bgw_v + bgw_w / ju) + ju * (m13[c] * bgw_u + m23[c] * bgw_v + bgw_w / jc)) /
         (jc + ju) / uref;
    alpha = (ju * alphar[c] + jc * alphar[up]) / (jc + ju);
  }
  double old = q[c];
  double beta = 1.0 / (1.0 + alpha * dt);
  q[c] = (1.0 - beta) * bg + beta * old;
}

// the tracer's LHS sponge towards its initial state (apply_lhs_sponge!.jl:615-636)
__global__ void k_lhs_sponge_field(double* q, const double* alphar, const double* bg, Geo g, Range3 R, double dt) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k);
  double alpha = alphar[c];
  double old = q[c];
  double beta = 1.0 / (1.0 + alpha * dt);
  q[c] = (1.0 - beta) * bg[c] + beta * old;
}

// horizontal mean of one layer for relax_to_mean (sequential sum over the layer)
__global__ void k_layer_mean(const double* f, double* mean, Geo g, double nxy_x, double nxy_y) {
  const int k = g.k0 + blockIdx.x;
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int j = g.j0; j <= g.j1; ++j)
    for (int i = g.i0; i <= g.i1; ++i) s += f[g.h(i, j, k)] / nxy_x / nxy_y;
  mean[k - g.k0] = s;
}

__global__ void k_lhs_sponge_mean(double* q, const double* alphar, const double* jac, const double* mean, Geo g,
                                  Range3 R, double dt, int var) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k);
  double alpha;
  if (var == PCF_VAR_U) alpha = 0.5 * (alphar[c] + alphar[g.h(i + 1, j, k)]);
  else if (var == PCF_VAR_V) alpha = 0.5 * (alphar[c] + alphar[g.h(i, j + 1, k)]);
  else {
    const size_t up = g.h(i, j, k + 1);
    alpha = (jac[up] * alphar[c] + jac[c] * alphar[up]) / (jac[c] + jac[up]);
  }
  double bg = mean[k - g.k0];
  double old = q[c];
  double beta = 1.0 / (1.0 + alpha * dt);
  q[c] = (1.0 - beta) * bg + beta * old;
}

// maxima of |u|, |v|, |w| over the interior (compute_time_step.jl:71-73)
__global__ void k_cfl(const double* u, const double* v, const double* w, Geo g, double* partials) {
  double m[3] = {0.0, 0.0, 0.0};
  const long long n = (long long)g.nx * g.ny * g.nz;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    int i = g.i0 + (int)(t % g.nx);
    int j = g.j0 + (int)((t / g.nx) % g.ny);
    int k = g.k0 + (int)(t / ((long long)g.nx * g.ny));
    size_t c = g.h(i, j, k);
    m[0] = dmax(m[0], fabs(u[c]));
    m[1] = dmax(m[1], fabs(v[c]));
    m[2] = dmax(m[2], fabs(w[c]));
  }
  __shared__ double red[3][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int d = 0; d < 3; ++d) {
    double x = m[d];
    for (int off = 32; off > 0; off >>= 1) x = dmax(x, __shfl_down(x, off, 64));
    if (lane == 0) red[d][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) {
      double x = red[d][0];
      for (int q = 1; q < 4; ++q) x = dmax(x, red[d][q]);
      partials[d * gridDim.x + blockIdx.x] = x;
    }
}

__global__ void k_cfl_final(const double* partials, int nparts, double* out) {
  if (threadIdx.x != 0) return;
  for (int d = 0; d < 3; ++d) {
    double x = 0.0;
    for (int q = 0; q < nparts; ++q) x = dmax(x, partials[d * nparts + q]);
    out[d] = x;
  }
}

__global__ void k_copy_field(double* dst, const double* src, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

Fields fields_of(pcf_state* st) { return make_fields(st); }

template <int VAR, int DIR>
void launch_flux(pcf_state* st, const FluxArgs& A, Range3 R) {
  if (R.ni <= 0 || R.nj <= 0 || R.nk <= 0) return;
  hipLaunchKernelGGL((k_flux<VAR, DIR>), dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
}

}  // namespace

// reconstruct! (reconstruct!.jl:102-315): the reconstruction inputs phi of
// every variable and the Cartesian vertical wind they and the stress tensor
// need.  The limited face values themselves are evaluated by the flux
// kernels, so set_boundaries!(state, BoundaryReconstructions()) has nothing
// to fill: halo entries of phi are the periodic images of interior ones.
int reconstruct(pcf_state* st) {
  ProfScope prof(st, PCF_K_FLUXES);
  const Geo& g = st->g;
  Fields F = fields_of(st);
  const bool pi = st->cfg.model != PCF_BOUSSINESQ;
  hipLaunchKernelGGL(k_inputs<double>, dim3(blocks_for((long long)g.nxx * g.nyy, 256), g.nzz - 1), dim3(256), 0,
                     st->stream, F, g, st->phi[0], st->phi[1], st->phi[2], st->phi[3], pi ? 1 : 0, st->chi,
                     st->tracer ? st->phi[V_CHI] : nullptr, st->wcart, st->phi[4]);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// reconstruct! + compute_fluxes!(state, p0) + set_boundaries!(state,
// BoundaryFluxes()) (explicit_integration!.jl:53-58); the reconstruction is
// evaluated here again, so a caller of the reference's sequence (which calls
// reconstruct first) and a caller of this function alone get the same fluxes.
namespace {

// the six stress components compute_stress_tensor gives the viscous fluxes
// (compute_stress_tensor.jl:55-164), from the current winds
// (FP32 copies with flux_precision MIXED and f32)
int stress_tensor(pcf_state* st, bool f32 = false) {
  if (!st->ph.viscous) return PCF_OK;
  const Geo& g = st->g;
  Range3 R{g.i0 - 1, g.nx + 2, g.j0 - 1, g.ny + 2, g.k0, g.nz};
  if (f32)
    hipLaunchKernelGGL(k_stress<float>, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, fields_of(st),
                       g, st->wcart, st->stress32[0], st->stress32[1], st->stress32[2], st->stress32[3],
                       st->stress32[4], st->stress32[5], st->tstress32[0], st->tstress32[1], R);
  else
    hipLaunchKernelGGL(k_stress<double>, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, fields_of(st),
                       g, st->wcart, st->stress[0], st->stress[1], st->stress[2], st->stress[3], st->stress[4],
                       st->stress[5], st->tstress[0], st->tstress[1], R);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

bool flux32(const pcf_state* st) { return st->cfg.flux_precision == PCF_PRECISION_MIXED; }

// the face carriers of the p0 winds for the flux evaluation (FP32 copies with f32)
int carriers(pcf_state* st, bool f32) {
  const Geo& g = st->g;
  const dim3 grd(blocks_for(g.nxx, 64), g.nyy, g.nzz), blk(64);
  if (f32)
    hipLaunchKernelGGL(k_carriers<float>, grd, blk, 0, st->stream, fields_of(st), g, st->p0[2], st->p0[3], st->p0[4],
                       st->carrier32[0], st->carrier32[1], st->carrier32[2]);
  else
    hipLaunchKernelGGL(k_carriers<double>, grd, blk, 0, st->stream, fields_of(st), g, st->p0[2], st->p0[3],
                       st->p0[4], st->carrier[0], st->carrier[1], st->carrier[2]);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// the FP32 flux inputs of the RK stages (flux_precision MIXED)
FluxArgsT<float> flux_args32(pcf_state* st) {
  FluxArgsT<float> A;
  A.F = fields_of(st);
  for (int q = 0; q < 6; ++q) A.phi[q] = st->phi32[q];
  A.cx = st->carrier32[0];
  A.cy = st->carrier32[1];
  A.cz = st->carrier32[2];
  A.u0 = st->p0[2];
  A.v0 = st->p0[3];
  A.w0 = st->p0[4];
  for (int q = 0; q < 6; ++q) A.st[q] = st->stress32[q];
  A.t13 = st->tstress32[0];
  A.t23 = st->tstress32[1];
  for (int d = 0; d < 3; ++d) A.bg[d] = st->rho_bg[d];
  for (int q = 0; q < 7; ++q)
    for (int d = 0; d < 3; ++d) A.flux[q][d] = nullptr;
  A.g = st->g;
  A.inv_re = (float)(1 / st->ph.re);
  A.viscous = st->ph.viscous;
  return A;
}

FluxArgs flux_args(pcf_state* st) {
  FluxArgs A;
  A.F = fields_of(st);
  for (int q = 0; q < 6; ++q) A.phi[q] = st->phi[q];
  A.cx = st->carrier[0];
  A.cy = st->carrier[1];
  A.cz = st->carrier[2];
  A.u0 = st->p0[2];
  A.v0 = st->p0[3];
  A.w0 = st->p0[4];
  for (int q = 0; q < 6; ++q) A.st[q] = st->stress[q];
  A.t13 = st->tstress[0];
  A.t23 = st->tstress[1];
  for (int d = 0; d < 3; ++d) A.bg[d] = nullptr;  // the FP64 Rho flux carries its background itself
  for (int q = 0; q < 7; ++q)
    for (int d = 0; d < 3; ++d) A.flux[q][d] = st->flux[q][d];
  A.g = st->g;
  A.inv_re = 1 / st->ph.re;
  A.viscous = st->ph.viscous;
  return A;
}

// arguments of update!(state, dt, m, var, LHS()) common to every variable
UpdArgs upd_args(pcf_state* st, double dt, int m);

}  // namespace

// Whether the fused stage kernel of var applies its LHS sponge itself: the
// sponges of k_lhs_sponge (relaxation to fixed values; the compressible
// model's, the tracer's and relax_to_mean's, which needs the layer mean of
// the updated field, stay separate launches)
bool sponge_fused(const pcf_state* st, int var) {
  if (var == PCF_VAR_RHO) return st->cfg.model != PCF_BOUSSINESQ && !st->compressible;
  if (var == PCF_VAR_RHOP) return !st->compressible;
  if (var == PCF_VAR_U || var == PCF_VAR_V || var == PCF_VAR_W) return !st->cfg.relax_to_mean;
  return false;
}

namespace {

// the sponge arguments of the fused stage kernels (those of k_lhs_sponge)
void sponge_args(const pcf_state* st, int var, double sdt, UpdArgs& A) {
  const pcf_config& n = st->cfg;
  A.sponge = sponge_fused(st, var) ? 1 : 0;
  A.alphar = st->alphar;
  A.sdt = sdt;
  const double uref = st->uref;
  A.bgu = (n.relaxed_u + n.relaxed_u) / uref / 2;
  A.bgv = (n.relaxed_v + n.relaxed_v) / uref / 2;
  A.bgw_u = n.relaxed_u;
  A.bgw_v = n.relaxed_v;
  A.bgw_w = n.relaxed_w;
  A.uref = uref;
}

UpdArgs upd_args(pcf_state* st, double dt, int m) {
  UpdArgs A;
  A.F = fields_of(st);
  A.g = st->g;
  A.ph = st->ph;
  A.dt = dt;
  A.am = st->alphark[m - 1];
  A.bm = st->betark[m - 1];
  A.first = m == 1;
  A.force0 = 0;
  A.heat = 0;
  A.div_out = nullptr;
  A.div_u = A.div_v = nullptr;
  A.fdiv = nullptr;
  A.rhoold = st->rhoold;
  A.uold = st->uold;
  A.vold = st->vold;
  for (int d = 0; d < 3; ++d) {
    A.fu[d] = st->flux[V_U][d];
    A.fv[d] = st->flux[V_V][d];
    A.f[d] = nullptr;
  }
  A.q = A.dq = nullptr;
  A.sponge = 0;
  A.alphar = nullptr;
  A.sdt = A.bgu = A.bgv = A.bgw_u = A.bgw_v = A.bgw_w = A.uref = 0.0;
  return A;
}

}  // namespace

int compute_fluxes(pcf_state* st) {
  int rc = reconstruct(st);
  if (rc) return rc;
  ProfScope prof(st, PCF_K_FLUXES);
  const Geo& g = st->g;
  st->lhs_div_valid[0] = st->lhs_div_valid[1] = false;  // the stress arrays are rewritten below
  const bool pi = st->cfg.model != PCF_BOUSSINESQ;
  if ((rc = stress_tensor(st))) return rc;
  if ((rc = carriers(st, false))) return rc;
  FluxArgs A = flux_args(st);
  const int nx = g.nx, ny = g.ny, nz = g.nz, i0 = g.i0, j0 = g.j0, k0 = g.k0;
  // faces the updates consume (update!.jl loop ranges); vertical wall faces
  // are zeroed below
  if (pi) {
    launch_flux<V_RHO, 0>(st, A, Range3{i0 - 1, nx + 1, j0, ny, k0, nz});
    launch_flux<V_RHO, 1>(st, A, Range3{i0, nx, j0 - 1, ny + 1, k0, nz});
    launch_flux<V_RHO, 2>(st, A, Range3{i0, nx, j0, ny, k0, nz - 1});
  }
  launch_flux<V_RHOP, 0>(st, A, Range3{i0 - 1, nx + 1, j0, ny, k0, nz});
  launch_flux<V_RHOP, 1>(st, A, Range3{i0, nx, j0 - 1, ny + 1, k0, nz});
  launch_flux<V_RHOP, 2>(st, A, Range3{i0, nx, j0, ny, k0, nz - 1});
  launch_flux<V_U, 0>(st, A, Range3{i0 - 2, nx + 2, j0, ny, k0, nz});
  launch_flux<V_U, 1>(st, A, Range3{i0 - 1, nx + 1, j0 - 1, ny + 1, k0, nz});
  launch_flux<V_U, 2>(st, A, Range3{i0 - 1, nx + 1, j0, ny, k0, nz - 1});
  launch_flux<V_V, 0>(st, A, Range3{i0 - 1, nx + 1, j0 - 1, ny + 1, k0, nz});
  launch_flux<V_V, 1>(st, A, Range3{i0, nx, j0 - 2, ny + 2, k0, nz});
  launch_flux<V_V, 2>(st, A, Range3{i0, nx, j0 - 1, ny + 1, k0, nz - 1});
  launch_flux<V_W, 0>(st, A, Range3{i0 - 1, nx + 1, j0, ny, k0, nz - 1});
  launch_flux<V_W, 1>(st, A, Range3{i0, nx, j0 - 1, ny + 1, k0, nz - 1});
  launch_flux<V_W, 2>(st, A, Range3{i0, nx, j0, ny, k0 - 1, nz});
  if (st->tracer) {
    launch_flux<V_CHI, 0>(st, A, Range3{i0 - 1, nx + 1, j0, ny, k0, nz});
    launch_flux<V_CHI, 1>(st, A, Range3{i0, nx, j0 - 1, ny + 1, k0, nz});
    launch_flux<V_CHI, 2>(st, A, Range3{i0, nx, j0, ny, k0, nz - 1});
  }
  if (st->compressible) {
    Range3 rx{i0 - 1, nx + 1, j0, ny, k0, nz}, ry{i0, nx, j0 - 1, ny + 1, k0, nz}, rz{i0, nx, j0, ny, k0, nz - 1};
    hipLaunchKernelGGL(k_flux_p<0>, dim3(blocks_for(rx.count(), 256)), dim3(256), 0, st->stream, A, rx);
    hipLaunchKernelGGL(k_flux_p<1>, dim3(blocks_for(ry.count(), 256)), dim3(256), 0, st->stream, A, ry);
    if (rz.nk > 0) hipLaunchKernelGGL(k_flux_p<2>, dim3(blocks_for(rz.count(), 256)), dim3(256), 0, st->stream, A, rz);
  }
  PCF_HIP(hipGetLastError());
  return set_boundaries_fluxes(st);
}

// compute_fluxes!(state, p0, Theta()) (compute_fluxes!.jl:1592-1756),
// compressible model with a nonzero thermal conductivity
int compute_theta_fluxes(pcf_state* st) {
  if (!st->compressible || st->cfg.thermal_conductivity == 0.0) return PCF_OK;
  ProfScope prof(st, PCF_K_FLUXES);
  const Geo& g = st->g;
  ThetaArgs A;
  A.rho0 = st->p0[0]; A.rhobar = st->rhobar; A.pbar = st->pbar; A.jac = st->jac;
  A.m13 = st->met13; A.m23 = st->met23; A.m33 = st->met33;
  for (int d = 0; d < 3; ++d) A.f[d] = st->phitheta[d];
  A.g = g;
  A.mu = st->cfg.thermal_conductivity / st->uref / st->lref;
  Range3 rx{g.i0 - 1, g.nx + 1, g.j0, g.ny, g.k0, g.nz}, ry{g.i0, g.nx, g.j0 - 1, g.ny + 1, g.k0, g.nz};
  Range3 rz{g.i0, g.nx, g.j0, g.ny, g.k0 - 1, g.nz + 1};
  hipLaunchKernelGGL(k_flux_theta<0>, dim3(blocks_for(rx.count(), 256)), dim3(256), 0, st->stream, A, rx);
  hipLaunchKernelGGL(k_flux_theta<1>, dim3(blocks_for(ry.count(), 256)), dim3(256), 0, st->stream, A, ry);
  hipLaunchKernelGGL(k_flux_theta<2>, dim3(blocks_for(rz.count(), 256)), dim3(256), 0, st->stream, A, rz);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int modify_compressible_wind(pcf_state* st, bool multiply) {
  if (!st->compressible) return PCF_OK;
  ProfScope prof(st, PCF_K_UPDATES);
  const Geo& g = st->g;
  Range3 R{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
  hipLaunchKernelGGL(k_modify_wind, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, st->u, st->v, st->w,
                     st->p, st->jac, g, R, multiply ? 1 : 0);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int synchronize_compressible_atmosphere(pcf_state* st, const double* p) {
  if (!st->compressible) return PCF_OK;
  ProfScope prof(st, PCF_K_UPDATES);
  hipLaunchKernelGGL(k_sync_pbar, dim3(2048), dim3(256), 0, st->stream, st->pbar, st->fcscal, p, st->rhobar,
                     st->g.nh());
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// sqrt(pbar^2 / rhobar) after pbar or rhobar changed from the host
int refresh_fcscal(pcf_state* st) {
  hipLaunchKernelGGL(k_sync_pbar, dim3(2048), dim3(256), 0, st->stream, st->pbar, st->fcscal, st->pbar, st->rhobar,
                     st->g.nh());
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int synchronize_density_fluctuations(pcf_state* st) {
  ProfScope prof(st, PCF_K_UPDATES);
  if (st->cfg.model == PCF_PSEUDO_INCOMPRESSIBLE) return copy_field(st, st->rhop, st->rho);
  if (st->cfg.model == PCF_COMPRESSIBLE) {
    hipLaunchKernelGGL(k_sync_rhop, dim3(2048), dim3(256), 0, st->stream, st->rhop, st->rho, st->rhobar, st->pbar,
                       st->thetabar, st->g.nh());
    PCF_HIP(hipGetLastError());
  }
  return PCF_OK;
}

int update_lhs(pcf_state* st, double dt, int m, int var) {
  ProfScope prof(st, PCF_K_UPDATES);
  const Geo& g = st->g;
  const bool pi = st->cfg.model != PCF_BOUSSINESQ;
  UpdArgs A = upd_args(st, dt, m);
  Range3 R;
  switch (var) {
    case PCF_VAR_RHO:
    case PCF_VAR_RHOP: {
      if (var == PCF_VAR_RHO && !pi) return PCF_OK;
      const int q = var == PCF_VAR_RHO ? V_RHO : V_RHOP;
      A.q = var == PCF_VAR_RHO ? st->rho : st->rhop;
      A.dq = var == PCF_VAR_RHO ? st->drho : st->drhop;
      A.heat = var == PCF_VAR_RHOP && st->compressible ? 1 : 0;
      for (int d = 0; d < 3; ++d) A.f[d] = st->flux[q][d];
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
      hipLaunchKernelGGL(k_update_scalar, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    }
    case PCF_VAR_P: {  // update!.jl:1400-1448
      if (!st->compressible) return PCF_OK;
      A.q = st->p;
      A.dq = st->dp;
      A.heat = 2;
      for (int d = 0; d < 3; ++d) A.f[d] = st->flux[V_P][d];
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
      hipLaunchKernelGGL(k_update_scalar, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    }
    case PCF_VAR_CHI: {  // update!.jl:1460-1500
      if (!st->tracer) return PCF_OK;
      A.q = st->chi;
      A.dq = st->dchi;
      A.force0 = 1;
      for (int d = 0; d < 3; ++d) A.f[d] = st->flux[V_CHI][d];
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
      hipLaunchKernelGGL(k_update_scalar, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    }
    case PCF_VAR_U:
      A.q = st->u;
      A.dq = st->du;
      for (int d = 0; d < 3; ++d) A.f[d] = st->flux[V_U][d];
      // the stress arrays are free between flux computations: keep the
      // divergences there for the W update of this stage
      A.div_out = st->stress[0];
      R = Range3{g.i0 - 1, g.nx + 1, g.j0, g.ny, g.k0, g.nz};
      hipLaunchKernelGGL(k_update_u, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      st->lhs_div_valid[0] = true;
      break;
    case PCF_VAR_V:
      A.q = st->v;
      A.dq = st->dv;
      for (int d = 0; d < 3; ++d) A.f[d] = st->flux[V_V][d];
      A.div_out = st->stress[1];
      R = Range3{g.i0, g.nx, g.j0 - 1, g.ny + 1, g.k0, g.nz};
      hipLaunchKernelGGL(k_update_v, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      st->lhs_div_valid[1] = true;
      break;
    case PCF_VAR_W:
      A.q = st->w;
      A.dq = st->dw;
      if (st->lhs_div_valid[0] && st->lhs_div_valid[1]) {
        A.div_u = st->stress[0];
        A.div_v = st->stress[1];
      }
      for (int d = 0; d < 3; ++d) A.f[d] = st->flux[V_W][d];
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz - 1};
      if (R.nk > 0)
        hipLaunchKernelGGL(k_update_w, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    default:
      return PCF_E_ARG;
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// The inputs of a stage's fluxes: reconstruct! and the stress tensor from
// the current predictands (compute_fluxes without its flux kernels), and on
// the first stage (m = 1) the face carriers, which depend only on the p0
// winds, J and pbar, none of which the three RK stages change.
int stage_inputs(pcf_state* st, int m) {
  if (!flux32(st)) {
    int rc = reconstruct(st);
    if (rc) return rc;
    ProfScope prof(st, PCF_K_FLUXES);
    if (m == 1 && (rc = carriers(st, false))) return rc;
    return stress_tensor(st);
  }
  // flux_precision MIXED: the same inputs, computed in FP64 and stored in FP32
  ProfScope prof(st, PCF_K_FLUXES);
  const Geo& g = st->g;
  hipLaunchKernelGGL(k_inputs<float>, dim3(blocks_for((long long)g.nxx * g.nyy, 256), g.nzz - 1), dim3(256), 0,
                     st->stream, fields_of(st), g, st->phi32[0], st->phi32[1], st->phi32[2], st->phi32[3],
                     st->cfg.model != PCF_BOUSSINESQ ? 1 : 0, st->chi, st->tracer ? st->phi32[V_CHI] : nullptr,
                     st->wcart, st->phi32[4]);
  PCF_HIP(hipGetLastError());
  int rc = m == 1 ? carriers(st, true) : PCF_OK;
  if (rc) return rc;
  if (m == 1 && st->rho_bg[0]) {
    const dim3 grd(blocks_for(g.nxx, 64), g.nyy, g.nzz), blk(64);
    hipLaunchKernelGGL(k_rho_background, grd, blk, 0, st->stream, fields_of(st), g, st->p0[2], st->p0[3], st->p0[4],
                       st->rho_bg[0], st->rho_bg[1], st->rho_bg[2]);
    PCF_HIP(hipGetLastError());
  }
  return stress_tensor(st, true);
}

// update!(state, dt, m, var, LHS()) of one RK stage with the variable's
// fluxes (compute_fluxes!) evaluated in the same kernel (k_stage_update),
// after stage_inputs: the sequence of explicit_integration!.jl:53-99 without
// storing the fluxes.  The U and V updates leave their flux divergences in
// div_u / div_v for the W update of the same stage.
namespace {

// Launch shapes of the fused stage kernels: rows per thread (the template
// argument JC) x levels per thread (kc).  Every shape evaluates the same
// faces with the same arithmetic, so the choice never changes a bit.
constexpr int kTuneRows[3] = {1, 2, 4};
constexpr int kTuneLevels[3] = {4, 8, 16};

template <int VAR, class T>
void launch_stage_shape(pcf_state* st, const FluxArgsT<T>& FA, const UpdArgs& A, const Range3& R, int shape) {
  const int jc = kTuneRows[shape / 3], kc = kTuneLevels[shape % 3];
  dim3 grd(blocks_for((long long)R.ni * ((R.nj + jc - 1) / jc), 256), (unsigned)((R.nk + kc - 1) / kc));
  if (jc == 1) hipLaunchKernelGGL((k_stage_update<VAR, T, 1>), grd, dim3(256), 0, st->stream, FA, A, R, kc);
  else if (jc == 2) hipLaunchKernelGGL((k_stage_update<VAR, T, 2>), grd, dim3(256), 0, st->stream, FA, A, R, kc);
  else hipLaunchKernelGGL((k_stage_update<VAR, T, 4>), grd, dim3(256), 0, st->stream, FA, A, R, kc);
}

// The project's autotuner for the fused stage kernels: the first time a
// variable's kernel runs on a state, every launch shape is timed on copies
// of the variable and its increment (the state itself is not touched) and
// the fastest is kept for the rest of the run.
template <int VAR, class T>
int launch_stage(pcf_state* st, const FluxArgsT<T>& FA, const UpdArgs& A, const Range3& R) {
  if (R.ni <= 0 || R.nj <= 0 || R.nk <= 0) return PCF_OK;
  const int slot = VAR * 2 + (sizeof(T) == 4 ? 1 : 0);
  int& best = st->stage_shape[slot];
  if (best < 0) {
    UpdArgs T_args = A;
    const size_t bytes = st->g.nh() * sizeof(double);
    PCF_HIP(hipMemcpyAsync(st->tune_q, A.q, bytes, hipMemcpyDeviceToDevice, st->stream));
    PCF_HIP(hipMemcpyAsync(st->tune_dq, A.dq, bytes, hipMemcpyDeviceToDevice, st->stream));
    T_args.q = st->tune_q;
    T_args.dq = st->tune_dq;
    if (A.div_out) T_args.div_out = st->tune_div;
    hipEvent_t e0, e1;
    PCF_HIP(hipEventCreate(&e0));
    PCF_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    for (int shape = 0; shape < 9; ++shape) {
      launch_stage_shape<VAR, T>(st, FA, T_args, R, shape);  // warm-up
      PCF_HIP(hipEventRecord(e0, st->stream));
      for (int rep = 0; rep < 2; ++rep) launch_stage_shape<VAR, T>(st, FA, T_args, R, shape);
      PCF_HIP(hipEventRecord(e1, st->stream));
      PCF_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      PCF_HIP(hipEventElapsedTime(&ms, e0, e1));
      st->stage_shape_ms[slot][shape] = 0.5f * ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = shape;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  launch_stage_shape<VAR, T>(st, FA, A, R, best);
  return PCF_OK;
}

}  // namespace

int stage_update(pcf_state* st, double dt, int m, int var, double sdt) {
  const Geo& g = st->g;
  const bool pi = st->cfg.model != PCF_BOUSSINESQ;
  UpdArgs A = upd_args(st, dt, m);
  sponge_args(st, var, sdt, A);
  Range3 R{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
  const bool f32 = flux32(st);
  // k_stage_update<VAR, double, JC> or, with flux_precision MIXED, <VAR, float, JC>
  int trc = PCF_OK;
  auto launch = [&](auto tag) {
    constexpr int VAR = decltype(tag)::value;
    trc = f32 ? launch_stage<VAR, float>(st, flux_args32(st), A, R) : launch_stage<VAR, double>(st, flux_args(st), A, R);
  };
  if (var == PCF_VAR_P) {  // the carrier flux of P has no reconstruction: stored, then the update
    if (!st->compressible) return PCF_OK;
    FluxArgs FA = flux_args(st);
    Range3 rx{g.i0 - 1, g.nx + 1, g.j0, g.ny, g.k0, g.nz}, ry{g.i0, g.nx, g.j0 - 1, g.ny + 1, g.k0, g.nz},
        rz{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz - 1};
    {
      ProfScope prof(st, PCF_K_FLUXES);
      hipLaunchKernelGGL(k_flux_p<0>, dim3(blocks_for(rx.count(), 256)), dim3(256), 0, st->stream, FA, rx);
      hipLaunchKernelGGL(k_flux_p<1>, dim3(blocks_for(ry.count(), 256)), dim3(256), 0, st->stream, FA, ry);
      if (rz.nk > 0)
        hipLaunchKernelGGL(k_flux_p<2>, dim3(blocks_for(rz.count(), 256)), dim3(256), 0, st->stream, FA, rz);
      PCF_HIP(hipGetLastError());
      int rc = set_boundaries_fluxes(st);
      if (rc) return rc;
    }
    return update_lhs(st, dt, m, PCF_VAR_P);
  }
  ProfScope prof(st, PCF_K_FLUXES);
  switch (var) {
    case PCF_VAR_RHO:
      if (!pi) return PCF_OK;
      A.q = st->rho;
      A.dq = st->drho;
      launch(std::integral_constant<int, V_RHO>());
      if (trc) return trc;
      break;
    case PCF_VAR_RHOP:
      A.q = st->rhop;
      A.dq = st->drhop;
      A.heat = st->compressible ? 1 : 0;
      launch(std::integral_constant<int, V_RHOP>());
      if (trc) return trc;
      break;
    case PCF_VAR_CHI:
      if (!st->tracer) return PCF_OK;
      A.q = st->chi;
      A.dq = st->dchi;
      A.force0 = 1;
      launch(std::integral_constant<int, V_CHI>());
      if (trc) return trc;
      break;
    case PCF_VAR_U:
      A.q = st->u;
      A.dq = st->du;
      A.div_out = st->div_u;
      R = Range3{g.i0 - 1, g.nx + 1, g.j0, g.ny, g.k0, g.nz};
      launch(std::integral_constant<int, V_U>());
      if (trc) return trc;
      break;
    case PCF_VAR_V:
      A.q = st->v;
      A.dq = st->dv;
      A.div_out = st->div_v;
      R = Range3{g.i0, g.nx, g.j0 - 1, g.ny + 1, g.k0, g.nz};
      launch(std::integral_constant<int, V_V>());
      if (trc) return trc;
      break;
    case PCF_VAR_W:
      A.q = st->w;
      A.dq = st->dw;
      A.div_u = st->div_u;
      A.div_v = st->div_v;
      A.fdiv = st->fdiv_w;
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz - 1};
      launch(std::integral_constant<int, V_W>());
      if (trc) return trc;
      {
        // the update, and with a fused sponge the top interior level too
        Range3 rw{g.i0, g.nx, g.j0, g.ny, g.k0, A.sponge ? g.nz : g.nz - 1};
        if (rw.nk > 0)
          hipLaunchKernelGGL(k_update_w_fdiv, dim3(blocks_for(rw.count(), 256)), dim3(256), 0, st->stream, A, rw);
      }
      break;
    default:
      return PCF_E_ARG;
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int update_rhs(pcf_state* st, double dt, int var, bool implicit, double rf) {
  ProfScope prof(st, PCF_K_UPDATES);
  const Geo& g = st->g;
  RhsArgs A;
  A.F = fields_of(st);
  A.g = g;
  A.ph = st->ph;
  A.rhop = st->rhop; A.u = st->u; A.v = st->v; A.w = st->w;
  A.pip = st->pip;
  A.betar = st->betar;
  A.rhopold = st->rhopold;
  A.wold = st->wold;
  A.dt = dt;
  A.rf = rf;
  A.implicit = implicit ? 1 : 0;
  Range3 R;
  switch (var) {
    case PCF_VAR_U:
      R = Range3{g.i0 - 1, g.nx + 1, g.j0, g.ny, g.k0, g.nz};
      hipLaunchKernelGGL(k_rhs_u, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    case PCF_VAR_V:
      R = Range3{g.i0, g.nx, g.j0 - 1, g.ny + 1, g.k0, g.nz};
      hipLaunchKernelGGL(k_rhs_v, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    case PCF_VAR_W:
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz - 1};
      if (R.nk > 0) {
        if (st->compressible)
          hipLaunchKernelGGL(k_rhs_w<true>, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
        else
          hipLaunchKernelGGL(k_rhs_w<false>, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      }
      break;
    case PCF_VAR_PIP:  // update!.jl:1351-1392: explicit, compressible only
      if (!st->compressible || implicit) return PCF_OK;
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
      hipLaunchKernelGGL(k_rhs_pip, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, st->pip, st->uold,
                         st->vold, R);
      break;
    case PCF_VAR_RHOP:
      R = Range3{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
      if (implicit && st->compressible)
        hipLaunchKernelGGL(k_rhs_rhop_implicit<true>, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A,
                           R);
      else if (implicit)
        hipLaunchKernelGGL(k_rhs_rhop_implicit<false>, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A,
                           R);
      else
        hipLaunchKernelGGL(k_rhs_rhop_explicit, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, A, R);
      break;
    default:
      return PCF_E_ARG;
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int apply_lhs_sponge(pcf_state* st, double dt, double time, int var) {
  ProfScope prof(st, PCF_K_UPDATES);
  (void)time;
  const Geo& g = st->g;
  const pcf_config& n = st->cfg;
  if (var == PCF_VAR_PIP || var == PCF_VAR_P || (var == PCF_VAR_RHOP && st->compressible)) {
    if (!st->compressible) return PCF_OK;
    Range3 R{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
    hipLaunchKernelGGL(k_lhs_sponge_comp, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, fields_of(st),
                       st->ph, st->rhop, st->p, st->pip, st->alphar, g, R, dt, var);
    PCF_HIP(hipGetLastError());
    return PCF_OK;
  }
  if (var == PCF_VAR_RHO && n.model == PCF_BOUSSINESQ) return PCF_OK;
  if (var == PCF_VAR_CHI) {
    if (!st->tracer) return PCF_OK;
    Range3 R{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
    hipLaunchKernelGGL(k_lhs_sponge_field, dim3(blocks_for(R.count(), 256)), dim3(256), 0, st->stream, st->chi,
                       st->alphar, st->chi_init, g, R, dt);
    PCF_HIP(hipGetLastError());
    return PCF_OK;
  }
  double* q = var == PCF_VAR_RHO ? st->rho : var == PCF_VAR_RHOP ? st->rhop : var == PCF_VAR_U ? st->u
            : var == PCF_VAR_V ? st->v : st->w;
  Range3 R{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
  const unsigned blocks = blocks_for(R.count(), 256);
  if (n.relax_to_mean && (var == PCF_VAR_U || var == PCF_VAR_V || var == PCF_VAR_W)) {
    hipLaunchKernelGGL(k_layer_mean, dim3(g.nz), dim3(64), 0, st->stream, q, st->pc_q, g, (double)n.x_size,
                       (double)n.y_size);
    hipLaunchKernelGGL(k_lhs_sponge_mean, dim3(blocks), dim3(256), 0, st->stream, q, st->alphar, st->jac, st->pc_q, g,
                       R, dt, var);
  } else {
    const double uref = st->uref;
    hipLaunchKernelGGL(k_lhs_sponge, dim3(blocks), dim3(256), 0, st->stream, q, st->alphar, st->jac, st->met13,
                       st->met23, g, R, dt, var, (n.relaxed_u + n.relaxed_u) / uref / 2,
                       (n.relaxed_v + n.relaxed_v) / uref / 2, n.relaxed_u, n.relaxed_v, n.relaxed_w, uref);
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int compute_sponges(pcf_state* st, double dt, double time) {
  (void)time;
  size_t n = st->g.nh();
  hipLaunchKernelGGL(k_betar, dim3(2048), dim3(256), 0, st->stream, st->betar_s, st->betar, n, dt * st->tref,
                     st->tref);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int cfl_maxima(pcf_state* st, double out[3]) {
  const int nb = 512;
  hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, st->stream, st->u, st->v, st->w, st->g, st->partials);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(64), 0, st->stream, st->partials, nb, st->cfl_dev);
  PCF_HIP(hipGetLastError());
  PCF_HIP(hipMemcpyAsync(st->cfl_host, st->cfl_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  for (int d = 0; d < 3; ++d) out[d] = st->cfl_host[d];
  return PCF_OK;
}

int copy_field(pcf_state* st, double* dst, const double* src) {
  ProfScope prof(st, PCF_K_UPDATES);
  hipLaunchKernelGGL(k_copy_field, dim3(2048), dim3(256), 0, st->stream, dst, src, st->g.nh());
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

}  // namespace pcf
