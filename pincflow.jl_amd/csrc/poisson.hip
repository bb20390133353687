// Pressure solve on the device: src/PoissonSolver (compute_lhs!,
// compute_operator!, apply_operator!, apply_preconditioner!, apply_bicgstab!,
// solve_poisson!, apply_corrector!, correct!) and the global dot products of
// src/MPIOperations/compute_global_dot_product.jl.
//
// Reductions are two-stage and deterministic: each block reduces its values
// with wavefront shuffles and LDS and writes one partial per block; a
// one-block finalize kernel sums the partials in a fixed order and leaves the
// BiCGSTAB scalars (alpha, omega, beta, residual, convergence flag) in device
// memory, so an iteration needs no host round trip.
#include <cmath>
#include <cstring>

#include "pcf_device.hpp"
#include "solver_common.hpp"

namespace pcf {

namespace {

// ---------------------------------------------------------------------------
// compute_lhs! (compute_lhs!.jl:68-223)
// ---------------------------------------------------------------------------
__global__ void k_compute_lhs(Fields F, Phys ph, Geo g, double* rhs, double* partials) {
  int i = g.i0 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = g.j0 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = g.k0 + blockIdx.z;
  double val[2] = {0.0, 0.0};
  if (i <= g.i1 && j <= g.j1) {
    size_t c = g.h(i, j, k);
    const double *jac = F.jac, *pbar = F.pbar;
    double fcscal = sqrt(pbar[c] * pbar[c] / F.rhobar[c]);
    double ur = F.u[c], ul = F.u[g.h(i - 1, j, k)], vf = F.v[c], vb = F.v[g.h(i, j - 1, k)];
    double wu = F.w[c], wd = F.w[g.h(i, j, k - 1)];
    if (F.comp) {  // compute_lhs!.jl:162-223: the winds carry J P already
      double heating = conductive_heating(F, g, i, j, k) * ph.ma2 * ph.kappa;
      double bu = (ur - ul) / g.dx / jac[c] * ph.ma2 * ph.kappa;
      double bv = (vf - vb) / g.dy / jac[c] * ph.ma2 * ph.kappa;
      double bw = (wu - wd) / g.dz / jac[c] * ph.ma2 * ph.kappa;
      bu /= fcscal;
      bv /= fcscal;
      bw /= fcscal;
      heating /= fcscal;
      double bb = bu + bv + bw - heating;
      rhs[g.v(i - g.i0 + 1, j - g.j0 + 1, k - g.k0 + 1)] = bb;
      val[0] = bb * bb;
      val[1] = bu * bu + bv * bv + bw * bw + heating * heating;
    } else {
    size_t r = g.h(i + 1, j, k), l = g.h(i - 1, j, k), f = g.h(i, j + 1, k), b = g.h(i, j - 1, k);
    size_t up = g.h(i, j, k + 1), dn = g.h(i, j, k - 1);
    double pedger = 0.5 * (jac[c] * pbar[c] + jac[r] * pbar[r]);
    double pedgel = 0.5 * (jac[c] * pbar[c] + jac[l] * pbar[l]);
    double pedgef = 0.5 * (jac[c] * pbar[c] + jac[f] * pbar[f]);
    double pedgeb = 0.5 * (jac[c] * pbar[c] + jac[b] * pbar[b]);
    double pedgeu = jac[c] * jac[up] * (pbar[c] + pbar[up]) / (jac[c] + jac[up]);
    double pedged = jac[c] * jac[dn] * (pbar[c] + pbar[dn]) / (jac[c] + jac[dn]);
    double bu = (pedger * ur - pedgel * ul) / g.dx / jac[c] * ph.ma2 * ph.kappa;
    double bv = (pedgef * vf - pedgeb * vb) / g.dy / jac[c] * ph.ma2 * ph.kappa;
    double bw = (pedgeu * wu - pedged * wd) / g.dz / jac[c] * ph.ma2 * ph.kappa;
    bu /= fcscal;
    bv /= fcscal;
    bw /= fcscal;
    double bb = bu + bv + bw;
    rhs[g.v(i - g.i0 + 1, j - g.j0 + 1, k - g.k0 + 1)] = bb;
    val[0] = bb * bb;
    val[1] = bu * bu + bv * bv + bw * bw;
    }
  }
  int block = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  block_partials<2>(val, partials, block);
}

__global__ void k_finalize_lhs(const double* partials, int nparts, SolverScalars* sc, double nxyz) {
  double a = finalize_sum(partials, 0, nparts);
  double b = finalize_sum(partials, 1, nparts);
  if (threadIdx.x == 0) {
    sc->divl2 = sqrt(a / nxyz);
    sc->divl2_norm = sqrt(b / nxyz);
  }
}

// ---------------------------------------------------------------------------
// compute_operator! (compute_operator!.jl:40-1606)
// ---------------------------------------------------------------------------
// COMP: the compressible model (its buoyancy factor and centre shift); a
// template parameter so that each model's kernel carries only its own branch
template <bool COMP>
__global__ void __launch_bounds__(256) k_compute_operator(Fields F, const double* betar, const double* fcs, Geo g,
                                                          Tensor T, double dt, double rf, int damp,
                                                          const double* __restrict__ shift) {
  int i = g.i0 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = g.j0 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = g.k0 + blockIdx.z;
  if (i > g.i1 || j > g.j1) return;
  const int nbz = g.nbz, z_size = g.z_size, k0 = g.k0;
  const bool kbot = k == k0, kbot1 = k == k0 + 1;
  const bool ktop = k == z_size + nbz, ktop1 = k == z_size + nbz - 1;
  const double dx = g.dx, dy = g.dy, dz = g.dz, dt2 = dt * dt;
  const double *jac = F.jac, *pbar = F.pbar, *rhobar = F.rhobar, *rho = F.rho, *n2 = F.n2;
  const double *m13 = F.m13, *m23 = F.m23, *m33 = F.m33;
#define H(a, b, c) g.h(i + (a), j + (b), k + (c))
  const size_t C = H(0, 0, 0), R = H(1, 0, 0), L = H(-1, 0, 0), Fw = H(0, 1, 0), B = H(0, -1, 0);
  const size_t U = H(0, 0, 1), D = H(0, 0, -1);
  const size_t RU = H(1, 0, 1), RD = H(1, 0, -1), LU = H(-1, 0, 1), LD = H(-1, 0, -1);
  const size_t FU = H(0, 1, 1), FD = H(0, 1, -1), BU = H(0, -1, 1), BD = H(0, -1, -1);
  const double fcscal = fcs[C];
  const double jacinv = 1.0 / jac[C];
  const double jc = jac[C], ju = jac[U], jd = jac[D];
  // P coefficients (divergence)
  const double pedgerdiv = 0.5 * (jc * pbar[C] + jac[R] * pbar[R]);
  const double pedgeldiv = 0.5 * (jc * pbar[C] + jac[L] * pbar[L]);
  const double pedgefdiv = 0.5 * (jc * pbar[C] + jac[Fw] * pbar[Fw]);
  const double pedgebdiv = 0.5 * (jc * pbar[C] + jac[B] * pbar[B]);
  const double pedgeudiv = jc * ju * (pbar[C] + pbar[U]) / (jc + ju);
  const double pedgeddiv = jc * jd * (pbar[C] + pbar[D]) / (jc + jd);
  // P coefficients (pressure gradient)
  const double pedgergra = 0.5 * (pbar[C] + pbar[R]);
  const double pedgelgra = 0.5 * (pbar[C] + pbar[L]);
  const double pedgefgra = 0.5 * (pbar[C] + pbar[Fw]);
  const double pedgebgra = 0.5 * (pbar[C] + pbar[B]);
  const double pedgeugra = (ju * pbar[C] + jc * pbar[U]) / (jc + ju);
  const double pedgedgra = (jd * pbar[C] + jc * pbar[D]) / (jc + jd);
  const double puedgergra = 0.5 * (pbar[U] + pbar[RU]);
  const double puedgelgra = 0.5 * (pbar[U] + pbar[LU]);
  const double puedgefgra = 0.5 * (pbar[U] + pbar[FU]);
  const double puedgebgra = 0.5 * (pbar[U] + pbar[BU]);
  const double pdedgergra = 0.5 * (pbar[D] + pbar[RD]);
  const double pdedgelgra = 0.5 * (pbar[D] + pbar[LD]);
  const double pdedgefgra = 0.5 * (pbar[D] + pbar[FD]);
  const double pdedgebgra = 0.5 * (pbar[D] + pbar[BD]);
  // density coefficients
  const double rhoedger = 0.5 * (rho[C] + rho[R]) + 0.5 * (rhobar[C] + rhobar[R]);
  const double rhoedgel = 0.5 * (rho[C] + rho[L]) + 0.5 * (rhobar[C] + rhobar[L]);
  const double rhoedgef = 0.5 * (rho[C] + rho[Fw]) + 0.5 * (rhobar[C] + rhobar[Fw]);
  const double rhoedgeb = 0.5 * (rho[C] + rho[B]) + 0.5 * (rhobar[C] + rhobar[B]);
  const double rhoedgeu = (ju * rho[C] + jc * rho[U]) / (jc + ju) + (ju * rhobar[C] + jc * rhobar[U]) / (jc + ju);
  const double rhoedged = (jd * rho[C] + jc * rho[D]) / (jc + jd) + (jd * rhobar[C] + jc * rhobar[D]) / (jc + jd);
  F.comp = COMP ? 1 : 0;
  const double fwu = buoyancy_factor_w(F, g, i, j, k);
  const double fwd = buoyancy_factor_w(F, g, i, j, k - 1);
  const double rhouedger = 0.5 * (rho[U] + rho[RU] + rhobar[U] + rhobar[RU]);
  const double rhouedgel = 0.5 * (rho[U] + rho[LU] + rhobar[U] + rhobar[LU]);
  const double rhouedgef = 0.5 * (rho[U] + rho[FU] + rhobar[U] + rhobar[FU]);
  const double rhouedgeb = 0.5 * (rho[U] + rho[BU] + rhobar[U] + rhobar[BU]);
  const double rhodedger = 0.5 * (rho[D] + rho[RD] + rhobar[D] + rhobar[RD]);
  const double rhodedgel = 0.5 * (rho[D] + rho[LD] + rhobar[D] + rhobar[LD]);
  const double rhodedgef = 0.5 * (rho[D] + rho[FD] + rhobar[D] + rhobar[FD]);
  const double rhodedgeb = 0.5 * (rho[D] + rho[BD] + rhobar[D] + rhobar[BD]);
  const double n2edgeu = (ju * n2[C] + jc * n2[U]) / (jc + ju);
  const double n2edged = (jd * n2[C] + jc * n2[D]) / (jc + jd);
  // metric-tensor interpolation
  const double met13edger = 0.5 * (m13[C] + m13[R]);
  const double met13edgel = 0.5 * (m13[C] + m13[L]);
  const double met23edgef = 0.5 * (m23[C] + m23[Fw]);
  const double met23edgeb = 0.5 * (m23[C] + m23[B]);
  const double met13edgeu = (ju * m13[C] + jc * m13[U]) / (jc + ju);
  const double met23edgeu = (ju * m23[C] + jc * m23[U]) / (jc + ju);
  const double met33edgeu = (ju * m33[C] + jc * m33[U]) / (jc + ju);
  const double met13edged = (jd * m13[C] + jc * m13[D]) / (jc + jd);
  const double met23edged = (jd * m23[C] + jc * m23[D]) / (jc + jd);
  const double met33edged = (jd * m33[C] + jc * m33[D]) / (jc + jd);
  const double met13uedger = 0.5 * (m13[U] + m13[RU]);
  const double met13uedgel = 0.5 * (m13[U] + m13[LU]);
  const double met23uedgef = 0.5 * (m23[U] + m23[FU]);
  const double met23uedgeb = 0.5 * (m23[U] + m23[BU]);
  const double met13dedger = 0.5 * (m13[D] + m13[RD]);
  const double met13dedgel = 0.5 * (m13[D] + m13[LD]);
  const double met23dedgef = 0.5 * (m23[D] + m23[FD]);
  const double met23dedgeb = 0.5 * (m23[D] + m23[BD]);
  // Rayleigh damping (compute_operator!.jl:283-378)
  double facedger = 1.0, facedgel = 1.0, facedgef = 1.0, facedgeb = 1.0;
  double facuedger = 1.0, facuedgel = 1.0, facuedgef = 1.0, facuedgeb = 1.0;
  double facdedger = 1.0, facdedgel = 1.0, facdedgef = 1.0, facdedgeb = 1.0;
  double facedgeu = 1.0, facedged = 1.0;
  if (damp) {
    facedger += dt * 0.5 * (betar[C] + betar[R]) * rf;
    facedgel += dt * 0.5 * (betar[C] + betar[L]) * rf;
    facedgef += dt * 0.5 * (betar[C] + betar[Fw]) * rf;
    facedgeb += dt * 0.5 * (betar[C] + betar[B]) * rf;
    facuedger += dt * 0.5 * (betar[U] + betar[RU]) * rf;
    facuedgel += dt * 0.5 * (betar[U] + betar[LU]) * rf;
    facuedgef += dt * 0.5 * (betar[U] + betar[FU]) * rf;
    facuedgeb += dt * 0.5 * (betar[U] + betar[BU]) * rf;
    facdedger += dt * 0.5 * (betar[D] + betar[RD]) * rf;
    facdedgel += dt * 0.5 * (betar[D] + betar[LD]) * rf;
    facdedgef += dt * 0.5 * (betar[D] + betar[FD]) * rf;
    facdedgeb += dt * 0.5 * (betar[D] + betar[BD]) * rf;
  }
  facedgeu += dt * (ju * betar[C] + jc * betar[U]) / (jc + ju) * rf;
  facedged += dt * (jd * betar[C] + jc * betar[D]) / (jc + jd) * rf;
  const double imphoredger = 1.0 / facedger, imphoredgel = 1.0 / facedgel;
  const double imphoredgef = 1.0 / facedgef, imphoredgeb = 1.0 / facedgeb;
  const double imphoruedger = 1.0 / facuedger, imphoruedgel = 1.0 / facuedgel;
  const double imphoruedgef = 1.0 / facuedgef, imphoruedgeb = 1.0 / facuedgeb;
  const double imphordedger = 1.0 / facdedger, imphordedgel = 1.0 / facdedgel;
  const double imphordedgef = 1.0 / facdedgef, imphordedgeb = 1.0 / facdedgeb;
  const double impveredgeu = 1.0 / (facedgeu + fwu * n2edgeu * dt2);
  const double impveredged = 1.0 / (facedged + fwd * n2edged * dt2);
  // N^2-coupled vertical terms of the horizontal gradients
  auto Tu = [&](double met, double jfac, double imp, double rhoe) {
    return jacinv / dz * pedgeudiv * impveredgeu * fwu * n2edgeu * dt2 * 0.5 * met * jfac / (jc + ju) * imp / rhoe;
  };
  auto Td = [&](double met, double jfac, double imp, double rhoe) {
    return jacinv / dz * pedgeddiv * impveredged * fwd * n2edged * dt2 * 0.5 * met * jfac / (jc + jd) * imp / rhoe;
  };
  // gradient coefficients (compute_operator!.jl:380-719)
  double gedger, gedgel, gedgef, gedgeb;
  {
    const double hr = jacinv / dx * pedgerdiv * imphoredger / rhoedger;
    const double hl = -jacinv / dx * pedgeldiv * imphoredgel / rhoedgel;
    const double hf = jacinv / dy * pedgefdiv * imphoredgef / rhoedgef;
    const double hb = -jacinv / dy * pedgebdiv * imphoredgeb / rhoedgeb;
    if (kbot) {
      gedger = hr + Tu(m13[C], ju, imphoredger, rhoedger);
      gedgel = hl + Tu(m13[C], ju, imphoredgel, rhoedgel);
      gedgef = hf + Tu(m23[C], ju, imphoredgef, rhoedgef);
      gedgeb = hb + Tu(m23[C], ju, imphoredgeb, rhoedgeb);
    } else if (ktop) {
      gedger = hr - Td(m13[C], jd, imphoredger, rhoedger);
      gedgel = hl - Td(m13[C], jd, imphoredgel, rhoedgel);
      gedgef = hf - Td(m23[C], jd, imphoredgef, rhoedgef);
      gedgeb = hb - Td(m23[C], jd, imphoredgeb, rhoedgeb);
    } else {
      gedger = hr + Tu(m13[C], ju, imphoredger, rhoedger) - Td(m13[C], jd, imphoredger, rhoedger);
      gedgel = hl + Tu(m13[C], ju, imphoredgel, rhoedgel) - Td(m13[C], jd, imphoredgel, rhoedgel);
      gedgef = hf + Tu(m23[C], ju, imphoredgef, rhoedgef) - Td(m23[C], jd, imphoredgef, rhoedgef);
      gedgeb = hb + Tu(m23[C], ju, imphoredgeb, rhoedgeb) - Td(m23[C], jd, imphoredgeb, rhoedgeb);
    }
  }
  const double gedgeu = ktop ? 0.0 : jacinv / dz * pedgeudiv * impveredgeu / rhoedgeu;
  const double gedged = kbot ? 0.0 : -jacinv / dz * pedgeddiv * impveredged / rhoedged;
  const double guedger = ktop ? 0.0 : Tu(m13[U], jc, imphoruedger, rhouedger);
  const double guedgel = ktop ? 0.0 : Tu(m13[U], jc, imphoruedgel, rhouedgel);
  const double guedgef = ktop ? 0.0 : Tu(m23[U], jc, imphoruedgef, rhouedgef);
  const double guedgeb = ktop ? 0.0 : Tu(m23[U], jc, imphoruedgeb, rhouedgeb);
  const double gdedger = kbot ? 0.0 : -Td(m13[D], jc, imphordedger, rhodedger);
  const double gdedgel = kbot ? 0.0 : -Td(m13[D], jc, imphordedgel, rhodedgel);
  const double gdedgef = kbot ? 0.0 : -Td(m23[D], jc, imphordedgef, rhodedgef);
  const double gdedgeb = kbot ? 0.0 : -Td(m23[D], jc, imphordedgeb, rhodedgeb);

  // products G * P of compute_operator!.jl:721-1467
  const double Rr = gedger * pedgergra, Ll = gedgel * pedgelgra;
  const double Ff = gedgef * pedgefgra, Bb = gedgeb * pedgebgra;
  const double Uu = gedgeu * pedgeugra, Dd = gedged * pedgedgra;
  const double UR = guedger * puedgergra, UL = guedgel * puedgelgra;
  const double UF = guedgef * puedgefgra, UB = guedgeb * puedgebgra;
  const double DR = gdedger * pdedgergra, DL = gdedgel * pdedgelgra;
  const double DF = gdedgef * pdedgefgra, DB = gdedgeb * pdedgebgra;
  auto q4 = [&](double X, double m) { return X * 0.25 * m / dz; };
  auto q1 = [&](double X, double m) { return X * m / dz; };
  auto hx = [&](double X, double m, double d, double ja, double jb) { return X * 0.5 * m / d * ja / jb; };
  const double jr = jac[R], jru = jac[RU], jrd = jac[RD];
  const double jl = jac[L], jlu = jac[LU], jld = jac[LD];
  const double jf = jac[Fw], jfu = jac[FU], jfd = jac[FD];
  const double jb = jac[B], jbu = jac[BU], jbd = jac[BD];

  double ac, ar, al, af, ab, au, ad, aru, ard, alu, ald, afu, afd, abu, abd;
  double auu, add, aruu, ardd, aluu, aldd, afuu, afdd, abuu, abdd;
  if (kbot) {
    ac = -Rr * (1.0 / dx + 0.75 * met13edger / dz) + Ll * (1.0 / dx - 0.75 * met13edgel / dz) -
         Ff * (1.0 / dy + 0.75 * met23edgef / dz) + Bb * (1.0 / dy - 0.75 * met23edgeb / dz) - Uu * met33edgeu / dz -
         q4(UR, met13uedger) - q4(UL, met13uedgel) - q4(UF, met23uedgef) - q4(UB, met23uedgeb);
  } else if (kbot1) {
    ac = -Rr / dx + Ll / dx - Ff / dy + Bb / dy - Uu * met33edgeu / dz + Dd * met33edged / dz - q4(UR, met13uedger) -
         q4(UL, met13uedgel) - q4(UF, met23uedgef) - q4(UB, met23uedgeb) + q1(DR, met13dedger) + q1(DL, met13dedgel) +
         q1(DF, met23dedgef) + q1(DB, met23dedgeb);
  } else if (ktop1) {
    ac = -Rr / dx + Ll / dx - Ff / dy + Bb / dy - Uu * met33edgeu / dz + Dd * met33edged / dz - q1(UR, met13uedger) -
         q1(UL, met13uedgel) - q1(UF, met23uedgef) - q1(UB, met23uedgeb) + q4(DR, met13dedger) + q4(DL, met13dedgel) +
         q4(DF, met23dedgef) + q4(DB, met23dedgeb);
  } else if (ktop) {
    ac = -Rr * (1.0 / dx - 0.75 * met13edger / dz) + Ll * (1.0 / dx + 0.75 * met13edgel / dz) -
         Ff * (1.0 / dy - 0.75 * met23edgef / dz) + Bb * (1.0 / dy + 0.75 * met23edgeb / dz) + Dd * met33edged / dz +
         q4(DR, met13dedger) + q4(DL, met13dedgel) + q4(DF, met23dedgef) + q4(DB, met23dedgeb);
  } else {
    ac = -Rr / dx + Ll / dx - Ff / dy + Bb / dy - Uu * met33edgeu / dz + Dd * met33edged / dz - q4(UR, met13uedger) -
         q4(UL, met13uedgel) - q4(UF, met23uedgef) - q4(UB, met23uedgeb) + q4(DR, met13dedger) + q4(DL, met13dedgel) +
         q4(DF, met23dedgef) + q4(DB, met23dedgeb);
  }
  if (COMP) ac -= shift[g.v(i - g.i0 + 1, j - g.j0 + 1, k - g.k0 + 1)];  // compute_operator!.jl:854-860
  if (kbot) {
    ar = Rr * (1.0 / dx - 0.75 * met13edger / dz) + hx(Uu, met13edgeu, dx, jru, jr + jru) - q4(UR, met13uedger);
    al = -Ll * (1.0 / dx + 0.75 * met13edgel / dz) - hx(Uu, met13edgeu, dx, jlu, jl + jlu) - q4(UL, met13uedgel);
    af = Ff * (1.0 / dy - 0.75 * met23edgef / dz) + hx(Uu, met23edgeu, dy, jfu, jf + jfu) - q4(UF, met23uedgef);
    ab = -Bb * (1.0 / dy + 0.75 * met23edgeb / dz) - hx(Uu, met23edgeu, dy, jbu, jb + jbu) - q4(UB, met23uedgeb);
  } else if (kbot1) {
    ar = Rr / dx + hx(Uu, met13edgeu, dx, jru, jr + jru) + hx(Dd, met13edged, dx, jrd, jr + jrd) - q4(UR, met13uedger) +
         q1(DR, met13dedger);
    al = -Ll / dx - hx(Uu, met13edgeu, dx, jlu, jl + jlu) - hx(Dd, met13edged, dx, jld, jl + jld) - q4(UL, met13uedgel) +
         q1(DL, met13dedgel);
    af = Ff / dy + hx(Uu, met23edgeu, dy, jfu, jf + jfu) + hx(Dd, met23edged, dy, jfd, jf + jfd) - q4(UF, met23uedgef) +
         q1(DF, met23dedgef);
    ab = -Bb / dy - hx(Uu, met23edgeu, dy, jbu, jb + jbu) - hx(Dd, met23edged, dy, jbd, jb + jbd) - q4(UB, met23uedgeb) +
         q1(DB, met23dedgeb);
  } else if (ktop1) {
    ar = Rr / dx + hx(Uu, met13edgeu, dx, jru, jr + jru) + hx(Dd, met13edged, dx, jrd, jr + jrd) - q1(UR, met13uedger) +
         q4(DR, met13dedger);
    al = -Ll / dx - hx(Uu, met13edgeu, dx, jlu, jl + jlu) - hx(Dd, met13edged, dx, jld, jl + jld) - q1(UL, met13uedgel) +
         q4(DL, met13dedgel);
    af = Ff / dy + hx(Uu, met23edgeu, dy, jfu, jf + jfu) + hx(Dd, met23edged, dy, jfd, jf + jfd) - q1(UF, met23uedgef) +
         q4(DF, met23dedgef);
    ab = -Bb / dy - hx(Uu, met23edgeu, dy, jbu, jb + jbu) - hx(Dd, met23edged, dy, jbd, jb + jbd) - q1(UB, met23uedgeb) +
         q4(DB, met23dedgeb);
  } else if (ktop) {
    ar = Rr * (1.0 / dx + 0.75 * met13edger / dz) + hx(Dd, met13edged, dx, jrd, jr + jrd) + q4(DR, met13dedger);
    al = -Ll * (1.0 / dx - 0.75 * met13edgel / dz) - hx(Dd, met13edged, dx, jld, jl + jld) + q4(DL, met13dedgel);
    af = Ff * (1.0 / dy + 0.75 * met23edgef / dz) + hx(Dd, met23edged, dy, jfd, jf + jfd) + q4(DF, met23dedgef);
    ab = -Bb * (1.0 / dy - 0.75 * met23edgeb / dz) - hx(Dd, met23edged, dy, jbd, jb + jbd) + q4(DB, met23dedgeb);
  } else {
    ar = Rr / dx + hx(Uu, met13edgeu, dx, jru, jr + jru) + hx(Dd, met13edged, dx, jrd, jr + jrd) - q4(UR, met13uedger) +
         q4(DR, met13dedger);
    al = -Ll / dx - hx(Uu, met13edgeu, dx, jlu, jl + jlu) - hx(Dd, met13edged, dx, jld, jl + jld) - q4(UL, met13uedgel) +
         q4(DL, met13dedgel);
    af = Ff / dy + hx(Uu, met23edgeu, dy, jfu, jf + jfu) + hx(Dd, met23edged, dy, jfd, jf + jfd) - q4(UF, met23uedgef) +
         q4(DF, met23dedgef);
    ab = -Bb / dy - hx(Uu, met23edgeu, dy, jbu, jb + jbu) - hx(Dd, met23edged, dy, jbd, jb + jbd) - q4(UB, met23uedgeb) +
         q4(DB, met23dedgeb);
  }
  // A(i,j,k+1)
  if (kbot) {
    au = q1(Rr, met13edger) + q1(Ll, met13edgel) + q1(Ff, met23edgef) + q1(Bb, met23edgeb) + Uu * met33edgeu / dz -
         UR / dx + UL / dx - UF / dy + UB / dy;
  } else if (kbot1) {
    au = q4(Rr, met13edger) + q4(Ll, met13edgel) + q4(Ff, met23edgef) + q4(Bb, met23edgeb) + Uu * met33edgeu / dz -
         UR / dx + UL / dx - UF / dy + UB / dy - q4(DR, met13dedger) - q4(DL, met13dedgel) - q4(DF, met23dedgef) -
         q4(DB, met23dedgeb);
  } else if (ktop1) {
    au = q4(Rr, met13edger) + q4(Ll, met13edgel) + q4(Ff, met23edgef) + q4(Bb, met23edgeb) + Uu * met33edgeu / dz -
         UR * (1.0 / dx - 0.75 * met13uedger / dz) + UL * (1.0 / dx + 0.75 * met13uedgel / dz) -
         UF * (1.0 / dy - 0.75 * met23uedgef / dz) + UB * (1.0 / dy + 0.75 * met23uedgeb / dz);
  } else if (ktop) {
    au = 0.0;
  } else {
    au = q4(Rr, met13edger) + q4(Ll, met13edgel) + q4(Ff, met23edgef) + q4(Bb, met23edgeb) + Uu * met33edgeu / dz -
         UR / dx + UL / dx - UF / dy + UB / dy;
  }
  // A(i,j,k-1)
  if (kbot) {
    ad = 0.0;
  } else if (kbot1) {
    ad = -q4(Rr, met13edger) - q4(Ll, met13edgel) - q4(Ff, met23edgef) - q4(Bb, met23edgeb) - Dd * met33edged / dz -
         DR * (1.0 / dx + 0.75 * met13dedger / dz) + DL * (1.0 / dx - 0.75 * met13dedgel / dz) -
         DF * (1.0 / dy + 0.75 * met23dedgef / dz) + DB * (1.0 / dy - 0.75 * met23dedgeb / dz);
  } else if (ktop1) {
    ad = -q4(Rr, met13edger) - q4(Ll, met13edgel) - q4(Ff, met23edgef) - q4(Bb, met23edgeb) - Dd * met33edged / dz -
         DR / dx + DL / dx - DF / dy + DB / dy + q4(UR, met13uedger) + q4(UL, met13uedgel) + q4(UF, met23uedgef) +
         q4(UB, met23uedgeb);
  } else if (ktop) {
    ad = -q1(Rr, met13edger) - q1(Ll, met13edgel) - q1(Ff, met23edgef) - q1(Bb, met23edgeb) - Dd * met33edged / dz -
         DR / dx + DL / dx - DF / dy + DB / dy;
  } else {
    ad = -q4(Rr, met13edger) - q4(Ll, met13edgel) - q4(Ff, met23edgef) - q4(Bb, met23edgeb) - Dd * met33edged / dz -
         DR / dx + DL / dx - DF / dy + DB / dy;
  }
  // diagonal neighbours A(i+-1, j, k+-1), A(i, j+-1, k+-1)
  if (kbot) {
    aru = q1(Rr, met13edger) + hx(Uu, met13edgeu, dx, jr, jr + jru) + UR / dx;
    ard = 0.0;
    alu = q1(Ll, met13edgel) - hx(Uu, met13edgeu, dx, jl, jl + jlu) - UL / dx;
    ald = 0.0;
    afu = q1(Ff, met23edgef) + hx(Uu, met23edgeu, dy, jf, jf + jfu) + UF / dy;
    afd = 0.0;
    abu = q1(Bb, met23edgeb) - hx(Uu, met23edgeu, dy, jb, jb + jbu) - UB / dy;
    abd = 0.0;
  } else if (kbot1) {
    aru = q4(Rr, met13edger) + hx(Uu, met13edgeu, dx, jr, jr + jru) + UR / dx - q4(DR, met13dedger);
    ard = -q4(Rr, met13edger) + hx(Dd, met13edged, dx, jr, jr + jrd) + DR * (1.0 / dx - 0.75 * met13dedger / dz);
    alu = q4(Ll, met13edgel) - hx(Uu, met13edgeu, dx, jl, jl + jlu) - UL / dx - q4(DL, met13dedgel);
    ald = -q4(Ll, met13edgel) - hx(Dd, met13edged, dx, jl, jl + jld) - DL * (1.0 / dx + 0.75 * met13dedgel / dz);
    afu = q4(Ff, met23edgef) + hx(Uu, met23edgeu, dy, jf, jf + jfu) + UF / dy - q4(DF, met23dedgef);
    afd = -q4(Ff, met23edgef) + hx(Dd, met23edged, dy, jf, jf + jfd) + DF * (1.0 / dy - 0.75 * met23dedgef / dz);
    abu = q4(Bb, met23edgeb) - hx(Uu, met23edgeu, dy, jb, jb + jbu) - UB / dy - q4(DB, met23dedgeb);
    abd = -q4(Bb, met23edgeb) - hx(Dd, met23edged, dy, jb, jb + jbd) - DB * (1.0 / dy + 0.75 * met23dedgeb / dz);
  } else if (ktop1) {
    aru = q4(Rr, met13edger) + hx(Uu, met13edgeu, dx, jr, jr + jru) + UR * (1.0 / dx + 0.75 * met13uedger / dz);
    ard = -q4(Rr, met13edger) + hx(Dd, met13edged, dx, jr, jr + jrd) + DR / dx + q4(UR, met13uedger);
    alu = q4(Ll, met13edgel) - hx(Uu, met13edgeu, dx, jl, jl + jlu) - UL * (1.0 / dx - 0.75 * met13uedgel / dz);
    ald = -q4(Ll, met13edgel) - hx(Dd, met13edged, dx, jl, jl + jld) - DL / dx + q4(UL, met13uedgel);
    afu = q4(Ff, met23edgef) + hx(Uu, met23edgeu, dy, jf, jf + jfu) + UF * (1.0 / dy + 0.75 * met23uedgef / dz);
    afd = -q4(Ff, met23edgef) + hx(Dd, met23edged, dy, jf, jf + jfd) + DF / dy + q4(UF, met23uedgef);
    abu = q4(Bb, met23edgeb) - hx(Uu, met23edgeu, dy, jb, jb + jbu) - UB * (1.0 / dy - 0.75 * met23uedgeb / dz);
    abd = -q4(Bb, met23edgeb) - hx(Dd, met23edged, dy, jb, jb + jbd) - DB / dy + q4(UB, met23uedgeb);
  } else if (ktop) {
    aru = 0.0;
    ard = -q1(Rr, met13edger) + hx(Dd, met13edged, dx, jr, jr + jrd) + DR / dx;
    alu = 0.0;
    ald = -q1(Ll, met13edgel) - hx(Dd, met13edged, dx, jl, jl + jld) - DL / dx;
    afu = 0.0;
    afd = -q1(Ff, met23edgef) + hx(Dd, met23edged, dy, jf, jf + jfd) + DF / dy;
    abu = 0.0;
    abd = -q1(Bb, met23edgeb) - hx(Dd, met23edged, dy, jb, jb + jbd) - DB / dy;
  } else {
    aru = q4(Rr, met13edger) + hx(Uu, met13edgeu, dx, jr, jr + jru) + UR / dx;
    ard = -q4(Rr, met13edger) + hx(Dd, met13edged, dx, jr, jr + jrd) + DR / dx;
    alu = q4(Ll, met13edgel) - hx(Uu, met13edgeu, dx, jl, jl + jlu) - UL / dx;
    ald = -q4(Ll, met13edgel) - hx(Dd, met13edged, dx, jl, jl + jld) - DL / dx;
    afu = q4(Ff, met23edgef) + hx(Uu, met23edgeu, dy, jf, jf + jfu) + UF / dy;
    afd = -q4(Ff, met23edgef) + hx(Dd, met23edged, dy, jf, jf + jfd) + DF / dy;
    abu = q4(Bb, met23edgeb) - hx(Uu, met23edgeu, dy, jb, jb + jbu) - UB / dy;
    abd = -q4(Bb, met23edgeb) - hx(Dd, met23edged, dy, jb, jb + jbd) - DB / dy;
  }
  // two levels up and down
  const bool top2 = k >= z_size + nbz - 1, bot2 = k <= k0 + 1;
  if (kbot) {
    auu = -q4(Rr, met13edger) - q4(Ll, met13edgel) - q4(Ff, met23edgef) - q4(Bb, met23edgeb) + q4(UR, met13uedger) +
          q4(UL, met13uedgel) + q4(UF, met23uedgef) + q4(UB, met23uedgeb);
    aruu = -q4(Rr, met13edger) + q4(UR, met13uedger);
    aluu = -q4(Ll, met13edgel) + q4(UL, met13uedgel);
    afuu = -q4(Ff, met23edgef) + q4(UF, met23uedgef);
    abuu = -q4(Bb, met23edgeb) + q4(UB, met23uedgeb);
  } else if (top2) {
    auu = 0.0; aruu = 0.0; aluu = 0.0; afuu = 0.0; abuu = 0.0;
  } else {
    auu = q4(UR, met13uedger) + q4(UL, met13uedgel) + q4(UF, met23uedgef) + q4(UB, met23uedgeb);
    aruu = q4(UR, met13uedger);
    aluu = q4(UL, met13uedgel);
    afuu = q4(UF, met23uedgef);
    abuu = q4(UB, met23uedgeb);
  }
  if (bot2) {
    add = 0.0; ardd = 0.0; aldd = 0.0; afdd = 0.0; abdd = 0.0;
  } else if (ktop) {
    add = q4(Rr, met13edger) + q4(Ll, met13edgel) + q4(Ff, met23edgef) + q4(Bb, met23edgeb) - q4(DR, met13dedger) -
          q4(DL, met13dedgel) - q4(DF, met23dedgef) - q4(DB, met23dedgeb);
    ardd = q4(Rr, met13edger) - q4(DR, met13dedger);
    aldd = q4(Ll, met13edgel) - q4(DL, met13dedgel);
    afdd = q4(Ff, met23edgef) - q4(DF, met23dedgef);
    abdd = q4(Bb, met23edgeb) - q4(DB, met23dedgeb);
  } else {
    add = -q4(DR, met13dedger) - q4(DL, met13dedgel) - q4(DF, met23dedgef) - q4(DB, met23dedgeb);
    ardd = -q4(DR, met13dedger);
    aldd = -q4(DL, met13dedgel);
    afdd = -q4(DF, met23dedgef);
    abdd = -q4(DB, met23dedgeb);
  }
  // scaling (compute_operator!.jl:1545-1570) and store
  const size_t o = g.v(i - g.i0 + 1, j - g.j0 + 1, k - g.k0 + 1);
  T.a[T_AC][o] = ac / (fcscal * fcscal);
  T.a[T_AR][o] = ar / fcscal / fcs[R];
  T.a[T_AL][o] = al / fcscal / fcs[L];
  T.a[T_AF][o] = af / fcscal / fcs[Fw];
  T.a[T_AB][o] = ab / fcscal / fcs[B];
  T.a[T_AU][o] = au / fcscal / fcs[U];
  T.a[T_AD][o] = ad / fcscal / fcs[D];
  T.a[T_ARU][o] = aru / fcscal / fcs[RU];
  T.a[T_ARD][o] = ard / fcscal / fcs[RD];
  T.a[T_ALU][o] = alu / fcscal / fcs[LU];
  T.a[T_ALD][o] = ald / fcscal / fcs[LD];
  T.a[T_AFU][o] = afu / fcscal / fcs[FU];
  T.a[T_AFD][o] = afd / fcscal / fcs[FD];
  T.a[T_ABU][o] = abu / fcscal / fcs[BU];
  T.a[T_ABD][o] = abd / fcscal / fcs[BD];
  T.a[T_AUU][o] = auu / fcscal / fcs[H(0, 0, 2)];
  T.a[T_ADD][o] = add / fcscal / fcs[H(0, 0, -2)];
  T.a[T_ARUU][o] = aruu / fcscal / fcs[H(1, 0, 2)];
  T.a[T_ARDD][o] = ardd / fcscal / fcs[H(1, 0, -2)];
  T.a[T_ALUU][o] = aluu / fcscal / fcs[H(-1, 0, 2)];
  T.a[T_ALDD][o] = aldd / fcscal / fcs[H(-1, 0, -2)];
  T.a[T_AFUU][o] = afuu / fcscal / fcs[H(0, 1, 2)];
  T.a[T_AFDD][o] = afdd / fcscal / fcs[H(0, 1, -2)];
  T.a[T_ABUU][o] = abuu / fcscal / fcs[H(0, -1, 2)];
  T.a[T_ABDD][o] = abdd / fcscal / fcs[H(0, -1, -2)];
#undef H
}

// ---------------------------------------------------------------------------
// apply_operator! (apply_operator!.jl:48-479) on interior vectors; the
// neighbour of a boundary cell is the periodic image in x and y and zero
// beyond the walls (the reference's halo of `s` holds exactly that).
// ---------------------------------------------------------------------------
struct ApplyArgs {
  Tensor T;
  const double* s;
  double* ls;
  int nx, ny, nz;
};

template <bool TOTAL, int NDOT>
__global__ void __launch_bounds__(256) k_apply(ApplyArgs A, const double* d1, const double* d2, double* partials,
                                               const SolverScalars* sc) {
  if (sc && sc->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based interior
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  double val[NDOT > 0 ? NDOT : 1];
#pragma unroll
  for (int d = 0; d < (NDOT > 0 ? NDOT : 1); ++d) val[d] = 0.0;
  if (i < nx && j < ny) {
    const size_t sxy = (size_t)nx * ny;
    const size_t o = (size_t)i + (size_t)nx * j + sxy * k;
    const int ip = i + 1 == nx ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
    const int jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
    const double* s = A.s;
    auto S = [&](int ii, int jj, int kk) -> double {
      return (kk < 0 || kk >= nz) ? 0.0 : s[(size_t)ii + (size_t)nx * jj + sxy * kk];
    };
    const double* const* T = A.T.a;
    double v;
    if (TOTAL) {
      v = T[T_AL][o] * S(im, j, k) + T[T_AR][o] * S(ip, j, k) + T[T_AF][o] * S(i, jp, k) +
          T[T_AB][o] * S(i, jm, k) + T[T_AU][o] * S(i, j, k + 1) + T[T_AD][o] * S(i, j, k - 1) +
          T[T_AC][o] * s[o];
    } else {
      v = T[T_AL][o] * S(im, j, k) + T[T_AR][o] * S(ip, j, k) + T[T_AF][o] * S(i, jp, k) + T[T_AB][o] * S(i, jm, k);
    }
    v += T[T_ARU][o] * S(ip, j, k + 1) + T[T_ARD][o] * S(ip, j, k - 1) + T[T_ALU][o] * S(im, j, k + 1) +
         T[T_ALD][o] * S(im, j, k - 1) + T[T_AFU][o] * S(i, jp, k + 1) + T[T_AFD][o] * S(i, jp, k - 1) +
         T[T_ABU][o] * S(i, jm, k + 1) + T[T_ABD][o] * S(i, jm, k - 1) + T[T_AUU][o] * S(i, j, k + 2) +
         T[T_ADD][o] * S(i, j, k - 2) + T[T_ARUU][o] * S(ip, j, k + 2) + T[T_ARDD][o] * S(ip, j, k - 2) +
         T[T_ALUU][o] * S(im, j, k + 2) + T[T_ALDD][o] * S(im, j, k - 2) + T[T_AFUU][o] * S(i, jp, k + 2) +
         T[T_AFDD][o] * S(i, jp, k - 2) + T[T_ABUU][o] * S(i, jm, k + 2) + T[T_ABDD][o] * S(i, jm, k - 2);
    A.ls[o] = v;
    if (NDOT == 1) {
      val[0] = v * d1[o];  // <v, r0>
    } else if (NDOT == 2) {
      val[0] = v * d1[o];  // <t, s>
      val[1] = v * v;      // <t, t>
    }
  }
  if (NDOT > 0) {
    int block = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_partials<(NDOT > 0 ? NDOT : 1)>(val, partials, block);
  }
  (void)d2;
}

// ---------------------------------------------------------------------------
// apply_preconditioner! (apply_preconditioner!.jl:40-124)
// ---------------------------------------------------------------------------
// Factors of the line solve in z of (1 - deta L_v) s = rhs
// (apply_preconditioner!.jl:66-111), which depend only on the operator:
// P = the divisor 1 - deta ac on the lowest level and the reciprocal pivot
// 1 / (1 - deta ac - deta ad q) above it, E = deta ad, Q = deta au p.  A sweep
// with them rounds exactly like the reference's loop, which evaluates
// deta * ad * s as (deta * ad) * s.  One column per thread.
__global__ void __launch_bounds__(64) k_pc_factors(const double* ac, const double* ad, const double* au, double* P,
                                                   double* E, double* Q, int nx, int ny, int nz, double deta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const size_t sxy = (size_t)nx * ny;
  size_t o = (size_t)i + (size_t)nx * j;
  const double denom = 1 - deta * ac[o];
  double qk = deta * au[o] / denom;
  P[o] = denom;
  E[o] = deta * ad[o];
  Q[o] = qk;
  for (int k = 1; k < nz; ++k) {
    o += sxy;
    const double e = deta * ad[o];
    const double p = 1 / (1 - deta * ac[o] - e * qk);
    qk = deta * au[o] * p;
    P[o] = p;
    E[o] = e;
    Q[o] = qk;
  }
}

// The line solve with precomputed factors, one column per thread and 64
// adjacent columns per wavefront (blockDim.x must be 64); the loads of 16
// levels are issued before their (sequential) recurrence steps, and a
// compiler barrier keeps them there.  The upward sweep's values of the
// lowest NLDS levels stay in LDS; those of higher levels go through `out`,
// which the downward sweep overwrites with the solution.  NLDS is chosen per
// launch (launch_thomas): 80 levels (40 KiB per one-wavefront block) while
// the grid has at most four blocks per CU, as a 256 x 256 grid has, so that
// four blocks fill the CU's 160 KiB; 64 levels (32 KiB, five blocks per CU)
// on larger grids, which have blocks enough to use the fifth slot.
// FIRST: rhs = 0 + deta (0 - src), the first pseudo-time step from s = 0,
// whose horizontal apply is zero.
// The body takes the block's (x, y) index, so that the ensemble kernel can
// run it for a block of any member.
template <bool FIRST, int NLDS>
__device__ __forceinline__ void pc_thomas(const double* P, const double* E, const double* Q, const double* src,
                                          double* out, int nx, int ny, int nz, double deta, int bx, int by) {
  constexpr int kThomasLds = NLDS;
  __shared__ double up[kThomasLds][64];
  constexpr int U = 16;
  const int tid = threadIdx.x;
  const int i = bx * 64 + tid;
  const int j = by;
  if (i >= nx) return;
  const size_t sxy = (size_t)nx * ny;
  const size_t c = (size_t)i + (size_t)nx * j;
  auto rhs_of = [&](double v) -> double { return FIRST ? 0.0 + deta * (0.0 - v) : v; };
  auto keep = [&](int k, double v) {
    if (k < kThomasLds)
      up[k][tid] = v;
    else
      out[c + sxy * k] = v;
  };
  auto kept = [&](int k) -> double { return k < kThomasLds ? up[k][tid] : out[c + sxy * k]; };
  double sk = rhs_of(src[c]) / P[c];
  keep(0, sk);
  // batches of U levels, the last one partial (its loads and steps
  // predicated on the level)
  for (int k = 1; k < nz; k += U) {
    double r[U], e[U], p[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = e[u] = p[u] = 0.0;
      if (k + u < nz) {
        const size_t o = c + sxy * (k + u);
        r[u] = src[o];
        e[u] = E[o];
        p[u] = P[o];
      }
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k + u < nz) {
        sk = (rhs_of(r[u]) + e[u] * sk) * p[u];
        keep(k + u, sk);
      }
    }
  }
  // the top level's upward value is its solution
  out[c + sxy * (nz - 1)] = sk;
  // downward sweep
  for (int k = nz - 2; k >= 0; k -= U) {
    double r[U], q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = q[u] = 0.0;
      if (k - u >= 0) {
        r[u] = kept(k - u);
        q[u] = Q[c + sxy * (k - u)];
      }
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k - u >= 0) {
        sk = r[u] + q[u] * sk;
        out[c + sxy * (k - u)] = sk;
      }
    }
  }
}
template <bool FIRST, int NLDS>
__global__ void __launch_bounds__(64) k_pc_thomas(const double* P, const double* E, const double* Q, const double* src,
                                                  double* out, int nx, int ny, int nz, double deta,
                                                  const SolverScalars* sc) {
  if (sc && sc->done) return;
  pc_thomas<FIRST, NLDS>(P, E, Q, src, out, nx, ny, nz, deta, blockIdx.x, blockIdx.y);
}

// rhs of a later pseudo-time step: s + deta (L_h s - sin) (apply_preconditioner!.jl:62-63)
__global__ void __launch_bounds__(256) k_pc_rhs(ApplyArgs A, const double* sin, double deta, const SolverScalars* sc) {
  if (sc && sc->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (i >= nx || j >= ny) return;
  const size_t sxy = (size_t)nx * ny;
  const size_t o = (size_t)i + (size_t)nx * j + sxy * k;
  const int ip = i + 1 == nx ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
  const int jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
  const double* s = A.s;
  auto S = [&](int ii, int jj, int kk) -> double {
    return (kk < 0 || kk >= nz) ? 0.0 : s[(size_t)ii + (size_t)nx * jj + sxy * kk];
  };
  const double* const* T = A.T.a;
  double v = T[T_AL][o] * S(im, j, k) + T[T_AR][o] * S(ip, j, k) + T[T_AF][o] * S(i, jp, k) + T[T_AB][o] * S(i, jm, k);
  v += T[T_ARU][o] * S(ip, j, k + 1) + T[T_ARD][o] * S(ip, j, k - 1) + T[T_ALU][o] * S(im, j, k + 1) +
       T[T_ALD][o] * S(im, j, k - 1) + T[T_AFU][o] * S(i, jp, k + 1) + T[T_AFD][o] * S(i, jp, k - 1) +
       T[T_ABU][o] * S(i, jm, k + 1) + T[T_ABD][o] * S(i, jm, k - 1) + T[T_AUU][o] * S(i, j, k + 2) +
       T[T_ADD][o] * S(i, j, k - 2) + T[T_ARUU][o] * S(ip, j, k + 2) + T[T_ARDD][o] * S(ip, j, k - 2) +
       T[T_ALUU][o] * S(im, j, k + 2) + T[T_ALDD][o] * S(im, j, k - 2) + T[T_AFUU][o] * S(i, jp, k + 2) +
       T[T_AFDD][o] * S(i, jp, k - 2) + T[T_ABUU][o] * S(i, jm, k + 2) + T[T_ABDD][o] * S(i, jm, k - 2);
  A.ls[o] = s[o] + deta * (v - sin[o]);
}

// ---------------------------------------------------------------------------
// The same stencils for an even nx, two x-adjacent cells per thread with
// 16-byte loads and stores: one wavefront covers 128 consecutive cells of a
// row, which halves the memory instructions per cell.  The coefficients are
// streamed once per launch and are loaded non-temporally, so they do not
// evict the neighbourhood of `s` from the caches.  Each cell's sum is
// evaluated in the order of the scalar kernels above.
//   MODE_TOTAL       ls = A s                          (apply_operator!, Total)
//   MODE_HORIZONTAL  ls = H s                          (apply_operator!, Horizontal)
//   MODE_PC_RHS      ls = s + deta (H s - sin)         (apply_preconditioner!.jl:62-63)
// NDOT = 1: partials of <ls, d1>; NDOT = 2: of <ls, d1> and <ls, ls>.
// Block 32 x 8 threads = 64 x 8 cells of one z-level.
// ---------------------------------------------------------------------------
enum { MODE_TOTAL = 0, MODE_HORIZONTAL = 1, MODE_PC_RHS = 2 };
typedef double dbl2 __attribute__((ext_vector_type(2)));
constexpr int kPairBlockY = 8;

// The body takes the block's index (bx, by, bz) and the index of its partial
// sums, so that the ensemble kernel can run it for a block of any member.
template <int MODE, int NDOT>
__device__ __forceinline__ void stencil_pair(const ApplyArgs& A, const double* __restrict__ d1,
                                             const double* __restrict__ sin, double deta, double* partials, int bx,
                                             int by, int bz, int block) {
  const int i0 = (bx * 32 + threadIdx.x) * 2;
  const int j = by * kPairBlockY + threadIdx.y;
  const int k = bz;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  double val[NDOT > 0 ? NDOT : 1];
#pragma unroll
  for (int d = 0; d < (NDOT > 0 ? NDOT : 1); ++d) val[d] = 0.0;
  if (i0 < nx && j < ny) {
    const size_t sxy = (size_t)nx * ny;
    const size_t o = (size_t)i0 + (size_t)nx * j + sxy * k;
    const int im = i0 == 0 ? nx - 1 : i0 - 1, ip = i0 + 2 == nx ? 0 : i0 + 2;
    const int jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
    const double* s = A.s;
    // s on levels k-2..k+2 (zero beyond the walls): own pair, the pairs of
    // the y neighbours, and the x neighbours left of i0 and right of i0+1
    dbl2 C[5], F[5], B[5];
    double L[5], R[5];
#pragma unroll
    for (int w = 0; w < 5; ++w) {
      const int kk = k - 2 + w;
      if (kk < 0 || kk >= nz) {
        C[w] = 0.0;
        F[w] = 0.0;
        B[w] = 0.0;
        L[w] = 0.0;
        R[w] = 0.0;
        continue;
      }
      const size_t pl = sxy * kk;
      C[w] = *(const dbl2*)(s + pl + (size_t)nx * j + i0);
      F[w] = *(const dbl2*)(s + pl + (size_t)nx * jp + i0);
      B[w] = *(const dbl2*)(s + pl + (size_t)nx * jm + i0);
      L[w] = s[pl + (size_t)nx * j + im];
      R[w] = s[pl + (size_t)nx * j + ip];
    }
    auto T = [&](int t) -> dbl2 { return __builtin_nontemporal_load((const dbl2*)(A.T.a[t] + o)); };
    const dbl2 al = T(T_AL), ar = T(T_AR), af = T(T_AF), ab = T(T_AB);
    dbl2 au = 0.0, ad = 0.0, ac = 0.0;
    if (MODE == MODE_TOTAL) {
      au = T(T_AU);
      ad = T(T_AD);
      ac = T(T_AC);
    }
    const dbl2 aru = T(T_ARU), ard = T(T_ARD), alu = T(T_ALU), ald = T(T_ALD), afu = T(T_AFU), afd = T(T_AFD);
    const dbl2 abu = T(T_ABU), abd = T(T_ABD), auu = T(T_AUU), add = T(T_ADD), aruu = T(T_ARUU), ardd = T(T_ARDD);
    const dbl2 aluu = T(T_ALUU), aldd = T(T_ALDD), afuu = T(T_AFUU), afdd = T(T_AFDD), abuu = T(T_ABUU),
               abdd = T(T_ABDD);
    dbl2 sn = 0.0;
    if (MODE == MODE_PC_RHS) sn = *(const dbl2*)(sin + o);
    double out[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      auto g = [&](const dbl2& a) -> double { return c == 0 ? a.x : a.y; };
      auto XL = [&](int w) -> double { return c == 0 ? L[w] : C[w].x; };
      auto XR = [&](int w) -> double { return c == 0 ? C[w].y : R[w]; };
      auto X0 = [&](int w) -> double { return g(C[w]); };
      auto YF = [&](int w) -> double { return g(F[w]); };
      auto YB = [&](int w) -> double { return g(B[w]); };
      double v;
      if (MODE == MODE_TOTAL) {
        v = g(al) * XL(2) + g(ar) * XR(2) + g(af) * YF(2) + g(ab) * YB(2) + g(au) * X0(3) + g(ad) * X0(1) +
            g(ac) * X0(2);
      } else {
        v = g(al) * XL(2) + g(ar) * XR(2) + g(af) * YF(2) + g(ab) * YB(2);
      }
      v += g(aru) * XR(3) + g(ard) * XR(1) + g(alu) * XL(3) + g(ald) * XL(1) + g(afu) * YF(3) + g(afd) * YF(1) +
           g(abu) * YB(3) + g(abd) * YB(1) + g(auu) * X0(4) + g(add) * X0(0) + g(aruu) * XR(4) + g(ardd) * XR(0) +
           g(aluu) * XL(4) + g(aldd) * XL(0) + g(afuu) * YF(4) + g(afdd) * YF(0) + g(abuu) * YB(4) + g(abdd) * YB(0);
      if (MODE == MODE_PC_RHS) v = X0(2) + deta * (v - g(sn));
      out[c] = v;
    }
    dbl2 res;
    res.x = out[0];
    res.y = out[1];
    __builtin_nontemporal_store(res, (dbl2*)(A.ls + o));
    if constexpr (NDOT >= 1) {
      const dbl2 dd = *(const dbl2*)(d1 + o);
      val[0] = out[0] * dd.x;
      val[0] += out[1] * dd.y;
    }
    if constexpr (NDOT == 2) {
      val[1] = out[0] * out[0];
      val[1] += out[1] * out[1];
    }
  }
  if (NDOT > 0) block_partials<(NDOT > 0 ? NDOT : 1)>(val, partials, block);
}
template <int MODE, int NDOT>
__global__ void __launch_bounds__(256) k_stencil_pair(ApplyArgs A, const double* __restrict__ d1,
                                                      const double* __restrict__ sin, double deta, double* partials,
                                                      const SolverScalars* sc) {
  if (sc && sc->done) return;
  stencil_pair<MODE, NDOT>(A, d1, sin, deta, partials, blockIdx.x, blockIdx.y, blockIdx.z,
                           blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// ---------------------------------------------------------------------------
// BiCGSTAB vector kernels (apply_bicgstab!.jl:24-150)
// ---------------------------------------------------------------------------
constexpr int kVecGrid = 2048;

// r0 = b - A 0, p = r = r0, partial <r, r>
__global__ void k_bicg_init(const double* b, double* x, double* r0, double* r, double* p, size_t n,
                            double* partials) {
  double val[1] = {0.0};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    x[q] = 0.0;
    double v = b[q] - 0.0;
    r0[q] = v;
    r[q] = v;
    p[q] = v;
    val[0] += v * v;
  }
  block_partials<1>(val, partials, blockIdx.x);
}

// x = c g (initial guess)
__global__ void k_bicg_guess(const double* g, double* x, size_t n, double c) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    x[q] = c * g[q];
}

// r0 = b - A x0 (ax = A x0), p = r = r0, partials <r, r> and <b, b>
__global__ void k_bicg_init_guessed(const double* b, const double* ax, double* r0, double* r, double* p, size_t n,
                                    double* partials) {
  double val[2] = {0.0, 0.0};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const double bq = b[q];
    const double v = bq - ax[q];
    r0[q] = v;
    r[q] = v;
    p[q] = v;
    val[0] += v * v;
    val[1] += bq * bq;
  }
  block_partials<2>(val, partials, blockIdx.x);
}

// The vector kernels below are grid-stride loops over kVecGrid blocks of 256
// threads; their bodies take the block index, so that the ensemble kernels
// run the same blocks for every member (and so give the same partial sums).

// s = r - alpha v
__device__ __forceinline__ void bicg_s(const double* r, const double* v, double* s, size_t n,
                                       const SolverScalars* sc, int b) {
  const double alpha = sc->alpha;
  for (size_t q = b * (size_t)kReduceBlock + threadIdx.x; q < n; q += (size_t)kVecGrid * kReduceBlock)
    s[q] = r[q] - alpha * v[q];
}
__global__ void k_bicg_s(const double* r, const double* v, double* s, size_t n, const SolverScalars* sc) {
  if (sc->done) return;
  bicg_s(r, v, s, n, sc, blockIdx.x);
}

// x += alpha p + omega s; r = s - omega t; partials <r, r>, <r, r0>
__device__ __forceinline__ void bicg_xr(double* x, const double* p, const double* s, const double* t, double* r,
                                        const double* r0, size_t n, double* partials, const SolverScalars* sc,
                                        int b) {
  const double alpha = sc->alpha, omega = sc->omega;
  double val[2] = {0.0, 0.0};
  for (size_t q = b * (size_t)kReduceBlock + threadIdx.x; q < n; q += (size_t)kVecGrid * kReduceBlock) {
    x[q] += alpha * p[q] + omega * s[q];
    double rn = s[q] - omega * t[q];
    r[q] = rn;
    val[0] += rn * rn;
    val[1] += rn * r0[q];
  }
  block_partials<2>(val, partials, b);
}
__global__ void k_bicg_xr(double* x, const double* p, const double* s, const double* t, double* r, const double* r0,
                          size_t n, double* partials, const SolverScalars* sc) {
  if (sc->done) return;
  bicg_xr(x, p, s, t, r, r0, n, partials, sc, blockIdx.x);
}

// The update of the textbook right-preconditioned form (van der Vorst 1992):
// x += alpha y + omega z with y = M^-1 p, z = M^-1 s; r = s - omega t;
// partials <r, r>, <r, r0>.  It equals the reference's form (x += alpha p +
// omega s, then x := M^-1 x at the end) for a linear M, and stays consistent
// with the residual recurrence when M is evaluated in FP32.
__device__ __forceinline__ void bicg_xr_yz(double* x, const double* y, const double* z, const double* s,
                                           const double* t, double* r, const double* r0, size_t n, double* partials,
                                           const SolverScalars* sc, int b) {
  const double alpha = sc->alpha, omega = sc->omega;
  double val[2] = {0.0, 0.0};
  for (size_t q = b * (size_t)kReduceBlock + threadIdx.x; q < n; q += (size_t)kVecGrid * kReduceBlock) {
    x[q] += alpha * y[q] + omega * z[q];
    double rn = s[q] - omega * t[q];
    r[q] = rn;
    val[0] += rn * rn;
    val[1] += rn * r0[q];
  }
  block_partials<2>(val, partials, b);
}
__global__ void k_bicg_xr_yz(double* x, const double* y, const double* z, const double* s, const double* t, double* r,
                             const double* r0, size_t n, double* partials, const SolverScalars* sc) {
  if (sc->done) return;
  bicg_xr_yz(x, y, z, s, t, r, r0, n, partials, sc, blockIdx.x);
}

// p = r + beta (p - omega v)
__device__ __forceinline__ void bicg_p(double* p, const double* r, const double* v, size_t n,
                                       const SolverScalars* sc, int b) {
  const double beta = sc->beta, omega = sc->omega;
  for (size_t q = b * (size_t)kReduceBlock + threadIdx.x; q < n; q += (size_t)kVecGrid * kReduceBlock)
    p[q] = r[q] + beta * (p[q] - omega * v[q]);
}
__global__ void k_bicg_p(double* p, const double* r, const double* v, size_t n, const SolverScalars* sc) {
  if (sc->done) return;
  bicg_p(p, r, v, n, sc, blockIdx.x);
}

__global__ void k_copy(double* dst, const double* src, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

// generic dot product partials
__global__ void k_dot(const double* a, const double* b, size_t n, double* partials) {
  double val[1] = {0.0};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    val[0] += a[q] * b[q];
  block_partials<1>(val, partials, blockIdx.x);
}

__global__ void k_finalize_one(const double* partials, int nparts, double* out) {
  double s = finalize_sum(partials, 0, nparts);
  if (threadIdx.x == 0) *out = s;
}

// solve_poisson!.jl:69-72: solution ./= sqrt(pbar^2 / rhobar); dpip = solution / dt
__global__ void k_solution_to_dpip(double* sol, const double* fcs, double* dpip, Geo g, double dtinv) {
  int i = g.i0 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = g.j0 + blockIdx.y;
  int k = g.k0 + blockIdx.z;
  if (i > g.i1) return;
  size_t o = g.v(i - g.i0 + 1, j - g.j0 + 1, k - g.k0 + 1), c = g.h(i, j, k);
  double s = sol[o] / fcs[c];
  sol[o] = s;
  dpip[c] = dtinv * s;
}

// ---------------------------------------------------------------------------
// correct! (correct!.jl:164-385)
// ---------------------------------------------------------------------------
__global__ void k_correct_u(Fields F, Phys ph, Geo g, const double* betar, const double* dpip, double* corx, double* u,
                            Range3 R, double dt, double rf) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  double factor = 1.0;
  if (ph.damp) factor += dt * 0.5 * (betar[g.h(i, j, k)] + betar[g.h(i + 1, j, k)]) * rf;
  double gradient = pressure_gradient_u(F, ph, g, dpip, i, j, k);
  size_t c = g.h(i, j, k);
  corx[c] = dt / factor * gradient;
  double jpedger = wind_factor_u(F, g, i, j, k);
  u[c] -= jpedger * corx[c];
}

__global__ void k_correct_v(Fields F, Phys ph, Geo g, const double* betar, const double* dpip, double* cory, double* v,
                            Range3 R, double dt, double rf) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  double factor = 1.0;
  if (ph.damp) factor += dt * 0.5 * (betar[g.h(i, j, k)] + betar[g.h(i, j + 1, k)]) * rf;
  double gradient = pressure_gradient_v(F, ph, g, dpip, i, j, k);
  size_t c = g.h(i, j, k);
  cory[c] = dt / factor * gradient;
  double jpedgef = wind_factor_v(F, g, i, j, k);
  v[c] -= jpedgef * cory[c];
}

template <bool COMP>
__global__ void k_correct_w(Fields F, Phys ph, Geo g, const double* betar, const double* dpip, const double* corx,
                            const double* cory, double* w, Range3 R, double dt, double rf) {
  F.comp = COMP ? 1 : 0;  // folds the model's branches
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k), up = g.h(i, j, k + 1);
  const double jc = F.jac[c], ju = F.jac[up], dt2 = dt * dt;
  double factor = 1.0;
  factor += dt * (ju * betar[c] + jc * betar[up]) / (jc + ju) * rf;
  double n2edgeu = (ju * F.n2[c] + jc * F.n2[up]) / (jc + ju);
  double gradient = pressure_gradient_w(F, ph, g, dpip, i, j, k);
  double jpedgeu = wind_factor_w(F, g, i, j, k);
  double fw = buoyancy_factor_w(F, g, i, j, k);
  w[c] += -dt / (factor + fw * n2edgeu * dt2) * jpedgeu * gradient -
          1.0 / (factor + fw * n2edgeu * dt2) * fw * n2edgeu * dt2 * jpedgeu * 0.5 *
              (ju * (F.m13[c] * (corx[c] + corx[g.h(i - 1, j, k)]) + F.m23[c] * (cory[c] + cory[g.h(i, j - 1, k)])) +
               jc * (F.m13[up] * (corx[up] + corx[g.h(i - 1, j, k + 1)]) +
                     F.m23[up] * (cory[up] + cory[g.h(i, j - 1, k + 1)]))) /
              (jc + ju);
}

template <bool COMP>
__global__ void k_correct_rhop(Fields F, Phys ph, Geo g, const double* betar, const double* dpip, const double* corx,
                               const double* cory, double* rhop, Range3 R, double dt, double rf) {
  F.comp = COMP ? 1 : 0;  // folds the model's branches
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  const size_t c = g.h(i, j, k);
  const double dt2 = dt * dt;
  double factor = 1.0;
  factor += dt * betar[c] * rf;
  double lower_gradient = pressure_gradient_w(F, ph, g, dpip, i, j, k - 1);
  double upper_gradient = pressure_gradient_w(F, ph, g, dpip, i, j, k);
  if (k == g.k0) lower_gradient = 0.0;
  else if (k == g.z_size + g.nbz) upper_gradient = 0.0;
  double gradient = 0.5 * (lower_gradient + upper_gradient);
  double fb = buoyancy_factor_rhop(F, g, i, j, k);
  const double n2c = F.n2[c], jc = F.jac[c];
  double db = -1.0 / (factor + fb * n2c * dt2) *
              (-fb * n2c * dt2 * jc * gradient +
               fb * n2c * dt * jc * factor * 0.5 *
                   (F.m13[c] * (corx[c] + corx[g.h(i - 1, j, k)]) + F.m23[c] * (cory[c] + cory[g.h(i, j - 1, k)])));
  rhop[c] -= (F.rho[c] + F.rhobar[c]) / ph.g_ndim * db;
}

__global__ void k_correct_pip(double* pip, const double* dpip, Range3 R, Geo g) {
  int i, j, k;
  if (!range_ijk(R, blockIdx.x * (long long)blockDim.x + threadIdx.x, i, j, k)) return;
  size_t c = g.h(i, j, k);
  pip[c] += dpip[c];
}

Fields fields_of(pcf_state* st) { return make_fields(st); }


// ---------------------------------------------------------------------------
// Ensemble kernels: one launch over the solves of a batched iteration
// (EnsList, members of an EnsTable).  A block finds its solve from the
// list's first-block offsets and runs the single-solve kernel's body for
// that member's block; the blocks of a solve that has stopped (sc->done, in
// the iteration after its last) return at once.  Every member runs the
// same blocks as its single solve, so it gets the same partial sums and the
// same bits.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ens_find(const int* start, int count, int b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// line solve of one pseudo-time step: v[out] = (line solve of v[src])
template <bool FIRST>
__global__ void __launch_bounds__(64) kE_thomas(const EnsTable* __restrict__ E, const EnsList* __restrict__ L,
                                                int src, int out) {
  const int a = ens_find(L->thomas_start, L->count, blockIdx.x);
  const EnsMember& M = E->m[L->member[a]];
  if (M.sc->done) return;
  const int b = blockIdx.x - L->thomas_start[a];
  const int gx = (M.nx + 63) / 64;
  pc_thomas<FIRST, 64>(M.P, M.E, M.Q, M.v[src], M.v[out], M.nx, M.ny, M.nz, M.deta, b % gx, b / gx);
}

// v[out] = A v[in] (MODE_TOTAL, with NDOT dots against v[d1]) or
// v[in] + deta (H v[in] - v[sin]) (MODE_PC_RHS)
template <int MODE, int NDOT>
__global__ void __launch_bounds__(256) kE_pair(const EnsTable* __restrict__ E, const EnsList* __restrict__ L, int in,
                                               int out, int d1, int sin) {
  const int a = ens_find(L->pair_start, L->count, blockIdx.x);
  const EnsMember& M = E->m[L->member[a]];
  if (M.sc->done) return;
  const int b = blockIdx.x - L->pair_start[a];
  const int gx = (M.nx + 63) / 64, gy = (M.ny + kPairBlockY - 1) / kPairBlockY;
  ApplyArgs A;
  A.T = M.T;
  A.s = M.v[in];
  A.ls = M.v[out];
  A.nx = M.nx;
  A.ny = M.ny;
  A.nz = M.nz;
  stencil_pair<MODE, NDOT>(A, d1 >= 0 ? M.v[d1] : nullptr, sin >= 0 ? M.v[sin] : nullptr, M.deta, M.partials,
                           b % gx, (b / gx) % gy, b / (gx * gy), b);
}

// the scalar steps, one block of 1024 threads per solve
__global__ void kE_alpha(const EnsTable* __restrict__ E, const EnsList* __restrict__ L) {
  const EnsMember& M = E->m[L->member[blockIdx.x]];
  bicg_alpha(M.partials, M.aparts, M.sc);
}
__global__ void kE_omega(const EnsTable* __restrict__ E, const EnsList* __restrict__ L) {
  const EnsMember& M = E->m[L->member[blockIdx.x]];
  bicg_omega(M.partials, M.aparts, M.sc);
}
__global__ void kE_check(const EnsTable* __restrict__ E, const EnsList* __restrict__ L) {
  const EnsMember& M = E->m[L->member[blockIdx.x]];
  bicg_check(M.partials, kVecGrid, M.sc, M.nxyz, M.maxit);
}

// the vector steps, kVecGrid blocks per solve
__global__ void __launch_bounds__(256) kE_s(const EnsTable* __restrict__ E, const EnsList* __restrict__ L) {
  const EnsMember& M = E->m[L->member[blockIdx.x / kVecGrid]];
  if (M.sc->done) return;
  bicg_s(M.v[EV_R], M.v[EV_V], M.v[EV_S], M.n, M.sc, blockIdx.x % kVecGrid);
}
// YZ: the textbook update x += alpha y + omega z (mixed); else x += alpha p + omega s
template <bool YZ>
__global__ void __launch_bounds__(256) kE_xr(const EnsTable* __restrict__ E, const EnsList* __restrict__ L) {
  const EnsMember& M = E->m[L->member[blockIdx.x / kVecGrid]];
  if (M.sc->done) return;
  const int b = blockIdx.x % kVecGrid;
  if (YZ)
    bicg_xr_yz(M.v[EV_X], M.v[EV_Y], M.v[EV_Z], M.v[EV_S], M.v[EV_T], M.v[EV_R], M.v[EV_R0], M.n, M.partials, M.sc,
               b);
  else
    bicg_xr(M.v[EV_X], M.v[EV_P], M.v[EV_S], M.v[EV_T], M.v[EV_R], M.v[EV_R0], M.n, M.partials, M.sc, b);
}
__global__ void __launch_bounds__(256) kE_p(const EnsTable* __restrict__ E, const EnsList* __restrict__ L) {
  const EnsMember& M = E->m[L->member[blockIdx.x / kVecGrid]];
  if (M.sc->done) return;
  bicg_p(M.v[EV_P], M.v[EV_R], M.v[EV_V], M.n, M.sc, blockIdx.x % kVecGrid);
}

// the solves' done flags, by table slot, for the host's look at the batch
__global__ void kE_flags(const EnsTable* __restrict__ E, const EnsList* __restrict__ L, int32_t* flags) {
  const int a = threadIdx.x;
  if (a < L->count) flags[L->member[a]] = E->m[L->member[a]].sc->done;
}

}  // namespace


int compute_lhs(pcf_state* st, double* tolref_out) {
  const Geo& g = st->g;
  dim3 blk(64, 4), grd(blocks_for(g.nx, 64), blocks_for(g.ny, 4), g.nz);
  int nparts = grd.x * grd.y * grd.z;
  hipLaunchKernelGGL(k_compute_lhs, grd, blk, 0, st->stream, fields_of(st), st->ph, g, st->rhs, st->partials);
  hipLaunchKernelGGL(k_finalize_lhs, dim3(1), dim3(1024), 0, st->stream, st->partials, nparts, st->sc,
                     (double)st->cfg.x_size * st->cfg.y_size * st->cfg.z_size);
  PCF_HIP(hipGetLastError());
  PCF_HIP(hipMemcpyAsync(st->sc_host, st->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  // compute_lhs!.jl:143-157 (the two global sums are RMS over x_size y_size z_size)
  const double divl2 = st->sc_host->divl2, divl2_norm = st->sc_host->divl2_norm;
  double tolref;
  if (divl2_norm != 0.0) {
    tolref = divl2 / divl2_norm;
  } else if (divl2 == 0.0) {
    tolref = 1.0;
  } else {
    set_error("Error in compute_rhs: divl2_norm = 0 while divl2 != 0!");
    return PCF_E_DIVNORM;
  }
  *tolref_out = tolref;
  return PCF_OK;
}

// The compressible model's centre-coefficient term (dP/dpi / dt) / (dt R / kappa)
// (compute_operator!.jl:854-860) per interior cell; evaluated apart so that
// its pow does not raise the register count of the operator kernel.
__global__ void k_operator_shift(const double* p, Phys ph, Geo g, double dt, double* shift) {
  const int i = g.i0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = g.j0 + blockIdx.y;
  const int k = g.k0 + blockIdx.z;
  if (i > g.i1) return;
  double d = dpdpi(ph, p[g.h(i, j, k)]);
  shift[g.v(i - g.i0 + 1, j - g.j0 + 1, k - g.k0 + 1)] = (d / dt) / (dt * ph.rsp / ph.kappa);
}

int compute_operator(pcf_state* st, double dt, double rf) {
  const Geo& g = st->g;
  dim3 blk(64, 4), grd(blocks_for(g.nx, 64), blocks_for(g.ny, 4), g.nz);
  ProfScope prof(st, PCF_K_COMPUTE_OPERATOR);
  const double* shift = nullptr;
  if (st->compressible) {  // pc_tmp is free outside the preconditioner
    hipLaunchKernelGGL(k_operator_shift, dim3(blocks_for(g.nx, 64), g.ny, g.nz), dim3(64), 0, st->stream, st->p,
                       st->ph, g, dt, st->pc_tmp);
    shift = st->pc_tmp;
  }
  if (st->compressible)
    hipLaunchKernelGGL(k_compute_operator<true>, grd, blk, 0, st->stream, fields_of(st), st->betar, st->fcscal, g,
                       st->T, dt, rf, st->ph.damp, shift);
  else
    hipLaunchKernelGGL(k_compute_operator<false>, grd, blk, 0, st->stream, fields_of(st), st->betar, st->fcscal, g,
                       st->T, dt, rf, st->ph.damp, shift);
  PCF_HIP(hipGetLastError());
  st->pc_factors_valid = false;
  st->t32_valid = false;
  return PCF_OK;
}

static ApplyArgs apply_args(pcf_state* st, const double* s, double* ls) {
  ApplyArgs A;
  A.T = st->T;
  A.s = s;
  A.ls = ls;
  A.nx = st->g.nx;
  A.ny = st->g.ny;
  A.nz = st->g.nz;
  return A;
}

static dim3 apply_grid(const Geo& g) { return dim3(blocks_for(g.nx, 64), blocks_for(g.ny, 4), g.nz); }

// The two-cells-per-thread kernels need an even nx (16-byte aligned pairs).
static bool use_pairs(const Geo& g) { return g.nx % 2 == 0; }
static dim3 pair_grid(const Geo& g) { return dim3(blocks_for(g.nx, 64), blocks_for(g.ny, kPairBlockY), g.nz); }
static dim3 pair_block() { return dim3(32, kPairBlockY); }

// Number of partials the dot-fused apply writes.
static int apply_parts(const Geo& g) {
  const dim3 grd = use_pairs(g) ? pair_grid(g) : apply_grid(g);
  return (int)(grd.x * grd.y * grd.z);
}

// The block partials of compute_lhs and of the dot-fused applies must fit
// the kPartialStride slots of one dot product (checked once, by pcf_create).
int check_reduction_sizes(const Geo& g, std::string& err) {
  const long long lhs = (long long)blocks_for(g.nx, 64) * blocks_for(g.ny, 4) * g.nz;
  const long long app = (long long)apply_parts(g);
  if (lhs > kPartialStride || app > kPartialStride) {
    err = "grid too large for the reductions: ceil(x_size / 64) * ceil(y_size / 4) * z_size = " +
          std::to_string(lhs > app ? lhs : app) + " block partials, at most " + std::to_string(kPartialStride);
    return PCF_E_CONFIG;
  }
  return PCF_OK;
}

// ls = A s (TOTAL) with NDOT fused dot products: <ls, d1> (and <ls, ls>).
template <int NDOT>
static void launch_total(pcf_state* st, const double* s, double* ls, const double* d1, const SolverScalars* sc) {
  const Geo& g = st->g;
  if (use_pairs(g))
    hipLaunchKernelGGL((k_stencil_pair<MODE_TOTAL, NDOT>), pair_grid(g), pair_block(), 0, st->stream,
                       apply_args(st, s, ls), d1, nullptr, 0.0, st->partials, sc);
  else
    hipLaunchKernelGGL((k_apply<true, NDOT>), apply_grid(g), dim3(64, 4), 0, st->stream, apply_args(st, s, ls), d1,
                       nullptr, st->partials, sc);
}

// out = A^T r for the total operator: the tap t of cell p, T[t][p] r[p],
// lands on cell p + offset_t (periodic in x and y, dropped beyond the z
// walls exactly where the apply drops it), so cell q gathers T[t][q - off_t]
// r[q - off_t] over the 25 taps.  Used to train the initial-guess network on
// the residual; the reference has no transposed apply.
__constant__ int kTapOffset[T_COUNT][3] = {
    {0, 0, 0},   {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},   {0, 0, -1},  {0, 0, 1},
    {-1, 0, -1}, {-1, 0, 1},  {1, 0, -1},  {1, 0, 1},   {0, -1, -1}, {0, -1, 1},  {0, 1, -1},
    {0, 1, 1},   {0, 0, -2},  {0, 0, 2},   {-1, 0, -2}, {-1, 0, 2},  {1, 0, -2},  {1, 0, 2},
    {0, -1, -2}, {0, -1, 2},  {0, 1, -2},  {0, 1, 2}};

__global__ void __launch_bounds__(256) k_apply_transpose(ApplyArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (i >= nx || j >= ny) return;
  const size_t sxy = (size_t)nx * ny;
  double v = 0.0;
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) {
    const int pk = k - kTapOffset[t][2];
    if (pk < 0 || pk >= nz) continue;
    int pi = i - kTapOffset[t][0], pj = j - kTapOffset[t][1];
    pi = pi < 0 ? pi + nx : (pi >= nx ? pi - nx : pi);
    pj = pj < 0 ? pj + ny : (pj >= ny ? pj - ny : pj);
    const size_t p = (size_t)pi + (size_t)nx * pj + sxy * pk;
    v += A.T.a[t][p] * A.s[p];
  }
  A.ls[(size_t)i + (size_t)nx * j + sxy * k] = v;
}

int apply_operator_transpose(pcf_state* st, const double* r, double* out) {
  const Geo& g = st->g;
  hipLaunchKernelGGL(k_apply_transpose, dim3(blocks_for(g.nx, 64), blocks_for(g.ny, 4), g.nz), dim3(64, 4), 0,
                     st->stream, apply_args(st, r, out));
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int apply_operator(pcf_state* st, const double* sin, double* ls, bool total) {
  const Geo& g = st->g;
  if (total) {
    launch_total<0>(st, sin, ls, nullptr, nullptr);
  } else if (use_pairs(g)) {
    hipLaunchKernelGGL((k_stencil_pair<MODE_HORIZONTAL, 0>), pair_grid(g), pair_block(), 0, st->stream,
                       apply_args(st, sin, ls), nullptr, nullptr, 0.0, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL((k_apply<false, 0>), apply_grid(g), dim3(64, 4), 0, st->stream, apply_args(st, sin, ls), nullptr,
                       nullptr, nullptr, nullptr);
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// Line-solve factors of the current operator (computed once per operator).
int pc_factors(pcf_state* st) {
  if (st->pc_factors_valid) return PCF_OK;
  const Geo& g = st->g;
  hipLaunchKernelGGL(k_pc_factors, dim3(blocks_for(g.nx, 64), g.ny), dim3(64), 0, st->stream, st->T.a[T_AC],
                     st->T.a[T_AD], st->T.a[T_AU], st->pc_P, st->pc_E, st->pc_Q, g.nx, g.ny, g.nz, st->deta);
  PCF_HIP(hipGetLastError());
  st->pc_factors_valid = true;
  return PCF_OK;
}

// The line solve of one pseudo-time step (k_pc_thomas) with the LDS depth
// for the grid: at most 4 x 256 one-wavefront blocks fill the 256 CUs at
// four blocks each (80 levels in LDS), more blocks are better served by five
// blocks per CU (64 levels).
constexpr int kThomasBlocksPerCuAt80 = 4;
constexpr int kCUs = 256;
template <bool FIRST>
static void launch_thomas(pcf_state* st, const double* src, double* out, const SolverScalars* sc) {
  const Geo& g = st->g;
  const dim3 tgrd(blocks_for(g.nx, 64), g.ny), tblk(64);
  if ((long long)tgrd.x * tgrd.y <= (long long)kThomasBlocksPerCuAt80 * kCUs)
    hipLaunchKernelGGL((k_pc_thomas<FIRST, 80>), tgrd, tblk, 0, st->stream, st->pc_P, st->pc_E, st->pc_Q, src, out,
                       g.nx, g.ny, g.nz, st->deta, sc);
  else
    hipLaunchKernelGGL((k_pc_thomas<FIRST, 64>), tgrd, tblk, 0, st->stream, st->pc_P, st->pc_E, st->pc_Q, src, out,
                       g.nx, g.ny, g.nz, st->deta, sc);
}

// M^-1 sin -> sout with preconditioner_iterations pseudo-time steps.  sout
// must not alias sin.  `sc` (may be null) lets a converged solve skip it.
static int precondition(pcf_state* st, const double* sin, double* sout, const SolverScalars* sc) {
  const Geo& g = st->g;
  int rc = pc_factors(st);
  if (rc) return rc;
  const int iters = st->cfg.preconditioner_iterations;
  // pseudo-time step 1 from s = 0
  double* cur = iters == 1 ? sout : st->pc_tmp;
  {
    ProfScope prof(st, PCF_K_PC_THOMAS);
    launch_thomas<true>(st, sin, cur, sc);
  }
  for (int it = 2; it <= iters; ++it) {
    {
      ProfScope prof(st, PCF_K_PC_RHS);
      if (use_pairs(g))
        hipLaunchKernelGGL((k_stencil_pair<MODE_PC_RHS, 0>), pair_grid(g), pair_block(), 0, st->stream,
                           apply_args(st, cur, st->pc_rhs), nullptr, sin, st->deta, nullptr, sc);
      else
        hipLaunchKernelGGL(k_pc_rhs, apply_grid(g), dim3(64, 4), 0, st->stream, apply_args(st, cur, st->pc_rhs), sin,
                           st->deta, sc);
    }
    double* next = it == iters ? sout : (cur == st->pc_tmp ? sout : st->pc_tmp);
    {
      ProfScope prof(st, PCF_K_PC_THOMAS);
      launch_thomas<false>(st, st->pc_rhs, next, sc);
    }
    cur = next;
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// apply_preconditioner! in the configured precision (and, for the mixed
// solve, the configured preconditioner).
int apply_preconditioner(pcf_state* st, const double* sin, double* sout) {
  if (st->cfg.precision == PCF_PRECISION_MIXED) return precondition_mixed(st, sin, sout, nullptr);
  return precondition(st, sin, sout, nullptr);
}

int dot(pcf_state* st, const double* a, const double* b, double* out) {
  size_t n = st->g.nv();
  hipLaunchKernelGGL(k_dot, dim3(kVecGrid), dim3(kReduceBlock), 0, st->stream, a, b, n, st->partials);
  hipLaunchKernelGGL(k_finalize_one, dim3(1), dim3(1024), 0, st->stream, st->partials, kVecGrid, st->cfl_dev);
  PCF_HIP(hipGetLastError());
  PCF_HIP(hipMemcpyAsync(st->cfl_host, st->cfl_dev, sizeof(double), hipMemcpyDeviceToHost, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  *out = st->cfl_host[0];
  return PCF_OK;
}

// apply_bicgstab!.jl:24-150.  The host enqueues iterations and looks at the
// convergence flag of iteration j - 1 while iteration j is queued; kernels of
// iterations past convergence see the flag and return at once.
//
// mixed: the preconditioner runs in FP32 (mixed.hip) and the solution is
// accumulated in the textbook form x += alpha M^-1 p + omega M^-1 s, which
// keeps it consistent with the FP64 residual recurrence; guess (may be null):
// the solve starts from x0 = guess_scale * guess instead of 0, unless x0
// leaves a residual no smaller than |b|, in which case it starts from 0.
int bicgstab_core(pcf_state* st, double tolref, bool mixed, const double* guess, double guess_scale,
                  int* niter_out) {
  const Geo& g = st->g;
  const pcf_config& n = st->cfg;
  const size_t N = g.nv();
  const double nxyz = (double)n.x_size * n.y_size * n.z_size;
  const double tol = n.tolerance_is_relative ? n.tolerance : n.tolerance / tolref;
  SolverScalars* sc = st->sc;
  const bool pc = n.preconditioner != 0;
  // the learned stage follows the preconditioner in the first
  // stage_iterations iterations (flexible BiCGSTAB: the textbook update below
  // keeps x consistent with the residual recurrence whatever M^-1 does)
  const bool stage = mixed && pc && st->stage_net && st->stage_iterations > 0;
  auto prec = [&](const double* in, double* out, const SolverScalars* s, int iter) -> int {
    int rc0 = mixed ? precondition_mixed(st, in, out, s) : precondition(st, in, out, s);
    if (rc0 == PCF_OK && stage && iter <= st->stage_iterations) rc0 = learned_stage(st, in, out, s);
    return rc0;
  };
  if (guess) {
    hipLaunchKernelGGL(k_bicg_guess, dim3(kVecGrid), dim3(256), 0, st->stream, guess, st->solution, N, guess_scale);
    launch_total<0>(st, st->solution, st->bt, nullptr, nullptr);
    hipLaunchKernelGGL(k_bicg_init_guessed, dim3(kVecGrid), dim3(kReduceBlock), 0, st->stream, st->rhs, st->bt,
                       st->br0, st->br, st->bp, N, st->partials);
  } else {
    hipLaunchKernelGGL(k_bicg_init, dim3(kVecGrid), dim3(kReduceBlock), 0, st->stream, st->rhs, st->solution,
                       st->br0, st->br, st->bp, N, st->partials);
  }
  hipLaunchKernelGGL(k_bicg_start, dim3(1), dim3(1024), 0, st->stream, st->partials, kVecGrid, sc, nxyz, tol,
                     guess ? 1 : 0);
  PCF_HIP(hipGetLastError());
  PCF_HIP(hipMemcpyAsync(st->sc_host, sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, st->stream));
  PCF_HIP(hipStreamSynchronize(st->stream));
  if (guess && !(st->sc_host->res < st->sc_host->b_norm)) {
    // the guess does not reduce the residual: start from zero, as the plain solve
    hipLaunchKernelGGL(k_bicg_init, dim3(kVecGrid), dim3(kReduceBlock), 0, st->stream, st->rhs, st->solution,
                       st->br0, st->br, st->bp, N, st->partials);
    hipLaunchKernelGGL(k_bicg_start, dim3(1), dim3(1024), 0, st->stream, st->partials, kVecGrid, sc, nxyz, tol, 0);
    PCF_HIP(hipGetLastError());
    PCF_HIP(hipMemcpyAsync(st->sc_host, sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, st->stream));
    PCF_HIP(hipStreamSynchronize(st->stream));
    st->guess_fallbacks += 1;
  }
  if (st->sc_host->done) {
    st->last_residual = st->sc_host->b_norm > 0.0 ? st->sc_host->res / st->sc_host->b_norm : 0.0;
    *niter_out = 0;
    return PCF_OK;
  }
  const int aparts = apply_parts(g);
  const int maxit = n.poisson_iterations;
  // the line-solve factors of this operator, before any capture of the
  // iteration (precondition() computes them on first use)
  if (pc && !mixed) {
    int rcf = pc_factors(st);
    if (rcf) return rcf;
  }
  // y = M^-1 p and z = M^-1 s are both kept for the textbook update
  double* ybuf = st->bvpc;
  double* zbuf = mixed ? st->bzpc : st->bvpc;
  const bool yz = mixed && pc;
  // final x := M^-1 x (apply_bicgstab!.jl:132-135); the textbook form has
  // accumulated M^-1 already
  auto finish = [&]() -> int {
    if (pc && !yz) {
      hipLaunchKernelGGL(k_copy, dim3(kVecGrid), dim3(256), 0, st->stream, st->bs, st->solution, N);
      int rc1 = precondition(st, st->bs, st->solution, nullptr);
      if (rc1) return rc1;
    }
    PCF_HIP(hipGetLastError());
    return PCF_OK;
  };
  if (st->ens && ensemble_batchable(st, mixed, stage)) {
    // the iterations run in the ensemble's batch, which leaves the final
    // scalars in sc_host[0]
    int rce = ensemble_solve(st);
    if (rce) return rce;
    const SolverScalars fin = st->sc_host[0];
    *niter_out = fin.niter;
    st->last_residual = fin.res / fin.b_norm;
    if (fin.done != 1) return PCF_E_NOCONV;
    return finish();
  }
  if (st->ens) {
    std::unique_lock<std::mutex> lk(st->ens->mu);
    st->ens->solves_alone += 1;
  }
  SolverScalars* hs = st->sc_host;  // pinned, two slots (no copy is pending here)
  std::memset(hs, 0, 2 * sizeof(SolverScalars));
  hipEvent_t ev[2];
  PCF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  PCF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int converged_at = -1, last_seen = 0;
  int rc = PCF_OK;
  // One BiCGSTAB iteration's launches.  Every argument is fixed for the
  // state (the scalars live in device memory), so the sequence is captured
  // once into a hipGraph and replayed as one launch per iteration; the
  // learned stage's iterations and profiled runs launch it directly.
  auto iteration = [&](int jb) -> int {
    int rc1 = PCF_OK;
    // v = A M^-1 p, <v, r0>
    if ((rc1 = pc ? prec(st->bp, ybuf, sc, jb) : PCF_OK)) return rc1;
    const double* vin = pc ? ybuf : st->bp;
    {
      ProfScope prof(st, PCF_K_APPLY_TOTAL_DOT);
      launch_total<1>(st, vin, st->bv, st->br0, sc);
    }
    {
      ProfScope prof(st, PCF_K_BICG_VECTOR);
      hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(1024), 0, st->stream, st->partials, aparts, sc);
      hipLaunchKernelGGL(k_bicg_s, dim3(kVecGrid), dim3(256), 0, st->stream, st->br, st->bv, st->bs, N, sc);
    }
    // t = A M^-1 s, <t, s>, <t, t>
    if ((rc1 = pc ? prec(st->bs, zbuf, sc, jb) : PCF_OK)) return rc1;
    const double* sin2 = pc ? zbuf : st->bs;
    {
      ProfScope prof(st, PCF_K_APPLY_TOTAL_DOT);
      launch_total<2>(st, sin2, st->bt, st->bs, sc);
    }
    {
      ProfScope prof(st, PCF_K_BICG_VECTOR);
      hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(1024), 0, st->stream, st->partials, aparts, sc);
      if (yz)
        hipLaunchKernelGGL(k_bicg_xr_yz, dim3(kVecGrid), dim3(kReduceBlock), 0, st->stream, st->solution, ybuf, zbuf,
                           st->bs, st->bt, st->br, st->br0, N, st->partials, sc);
      else
        hipLaunchKernelGGL(k_bicg_xr, dim3(kVecGrid), dim3(kReduceBlock), 0, st->stream, st->solution, st->bp,
                           st->bs, st->bt, st->br, st->br0, N, st->partials, sc);
      hipLaunchKernelGGL(k_bicg_check, dim3(1), dim3(1024), 0, st->stream, st->partials, kVecGrid, sc, nxyz);
      hipLaunchKernelGGL(k_bicg_p, dim3(kVecGrid), dim3(256), 0, st->stream, st->bp, st->br, st->bv, N, sc);
    }
    PCF_HIP(hipGetLastError());
    return PCF_OK;
  };
  for (int jb = 1; jb <= maxit; ++jb) {
    // (never for an ensemble member: its service thread synchronises
    // while this thread would capture)
    const bool graphed =
        st->solver_graphs && !st->ens && !st->prof_on && !(stage && jb <= st->stage_iterations);
    if (graphed) {
      hipGraphExec_t& exec = st->iter_graph[mixed ? 1 : 0];
      if (!exec) {
        hipGraph_t graph = nullptr;
        PCF_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
        rc = iteration(jb);
        hipError_t e = hipStreamEndCapture(st->stream, &graph);
        if (rc) {
          if (graph) (void)hipGraphDestroy(graph);
          break;
        }
        PCF_HIP(e);
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PCF_HIP(e);
      }
      PCF_HIP(hipGraphLaunch(exec, st->stream));
    } else if ((rc = iteration(jb))) {
      break;
    }
    PCF_HIP(hipMemcpyAsync(hs + (jb & 1), sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, st->stream));
    PCF_HIP(hipEventRecord(ev[jb & 1], st->stream));
    if (jb >= 2) {
      PCF_HIP(hipEventSynchronize(ev[(jb - 1) & 1]));
      last_seen = jb - 1;
      if (hs[(jb - 1) & 1].done) {
        converged_at = hs[(jb - 1) & 1].niter;
        break;
      }
    }
  }
  if (rc == PCF_OK && converged_at < 0) {
    PCF_HIP(hipStreamSynchronize(st->stream));
    int slot = (last_seen + 1) & 1;
    // the newest copy holds the final state
    SolverScalars fin = hs[slot];
    for (int s = 0; s < 2; ++s)
      if (hs[s].niter > fin.niter) fin = hs[s];
    if (fin.done) converged_at = fin.niter;
    else {
      *niter_out = fin.niter;
      st->last_residual = fin.res / fin.b_norm;
      (void)hipEventDestroy(ev[0]);
      (void)hipEventDestroy(ev[1]);
      return PCF_E_NOCONV;
    }
  }
  PCF_HIP(hipStreamSynchronize(st->stream));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  SolverScalars fin = hs[0].niter >= hs[1].niter ? hs[0] : hs[1];
  st->last_residual = fin.res / fin.b_norm;
  *niter_out = converged_at;
  return finish();
}

// ---------------------------------------------------------------------------
// Ensemble batch (pcf_ensemble): the BiCGSTAB iterations of every waiting
// member of one solve kind, one launch per kernel over all of them.
// ---------------------------------------------------------------------------

// The member table covers the line preconditioner (FP64, or FP32 in the
// mixed solve) on grids with an even x_size (the two-cell apply kernels; a
// multiple of 4 for the FP32 pc-rhs kernel).  Members with the same key
// share a table; -1: not covered.
int ensemble_key(const pcf_state* st) {
  const pcf_config& n = st->cfg;
  const bool mixed = n.precision == PCF_PRECISION_MIXED;
  if (n.preconditioner == 0 || !use_pairs(st->g)) return -1;
  if (mixed && (n.preconditioner_type == PCF_PC_MULTIGRID || st->g.nx % 4 != 0)) return -1;
  const int bf16 = mixed && n.preconditioner_storage == PCF_STORAGE_BF16 ? 1 : 0;
  return (mixed ? 1 : 0) | bf16 << 1 | n.preconditioner_iterations << 2;
}

// A solve joins the batch when its member is in a table and the solve is
// the member's own kind, without the learned stage or profiling.
bool ensemble_batchable(const pcf_state* st, bool mixed, bool stage) {
  return st->ens_group >= 0 && mixed == (st->cfg.precision == PCF_PRECISION_MIXED) && !stage && !st->prof_on;
}

// M^-1 v[in] -> v[out] for every solve of the list (FP64 line
// preconditioner; the pseudo-time steps and buffers of precondition())
static void ensemble_precondition(hipStream_t s, const EnsTable* dT, const EnsList* dL, const EnsList& L, int iters,
                                  int in, int out) {
  const int tb = L.thomas_start[L.count], pb = L.pair_start[L.count];
  int cur = iters == 1 ? out : EV_PCTMP;
  hipLaunchKernelGGL(kE_thomas<true>, dim3(tb), dim3(64), 0, s, dT, dL, in, cur);
  for (int it = 2; it <= iters; ++it) {
    hipLaunchKernelGGL((kE_pair<MODE_PC_RHS, 0>), dim3(pb), pair_block(), 0, s, dT, dL, cur, EV_PCRHS, -1, in);
    const int next = it == iters ? out : (cur == EV_PCTMP ? out : EV_PCTMP);
    hipLaunchKernelGGL(kE_thomas<false>, dim3(tb), dim3(64), 0, s, dT, dL, EV_PCRHS, next);
    cur = next;
  }
}

static void ensemble_fill(pcf_ensemble::Group& grp) {
  EnsTable& H = *grp.host;
  const std::vector<pcf_state*>& mem = grp.slots;
  H.count = (int)mem.size();
  grp.pair_blocks.assign(H.count, 0);
  grp.thomas_blocks.assign(H.count, 0);
  grp.rhs32_blocks.assign(H.count, 0);
  for (int q = 0; q < H.count; ++q) {
    pcf_state* st = mem[q];
    const Geo& g = st->g;
    EnsMember& M = H.m[q];
    std::memset(&M, 0, sizeof(M));
    M.T = st->T;
    double* v[EV_COUNT] = {st->solution, st->br0, st->br, st->bp, st->bv, st->bs, st->bt, st->bvpc,
                           st->bzpc ? st->bzpc : st->bvpc, st->pc_tmp, st->pc_rhs};
    for (int e = 0; e < EV_COUNT; ++e) M.v[e] = v[e];
    M.P = st->pc_P;
    M.E = st->pc_E;
    M.Q = st->pc_Q;
    M.P32 = st->pc32_P;
    M.E32 = st->pc32_E;
    M.Q32 = st->pc32_Q;
    for (int t = 0; t < T_COUNT; ++t) {
      M.T32[t] = st->T32[t];
      M.T16[t] = st->T16[t];
    }
    M.col32 = st->m_pc_tmp;
    M.cur32 = st->m_pc_cur;
    M.rhs32 = st->m_pc_rhs;
    M.partials = st->partials;
    M.sc = st->sc;
    M.n = g.nv();
    M.nxyz = (double)st->cfg.x_size * st->cfg.y_size * st->cfg.z_size;
    M.deta = st->deta;
    M.nx = g.nx;
    M.ny = g.ny;
    M.nz = g.nz;
    M.aparts = apply_parts(g);
    M.maxit = st->cfg.poisson_iterations;
    const dim3 pg = pair_grid(g);
    grp.pair_blocks[q] = (int)(pg.x * pg.y * pg.z);
    grp.thomas_blocks[q] = (int)(blocks_for(g.nx, 64) * g.ny);
    grp.rhs32_blocks[q] = ensemble_rhs32_blocks(st);
  }
}

// The group's table after a change of its members (between iterations:
// nothing of the ensemble is in flight afterwards).
int ensemble_upload(pcf_ensemble* e, pcf_ensemble::Group& grp) {
  hipStream_t s = e->stream;
  PCF_HIP(hipStreamSynchronize(s));
  if (!grp.table) PCF_HIP(hipMalloc((void**)&grp.table, sizeof(EnsTable)));
  if (!grp.list) PCF_HIP(hipMalloc((void**)&grp.list, sizeof(EnsList)));
  if (!grp.list_host) PCF_HIP(hipHostMalloc((void**)&grp.list_host, 2 * sizeof(EnsList), hipHostMallocDefault));
  if (!grp.host) grp.host.reset(new EnsTable());
  for (int q = 0; q < (int)grp.slots.size(); ++q) grp.slots[q]->ens_slot = q;
  ensemble_fill(grp);
  PCF_HIP(hipMemcpyAsync(grp.table, grp.host.get(), sizeof(EnsTable), hipMemcpyHostToDevice, s));
  PCF_HIP(hipStreamSynchronize(s));
  grp.stale = false;
  e->table_builds += 1;
  return PCF_OK;
}

// One BiCGSTAB iteration (bicgstab_core's loop body) of every solve in the
// group's batch, each kernel one launch whose grid covers these solves; the
// list goes through pinned slot `parity` (the copy of the slot's previous
// use has completed: the host has seen the iteration after it end); the
// done flags go to `flags` by table slot.
int ensemble_iterate(pcf_ensemble* e, pcf_ensemble::Group& grp, int parity, int32_t* flags) {
  const int cnt = (int)grp.active.size();
  if (cnt == 0) return PCF_OK;
  const bool mixed = (grp.key & 1) != 0;
  const bool bf16 = (grp.key & 2) != 0;
  const int iters = grp.key >> 2;
  hipStream_t s = e->stream;
  EnsList& L = grp.list_host[parity];
  L.count = cnt;
  int pb = 0, tb = 0, rb = 0;
  for (int a = 0; a < cnt; ++a) {
    const int m = grp.active[a]->ens_slot;
    L.member[a] = m;
    L.pair_start[a] = pb;
    L.thomas_start[a] = tb;
    L.rhs32_start[a] = rb;
    pb += grp.pair_blocks[m];
    tb += grp.thomas_blocks[m];
    rb += grp.rhs32_blocks[m];
  }
  L.pair_start[cnt] = pb;
  L.thomas_start[cnt] = tb;
  L.rhs32_start[cnt] = rb;
  PCF_HIP(hipMemcpyAsync(grp.list, &L, sizeof(EnsList), hipMemcpyHostToDevice, s));
  const EnsTable* dT = grp.table;
  const EnsList* dL = grp.list;
  const int vb = cnt * kVecGrid;
  auto prec = [&](int in, int out) {
    if (mixed) ensemble_precondition_mixed(s, dT, dL, L, iters, bf16, in, out);
    else ensemble_precondition(s, dT, dL, L, iters, in, out);
  };
  prec(EV_P, EV_Y);
  hipLaunchKernelGGL((kE_pair<MODE_TOTAL, 1>), dim3(pb), pair_block(), 0, s, dT, dL, EV_Y, EV_V, EV_R0, -1);
  hipLaunchKernelGGL(kE_alpha, dim3(cnt), dim3(1024), 0, s, dT, dL);
  hipLaunchKernelGGL(kE_s, dim3(vb), dim3(256), 0, s, dT, dL);
  const int z = mixed ? EV_Z : EV_Y;
  prec(EV_S, z);
  hipLaunchKernelGGL((kE_pair<MODE_TOTAL, 2>), dim3(pb), pair_block(), 0, s, dT, dL, z, EV_T, EV_S, -1);
  hipLaunchKernelGGL(kE_omega, dim3(cnt), dim3(1024), 0, s, dT, dL);
  if (mixed) hipLaunchKernelGGL(kE_xr<true>, dim3(vb), dim3(kReduceBlock), 0, s, dT, dL);
  else hipLaunchKernelGGL(kE_xr<false>, dim3(vb), dim3(kReduceBlock), 0, s, dT, dL);
  hipLaunchKernelGGL(kE_check, dim3(cnt), dim3(1024), 0, s, dT, dL);
  hipLaunchKernelGGL(kE_p, dim3(vb), dim3(256), 0, s, dT, dL);
  hipLaunchKernelGGL(kE_flags, dim3(1), dim3(kEnsMax), 0, s, dT, dL, flags);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int apply_bicgstab(pcf_state* st, double tolref, int* niter_out) {
  return bicgstab_core(st, tolref, false, nullptr, 0.0, niter_out);
}

// apply_bicgstab! in the configured precision; dt (0 if unknown) scales the
// warm start of the mixed solve.
int solve(pcf_state* st, double tolref, double dt, int* niter) {
  if (st->cfg.precision == PCF_PRECISION_MIXED) return apply_bicgstab_mixed(st, tolref, dt, niter);
  return apply_bicgstab(st, tolref, niter);
}

// apply_corrector!.jl:38-66 with solve_poisson!.jl:39-75
int apply_corrector(pcf_state* st, double dt, double rf, int* niter) {
  double tolref = 0.0;
  int rc = compute_lhs(st, &tolref);
  if (rc) return rc;
  if (dt == 0.0) {
    set_error("Error in solve_poisson!: dt = 0.0!");
    return PCF_E_DT_ZERO;
  }
  const double dtinv = 1.0 / dt;
  if ((rc = compute_operator(st, dt, rf))) return rc;
  rc = solve(st, tolref, dt, niter);
  if (rc) return rc;
  const Geo& g = st->g;
  hipLaunchKernelGGL(k_solution_to_dpip, dim3(blocks_for(g.nx, 64), g.ny, g.nz), dim3(64), 0, st->stream,
                     st->solution, st->fcscal, st->dpip, g, dtinv);
  PCF_HIP(hipGetLastError());
  double* f[1] = {st->dpip};
  if ((rc = bc_xy(st, f, 1))) return rc;
  if ((rc = bc_z(st, st->dpip, +1, false))) return rc;
  return correct(st, dt, rf);
}

int correct(pcf_state* st, double dt, double rf) {
  const Geo& g = st->g;
  Fields F = fields_of(st);
  // U and V: k0:k1 (one process holds the top), i0-1:i1 resp. j0-1:j1
  Range3 ru{g.i0 - 1, g.nx + 1, g.j0, g.ny, g.k0, g.nz};
  Range3 rv{g.i0, g.nx, g.j0 - 1, g.ny + 1, g.k0, g.nz};
  Range3 rw{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz - 1};
  Range3 rr{g.i0, g.nx, g.j0, g.ny, g.k0, g.nz};
  Range3 rp{g.i0 - 1, g.nx + 2, g.j0 - 1, g.ny + 2, g.k0 - 1, g.nz + 2};
  hipLaunchKernelGGL(k_correct_u, dim3(blocks_for(ru.count(), 256)), dim3(256), 0, st->stream, F, st->ph, g, st->betar,
                     st->dpip, st->corx, st->u, ru, dt, rf);
  hipLaunchKernelGGL(k_correct_v, dim3(blocks_for(rv.count(), 256)), dim3(256), 0, st->stream, F, st->ph, g, st->betar,
                     st->dpip, st->cory, st->v, rv, dt, rf);
  if (rw.nk > 0)
    hipLaunchKernelGGL(st->compressible ? k_correct_w<true> : k_correct_w<false>, dim3(blocks_for(rw.count(), 256)),
                       dim3(256), 0, st->stream, F, st->ph, g, st->betar, st->dpip, st->corx, st->cory, st->w, rw, dt,
                       rf);
  hipLaunchKernelGGL(st->compressible ? k_correct_rhop<true> : k_correct_rhop<false>, dim3(blocks_for(rr.count(), 256)),
                     dim3(256), 0, st->stream, F, st->ph, g, st->betar, st->dpip, st->corx, st->cory, st->rhop, rr, dt,
                     rf);
  hipLaunchKernelGGL(k_correct_pip, dim3(blocks_for(rp.count(), 256)), dim3(256), 0, st->stream, st->pip, st->dpip,
                     rp, g);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

}  // namespace pcf

// ---------------------------------------------------------------------------
// pcf_dot_noalloc: pcf_dot's kernels with their block partials in a
// file-scope device array instead of the state's buffer, for callers that
// cannot allocate device memory (graph capture, serving loops).  The array is
// addressed through the pointer hipGetSymbolAddress gives on each device; the
// array's host-side name is not a device address.
// ---------------------------------------------------------------------------
#include <mutex>

static __device__ __align__(256) double scratch[pcf::kVecGrid];  // 16 KiB per device

namespace {

constexpr int kNoallocMaxDevices = 64;

std::mutex g_scratch_mutex;
double* g_scratch_ptr[kNoallocMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
int scratch_on_current_device(double** out) {
  int dev = -1;
  PCF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoallocMaxDevices) {
    pcf::set_error("pcf_dot_noalloc: device index out of range");
    return PCF_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PCF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PCF_OK;
}

}  // namespace

extern "C" int pcf_dot_noalloc(const double* d_a, const double* d_b, int64_t n, double* d_result, void* stream) {
  using namespace pcf;
  if (!d_a || !d_b || !d_result) {
    set_error("pcf_dot_noalloc: null pointer argument");
    return PCF_E_ARG;
  }
  if (n < 0 || n > PCF_DOT_NOALLOC_MAX_N) {
    set_error("pcf_dot_noalloc: n = " + std::to_string((long long)n) + " outside [0, " +
              std::to_string((long long)PCF_DOT_NOALLOC_MAX_N) + "]");
    return PCF_E_ARG;
  }
  double* partials = nullptr;
  int rc = scratch_on_current_device(&partials);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the kernels and launch shape of pcf_dot, so the result has its bits
  hipLaunchKernelGGL(k_dot, dim3(kVecGrid), dim3(kReduceBlock), 0, s, d_a, d_b, (size_t)n, partials);
  hipLaunchKernelGGL(k_finalize_one, dim3(1), dim3(1024), 0, s, partials, kVecGrid, d_result);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}
