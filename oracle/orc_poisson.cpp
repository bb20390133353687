// ORACLE -- test infrastructure only (see orc.hpp).
// src/PoissonSolver/*.jl and src/MPIOperations/compute_global_dot_product.jl.
#include <cstdio>

#include "orc.hpp"

namespace orc {

// Julia's pairwise mapreduce (Base.mapreduce_impl, block size 1024): split
// at the midpoint until a block is shorter than 1024, sum blocks in order.
template <class F>
static double pairwise(F f, size_t lo, size_t hi) {  // [lo, hi]
  if (hi - lo < 1024) {
    double v = f(lo);
    for (size_t q = lo + 1; q <= hi; ++q) v += f(q);
    return v;
  }
  size_t mid = lo + ((hi - lo) >> 1);
  double a, b;
  if (hi - lo > (1u << 18)) {
#pragma omp task shared(a)
    a = pairwise(f, lo, mid);
    b = pairwise(f, mid + 1, hi);
#pragma omp taskwait
  } else {
    a = pairwise(f, lo, mid);
    b = pairwise(f, mid + 1, hi);
  }
  return a + b;
}

// Summation order of the global sums: 0 = Julia's pairwise order above (the
// default); B > 0 = blocks of B terms summed in order, the block sums added
// forwards, or backwards when g_order_backwards.  Only the iteration-count
// spread test (tests/test_oracle_invariants.py) changes it: the count of a
// BiCGSTAB solve near its tolerance depends on the rounding of these sums.
static size_t g_order_block = 0;
static bool g_order_backwards = false;

template <class F>
static double preduce(F f, size_t n) {
  double r = 0.0;
  if (g_order_block == 0) {
#pragma omp parallel
#pragma omp single
    r = pairwise(f, 0, n - 1);
    return r;
  }
  const size_t B = g_order_block, nb = (n + B - 1) / B;
  for (size_t bb = 0; bb < nb; ++bb) {
    const size_t b = g_order_backwards ? nb - 1 - bb : bb;
    double v = 0.0;
    for (size_t q = b * B; q < n && q < b * B + B; ++q) v += f(q);
    r += v;
  }
  return r;
}

// compute_global_dot_product.jl:20-36 (one process: no Allreduce)
double dot(const Arr3& a, const Arr3& b) {
  const double* x = a.a.data();
  const double* y = b.a.data();
  return preduce([&](size_t q) { return x[q] * y[q]; }, a.a.size());
}

// sum(a -> a^2, r) of apply_bicgstab!.jl:63,119
double sumsq(const Arr3& a) {
  const double* x = a.a.data();
  return preduce([&](size_t q) { return x[q] * x[q]; }, a.a.size());
}

// compute_lhs!.jl:68-223
int compute_lhs(State& st, double& tolref) {
  const double ma = st.c.ma, kappa = st.c.kappa, dx = st.dx, dy = st.dy, dz = st.dz;
  Arr3 &jac = st.jac, &pbar = st.pbar, &rhobar = st.rhobar;
  Arr3 &u = st.pr.u, &v = st.pr.v, &w = st.pr.w;
  double divl2 = 0.0, divl2_norm = 0.0;
  const double ma2k = ma * ma * kappa;
  (void)ma2k;
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double fcscal = std::sqrt(pbar(i, j, k) * pbar(i, j, k) / rhobar(i, j, k));
        double ur = u(i, j, k), ul = u(i - 1, j, k), vf = v(i, j, k), vb = v(i, j - 1, k);
        double wu = w(i, j, k), wd = w(i, j, k - 1);
        if (st.compressible) {  // compute_lhs!.jl:162-223: the winds carry J P already
          double heating = conductive_heating(st, i, j, k) * (ma * ma) * kappa;
          double bu = (ur - ul) / dx / jac(i, j, k) * (ma * ma) * kappa;
          double bv = (vf - vb) / dy / jac(i, j, k) * (ma * ma) * kappa;
          double bw = (wu - wd) / dz / jac(i, j, k) * (ma * ma) * kappa;
          bu /= fcscal;
          bv /= fcscal;
          bw /= fcscal;
          heating /= fcscal;
          double b = bu + bv + bw - heating;
          st.rhs(i - st.i0 + 1, j - st.j0 + 1, k - st.k0 + 1) = b;
          divl2 += b * b;
          divl2_norm += bu * bu + bv * bv + bw * bw + heating * heating;
          continue;
        }
        double pedger = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i + 1, j, k) * pbar(i + 1, j, k));
        double pedgel = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i - 1, j, k) * pbar(i - 1, j, k));
        double pedgef = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j + 1, k) * pbar(i, j + 1, k));
        double pedgeb = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j - 1, k) * pbar(i, j - 1, k));
        double pedgeu = jac(i, j, k) * jac(i, j, k + 1) * (pbar(i, j, k) + pbar(i, j, k + 1)) /
                        (jac(i, j, k) + jac(i, j, k + 1));
        double pedged = jac(i, j, k) * jac(i, j, k - 1) * (pbar(i, j, k) + pbar(i, j, k - 1)) /
                        (jac(i, j, k) + jac(i, j, k - 1));
        double bu = (pedger * ur - pedgel * ul) / dx / jac(i, j, k) * (ma * ma) * kappa;
        double bv = (pedgef * vf - pedgeb * vb) / dy / jac(i, j, k) * (ma * ma) * kappa;
        double bw = (pedgeu * wu - pedged * wd) / dz / jac(i, j, k) * (ma * ma) * kappa;
        bu /= fcscal;
        bv /= fcscal;
        bw /= fcscal;
        double b = bu + bv + bw;
        st.rhs(i - st.i0 + 1, j - st.j0 + 1, k - st.k0 + 1) = b;
        divl2 += b * b;
        divl2_norm += bu * bu + bv * bv + bw * bw;
      }
  const double n = (double)st.cfg.x_size * st.cfg.y_size * st.cfg.z_size;
  divl2 = std::sqrt(divl2 / st.cfg.x_size / st.cfg.y_size / st.cfg.z_size);
  divl2_norm = std::sqrt(divl2_norm / st.cfg.x_size / st.cfg.y_size / st.cfg.z_size);
  (void)n;
  if (divl2_norm != 0.0) {
    tolref = divl2 / divl2_norm;
  } else if (divl2 == 0.0) {
    tolref = 1.0;
  } else {
    st.err = "Error in compute_rhs: divl2_norm = 0 while divl2 != 0!";
    return PCF_E_DIVNORM;
  }
  return PCF_OK;
}

static inline double fcs(State& st, int i, int j, int k) {
  return std::sqrt(st.pbar(i, j, k) * st.pbar(i, j, k) / st.rhobar(i, j, k));
}

// compute_operator!.jl:40-1606
void compute_operator(State& st, double dt, double rf) {
  const int z_size = st.cfg.z_size, nbz = st.nbz, ko = st.ko, k0 = st.k0;
  const bool damp = st.cfg.damp_horizontal_wind_on_rhs != 0;
  const double dx = st.dx, dy = st.dy, dz = st.dz;
  Arr3 &jac = st.jac, &pbar = st.pbar, &rhobar = st.rhobar, &n2 = st.n2, &betar = st.betar;
  Arr3 &m13 = st.met13, &m23 = st.met23, &m33 = st.met33, &rho = st.pr.rho;
  const double dt2 = dt * dt;
#pragma omp parallel for schedule(static)
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        const bool kbot = ko + k == k0, kbot1 = ko + k == k0 + 1;
        const bool ktop = ko + k == z_size + nbz, ktop1 = ko + k == z_size + nbz - 1;
        double fcscal = fcs(st, i, j, k);
        double jacinv = 1.0 / jac(i, j, k);
        // P coefficients (divergence)
        double pedgerdiv = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i + 1, j, k) * pbar(i + 1, j, k));
        double pedgeldiv = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i - 1, j, k) * pbar(i - 1, j, k));
        double pedgefdiv = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j + 1, k) * pbar(i, j + 1, k));
        double pedgebdiv = 0.5 * (jac(i, j, k) * pbar(i, j, k) + jac(i, j - 1, k) * pbar(i, j - 1, k));
        double pedgeudiv = jac(i, j, k) * jac(i, j, k + 1) * (pbar(i, j, k) + pbar(i, j, k + 1)) /
                           (jac(i, j, k) + jac(i, j, k + 1));
        double pedgeddiv = jac(i, j, k) * jac(i, j, k - 1) * (pbar(i, j, k) + pbar(i, j, k - 1)) /
                           (jac(i, j, k) + jac(i, j, k - 1));
        // P coefficients (pressure gradient)
        double pedgergra = 0.5 * (pbar(i, j, k) + pbar(i + 1, j, k));
        double pedgelgra = 0.5 * (pbar(i, j, k) + pbar(i - 1, j, k));
        double pedgefgra = 0.5 * (pbar(i, j, k) + pbar(i, j + 1, k));
        double pedgebgra = 0.5 * (pbar(i, j, k) + pbar(i, j - 1, k));
        double pedgeugra = (jac(i, j, k + 1) * pbar(i, j, k) + jac(i, j, k) * pbar(i, j, k + 1)) /
                           (jac(i, j, k) + jac(i, j, k + 1));
        double pedgedgra = (jac(i, j, k - 1) * pbar(i, j, k) + jac(i, j, k) * pbar(i, j, k - 1)) /
                           (jac(i, j, k) + jac(i, j, k - 1));
        double puedgergra = 0.5 * (pbar(i, j, k + 1) + pbar(i + 1, j, k + 1));
        double puedgelgra = 0.5 * (pbar(i, j, k + 1) + pbar(i - 1, j, k + 1));
        double puedgefgra = 0.5 * (pbar(i, j, k + 1) + pbar(i, j + 1, k + 1));
        double puedgebgra = 0.5 * (pbar(i, j, k + 1) + pbar(i, j - 1, k + 1));
        double pdedgergra = 0.5 * (pbar(i, j, k - 1) + pbar(i + 1, j, k - 1));
        double pdedgelgra = 0.5 * (pbar(i, j, k - 1) + pbar(i - 1, j, k - 1));
        double pdedgefgra = 0.5 * (pbar(i, j, k - 1) + pbar(i, j + 1, k - 1));
        double pdedgebgra = 0.5 * (pbar(i, j, k - 1) + pbar(i, j - 1, k - 1));
        // density coefficients
        double rhobaredger = 0.5 * (rhobar(i, j, k) + rhobar(i + 1, j, k));
        double rhobaredgel = 0.5 * (rhobar(i, j, k) + rhobar(i - 1, j, k));
        double rhobaredgef = 0.5 * (rhobar(i, j, k) + rhobar(i, j + 1, k));
        double rhobaredgeb = 0.5 * (rhobar(i, j, k) + rhobar(i, j - 1, k));
        double rhobaredgeu = (jac(i, j, k + 1) * rhobar(i, j, k) + jac(i, j, k) * rhobar(i, j, k + 1)) /
                             (jac(i, j, k) + jac(i, j, k + 1));
        double rhobaredged = (jac(i, j, k - 1) * rhobar(i, j, k) + jac(i, j, k) * rhobar(i, j, k - 1)) /
                             (jac(i, j, k) + jac(i, j, k - 1));
        double rhoedger = 0.5 * (rho(i, j, k) + rho(i + 1, j, k)) + rhobaredger;
        double rhoedgel = 0.5 * (rho(i, j, k) + rho(i - 1, j, k)) + rhobaredgel;
        double rhoedgef = 0.5 * (rho(i, j, k) + rho(i, j + 1, k)) + rhobaredgef;
        double rhoedgeb = 0.5 * (rho(i, j, k) + rho(i, j - 1, k)) + rhobaredgeb;
        double rhoedgeu = (jac(i, j, k + 1) * rho(i, j, k) + jac(i, j, k) * rho(i, j, k + 1)) /
                              (jac(i, j, k) + jac(i, j, k + 1)) + rhobaredgeu;
        double rhoedged = (jac(i, j, k - 1) * rho(i, j, k) + jac(i, j, k) * rho(i, j, k - 1)) /
                              (jac(i, j, k) + jac(i, j, k - 1)) + rhobaredged;
        double fwu = buoyancy_factor_w(st, i, j, k);
        double fwd = buoyancy_factor_w(st, i, j, k - 1);
        double rhouedger = 0.5 * (rho(i, j, k + 1) + rho(i + 1, j, k + 1) + rhobar(i, j, k + 1) + rhobar(i + 1, j, k + 1));
        double rhouedgel = 0.5 * (rho(i, j, k + 1) + rho(i - 1, j, k + 1) + rhobar(i, j, k + 1) + rhobar(i - 1, j, k + 1));
        double rhouedgef = 0.5 * (rho(i, j, k + 1) + rho(i, j + 1, k + 1) + rhobar(i, j, k + 1) + rhobar(i, j + 1, k + 1));
        double rhouedgeb = 0.5 * (rho(i, j, k + 1) + rho(i, j - 1, k + 1) + rhobar(i, j, k + 1) + rhobar(i, j - 1, k + 1));
        double rhodedger = 0.5 * (rho(i, j, k - 1) + rho(i + 1, j, k - 1) + rhobar(i, j, k - 1) + rhobar(i + 1, j, k - 1));
        double rhodedgel = 0.5 * (rho(i, j, k - 1) + rho(i - 1, j, k - 1) + rhobar(i, j, k - 1) + rhobar(i - 1, j, k - 1));
        double rhodedgef = 0.5 * (rho(i, j, k - 1) + rho(i, j + 1, k - 1) + rhobar(i, j, k - 1) + rhobar(i, j + 1, k - 1));
        double rhodedgeb = 0.5 * (rho(i, j, k - 1) + rho(i, j - 1, k - 1) + rhobar(i, j, k - 1) + rhobar(i, j - 1, k - 1));
        double n2edgeu = (jac(i, j, k + 1) * n2(i, j, k) + jac(i, j, k) * n2(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
        double n2edged = (jac(i, j, k - 1) * n2(i, j, k) + jac(i, j, k) * n2(i, j, k - 1)) / (jac(i, j, k) + jac(i, j, k - 1));
        // metric-tensor interpolation
        double met13edger = 0.5 * (m13(i, j, k) + m13(i + 1, j, k));
        double met13edgel = 0.5 * (m13(i, j, k) + m13(i - 1, j, k));
        double met23edgef = 0.5 * (m23(i, j, k) + m23(i, j + 1, k));
        double met23edgeb = 0.5 * (m23(i, j, k) + m23(i, j - 1, k));
        double met13edgeu = (jac(i, j, k + 1) * m13(i, j, k) + jac(i, j, k) * m13(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
        double met23edgeu = (jac(i, j, k + 1) * m23(i, j, k) + jac(i, j, k) * m23(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
        double met33edgeu = (jac(i, j, k + 1) * m33(i, j, k) + jac(i, j, k) * m33(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
        double met13edged = (jac(i, j, k - 1) * m13(i, j, k) + jac(i, j, k) * m13(i, j, k - 1)) / (jac(i, j, k) + jac(i, j, k - 1));
        double met23edged = (jac(i, j, k - 1) * m23(i, j, k) + jac(i, j, k) * m23(i, j, k - 1)) / (jac(i, j, k) + jac(i, j, k - 1));
        double met33edged = (jac(i, j, k - 1) * m33(i, j, k) + jac(i, j, k) * m33(i, j, k - 1)) / (jac(i, j, k) + jac(i, j, k - 1));
        double met13uedger = 0.5 * (m13(i, j, k + 1) + m13(i + 1, j, k + 1));
        double met13uedgel = 0.5 * (m13(i, j, k + 1) + m13(i - 1, j, k + 1));
        double met23uedgef = 0.5 * (m23(i, j, k + 1) + m23(i, j + 1, k + 1));
        double met23uedgeb = 0.5 * (m23(i, j, k + 1) + m23(i, j - 1, k + 1));
        double met13dedger = 0.5 * (m13(i, j, k - 1) + m13(i + 1, j, k - 1));
        double met13dedgel = 0.5 * (m13(i, j, k - 1) + m13(i - 1, j, k - 1));
        double met23dedgef = 0.5 * (m23(i, j, k - 1) + m23(i, j + 1, k - 1));
        double met23dedgeb = 0.5 * (m23(i, j, k - 1) + m23(i, j - 1, k - 1));
        // Rayleigh damping (compute_operator!.jl:283-378)
        double facedger = 1.0, facedgel = 1.0, facedgef = 1.0, facedgeb = 1.0;
        double facuedger = 1.0, facuedgel = 1.0, facuedgef = 1.0, facuedgeb = 1.0;
        double facdedger = 1.0, facdedgel = 1.0, facdedgef = 1.0, facdedgeb = 1.0;
        double facedgeu = 1.0, facedged = 1.0;
        if (damp) {
          facedger += dt * 0.5 * (betar(i, j, k) + betar(i + 1, j, k)) * rf;
          facedgel += dt * 0.5 * (betar(i, j, k) + betar(i - 1, j, k)) * rf;
          facedgef += dt * 0.5 * (betar(i, j, k) + betar(i, j + 1, k)) * rf;
          facedgeb += dt * 0.5 * (betar(i, j, k) + betar(i, j - 1, k)) * rf;
          facuedger += dt * 0.5 * (betar(i, j, k + 1) + betar(i + 1, j, k + 1)) * rf;
          facuedgel += dt * 0.5 * (betar(i, j, k + 1) + betar(i - 1, j, k + 1)) * rf;
          facuedgef += dt * 0.5 * (betar(i, j, k + 1) + betar(i, j + 1, k + 1)) * rf;
          facuedgeb += dt * 0.5 * (betar(i, j, k + 1) + betar(i, j - 1, k + 1)) * rf;
          facdedger += dt * 0.5 * (betar(i, j, k - 1) + betar(i + 1, j, k - 1)) * rf;
          facdedgel += dt * 0.5 * (betar(i, j, k - 1) + betar(i - 1, j, k - 1)) * rf;
          facdedgef += dt * 0.5 * (betar(i, j, k - 1) + betar(i, j + 1, k - 1)) * rf;
          facdedgeb += dt * 0.5 * (betar(i, j, k - 1) + betar(i, j - 1, k - 1)) * rf;
        }
        facedgeu += dt * (jac(i, j, k + 1) * betar(i, j, k) + jac(i, j, k) * betar(i, j, k + 1)) /
                    (jac(i, j, k) + jac(i, j, k + 1)) * rf;
        facedged += dt * (jac(i, j, k - 1) * betar(i, j, k) + jac(i, j, k) * betar(i, j, k - 1)) /
                    (jac(i, j, k) + jac(i, j, k - 1)) * rf;
        double imphoredger = 1.0 / facedger, imphoredgel = 1.0 / facedgel;
        double imphoredgef = 1.0 / facedgef, imphoredgeb = 1.0 / facedgeb;
        double imphoruedger = 1.0 / facuedger, imphoruedgel = 1.0 / facuedgel;
        double imphoruedgef = 1.0 / facuedgef, imphoruedgeb = 1.0 / facuedgeb;
        double imphordedger = 1.0 / facdedger, imphordedgel = 1.0 / facdedgel;
        double imphordedgef = 1.0 / facdedgef, imphordedgeb = 1.0 / facdedgeb;
        double impveredgeu = 1.0 / (facedgeu + fwu * n2edgeu * dt2);
        double impveredged = 1.0 / (facedged + fwd * n2edged * dt2);
        // Common factors of the N^2-coupled vertical terms (the products of
        // compute_operator!.jl:385-395 and 396-405 up to the metric factor).
        const double ju = jac(i, j, k + 1), jd = jac(i, j, k - 1), jc = jac(i, j, k);
        auto Tu = [&](double met, double jfac, double imp, double rhoe) {
          return jacinv / dz * pedgeudiv * impveredgeu * fwu * n2edgeu * dt2 * 0.5 * met * jfac /
                 (jc + ju) * imp / rhoe;
        };
        auto Td = [&](double met, double jfac, double imp, double rhoe) {
          return jacinv / dz * pedgeddiv * impveredged * fwd * n2edged * dt2 * 0.5 * met * jfac /
                 (jc + jd) * imp / rhoe;
        };
        // gradient coefficients G (compute_operator!.jl:380-719)
        double gedger, gedgel, gedgef, gedgeb;
        {
          double hr = jacinv / dx * pedgerdiv * imphoredger / rhoedger;
          double hl = -jacinv / dx * pedgeldiv * imphoredgel / rhoedgel;
          double hf = jacinv / dy * pedgefdiv * imphoredgef / rhoedgef;
          double hb = -jacinv / dy * pedgebdiv * imphoredgeb / rhoedgeb;
          if (kbot) {
            gedger = hr + Tu(m13(i, j, k), ju, imphoredger, rhoedger);
            gedgel = hl + Tu(m13(i, j, k), ju, imphoredgel, rhoedgel);
            gedgef = hf + Tu(m23(i, j, k), ju, imphoredgef, rhoedgef);
            gedgeb = hb + Tu(m23(i, j, k), ju, imphoredgeb, rhoedgeb);
          } else if (ktop) {
            gedger = hr - Td(m13(i, j, k), jd, imphoredger, rhoedger);
            gedgel = hl - Td(m13(i, j, k), jd, imphoredgel, rhoedgel);
            gedgef = hf - Td(m23(i, j, k), jd, imphoredgef, rhoedgef);
            gedgeb = hb - Td(m23(i, j, k), jd, imphoredgeb, rhoedgeb);
          } else {
            gedger = hr + Tu(m13(i, j, k), ju, imphoredger, rhoedger) - Td(m13(i, j, k), jd, imphoredger, rhoedger);
            gedgel = hl + Tu(m13(i, j, k), ju, imphoredgel, rhoedgel) - Td(m13(i, j, k), jd, imphoredgel, rhoedgel);
            gedgef = hf + Tu(m23(i, j, k), ju, imphoredgef, rhoedgef) - Td(m23(i, j, k), jd, imphoredgef, rhoedgef);
            gedgeb = hb + Tu(m23(i, j, k), ju, imphoredgeb, rhoedgeb) - Td(m23(i, j, k), jd, imphoredgeb, rhoedgeb);
          }
        }
        double gedgeu = ktop ? 0.0 : jacinv / dz * pedgeudiv * impveredgeu / rhoedgeu;
        double gedged = kbot ? 0.0 : -jacinv / dz * pedgeddiv * impveredged / rhoedged;
        double guedger = ktop ? 0.0 : Tu(m13(i, j, k + 1), jc, imphoruedger, rhouedger);
        double guedgel = ktop ? 0.0 : Tu(m13(i, j, k + 1), jc, imphoruedgel, rhouedgel);
        double guedgef = ktop ? 0.0 : Tu(m23(i, j, k + 1), jc, imphoruedgef, rhouedgef);
        double guedgeb = ktop ? 0.0 : Tu(m23(i, j, k + 1), jc, imphoruedgeb, rhouedgeb);
        double gdedger = kbot ? 0.0 : -Td(m13(i, j, k - 1), jc, imphordedger, rhodedger);
        double gdedgel = kbot ? 0.0 : -Td(m13(i, j, k - 1), jc, imphordedgel, rhodedgel);
        double gdedgef = kbot ? 0.0 : -Td(m23(i, j, k - 1), jc, imphordedgef, rhodedgef);
        double gdedgeb = kbot ? 0.0 : -Td(m23(i, j, k - 1), jc, imphordedgeb, rhodedgeb);

        // Products G * P of compute_operator!.jl:721-1467.
        const double R = gedger * pedgergra, L = gedgel * pedgelgra;
        const double F = gedgef * pedgefgra, B = gedgeb * pedgebgra;
        const double U = gedgeu * pedgeugra, D = gedged * pedgedgra;
        const double UR = guedger * puedgergra, UL = guedgel * puedgelgra;
        const double UF = guedgef * puedgefgra, UB = guedgeb * puedgebgra;
        const double DR = gdedger * pdedgergra, DL = gdedgel * pdedgelgra;
        const double DF = gdedgef * pdedgefgra, DB = gdedgeb * pdedgebgra;
        auto q4 = [&](double X, double m) { return X * 0.25 * m / dz; };
        auto q1 = [&](double X, double m) { return X * m / dz; };
        auto hx = [&](double X, double m, double d, double ja, double jb) { return X * 0.5 * m / d * ja / jb; };
        const double jr = jac(i + 1, j, k), jru = jac(i + 1, j, k + 1), jrd = jac(i + 1, j, k - 1);
        const double jl = jac(i - 1, j, k), jlu = jac(i - 1, j, k + 1), jld = jac(i - 1, j, k - 1);
        const double jf = jac(i, j + 1, k), jfu = jac(i, j + 1, k + 1), jfd = jac(i, j + 1, k - 1);
        const double jb = jac(i, j - 1, k), jbu = jac(i, j - 1, k + 1), jbd = jac(i, j - 1, k - 1);

        double ac, ar, al, af, ab, au, ad, aru, ard, alu, ald, afu, afd, abu, abd;
        double auu, add, aruu, ardd, aluu, aldd, afuu, afdd, abuu, abdd;
        // A(i,j,k)
        if (kbot) {
          ac = -R * (1.0 / dx + 0.75 * met13edger / dz) + L * (1.0 / dx - 0.75 * met13edgel / dz) -
               F * (1.0 / dy + 0.75 * met23edgef / dz) + B * (1.0 / dy - 0.75 * met23edgeb / dz) -
               U * met33edgeu / dz - q4(UR, met13uedger) - q4(UL, met13uedgel) - q4(UF, met23uedgef) -
               q4(UB, met23uedgeb);
        } else if (kbot1) {
          ac = -R / dx + L / dx - F / dy + B / dy - U * met33edgeu / dz + D * met33edged / dz -
               q4(UR, met13uedger) - q4(UL, met13uedgel) - q4(UF, met23uedgef) - q4(UB, met23uedgeb) +
               q1(DR, met13dedger) + q1(DL, met13dedgel) + q1(DF, met23dedgef) + q1(DB, met23dedgeb);
        } else if (ktop1) {
          ac = -R / dx + L / dx - F / dy + B / dy - U * met33edgeu / dz + D * met33edged / dz -
               q1(UR, met13uedger) - q1(UL, met13uedgel) - q1(UF, met23uedgef) - q1(UB, met23uedgeb) +
               q4(DR, met13dedger) + q4(DL, met13dedgel) + q4(DF, met23dedgef) + q4(DB, met23dedgeb);
        } else if (ktop) {
          ac = -R * (1.0 / dx - 0.75 * met13edger / dz) + L * (1.0 / dx + 0.75 * met13edgel / dz) -
               F * (1.0 / dy - 0.75 * met23edgef / dz) + B * (1.0 / dy + 0.75 * met23edgeb / dz) +
               D * met33edged / dz + q4(DR, met13dedger) + q4(DL, met13dedgel) + q4(DF, met23dedgef) +
               q4(DB, met23dedgeb);
        } else {
          ac = -R / dx + L / dx - F / dy + B / dy - U * met33edgeu / dz + D * met33edged / dz -
               q4(UR, met13uedger) - q4(UL, met13uedgel) - q4(UF, met23uedgef) - q4(UB, met23uedgeb) +
               q4(DR, met13dedger) + q4(DL, met13dedgel) + q4(DF, met23dedgef) + q4(DB, met23dedgeb);
        }
        if (st.compressible) {  // compute_operator!.jl:854-860
          double dpdpi_ = dpdpi(st, st.pr.p(i, j, k));
          ac -= (dpdpi_ / dt) / (dt * st.c.rsp / st.c.kappa);
        }
        // A(i+1,j,k)
        if (kbot) {
          ar = R * (1.0 / dx - 0.75 * met13edger / dz) + hx(U, met13edgeu, dx, jru, jr + jru) - q4(UR, met13uedger);
        } else if (kbot1) {
          ar = R / dx + hx(U, met13edgeu, dx, jru, jr + jru) + hx(D, met13edged, dx, jrd, jr + jrd) -
               q4(UR, met13uedger) + q1(DR, met13dedger);
        } else if (ktop1) {
          ar = R / dx + hx(U, met13edgeu, dx, jru, jr + jru) + hx(D, met13edged, dx, jrd, jr + jrd) -
               q1(UR, met13uedger) + q4(DR, met13dedger);
        } else if (ktop) {
          ar = R * (1.0 / dx + 0.75 * met13edger / dz) + hx(D, met13edged, dx, jrd, jr + jrd) + q4(DR, met13dedger);
        } else {
          ar = R / dx + hx(U, met13edgeu, dx, jru, jr + jru) + hx(D, met13edged, dx, jrd, jr + jrd) -
               q4(UR, met13uedger) + q4(DR, met13dedger);
        }
        // A(i-1,j,k)
        if (kbot) {
          al = -L * (1.0 / dx + 0.75 * met13edgel / dz) - hx(U, met13edgeu, dx, jlu, jl + jlu) - q4(UL, met13uedgel);
        } else if (kbot1) {
          al = -L / dx - hx(U, met13edgeu, dx, jlu, jl + jlu) - hx(D, met13edged, dx, jld, jl + jld) -
               q4(UL, met13uedgel) + q1(DL, met13dedgel);
        } else if (ktop1) {
          al = -L / dx - hx(U, met13edgeu, dx, jlu, jl + jlu) - hx(D, met13edged, dx, jld, jl + jld) -
               q1(UL, met13uedgel) + q4(DL, met13dedgel);
        } else if (ktop) {
          al = -L * (1.0 / dx - 0.75 * met13edgel / dz) - hx(D, met13edged, dx, jld, jl + jld) + q4(DL, met13dedgel);
        } else {
          al = -L / dx - hx(U, met13edgeu, dx, jlu, jl + jlu) - hx(D, met13edged, dx, jld, jl + jld) -
               q4(UL, met13uedgel) + q4(DL, met13dedgel);
        }
        // A(i,j+1,k)
        if (kbot) {
          af = F * (1.0 / dy - 0.75 * met23edgef / dz) + hx(U, met23edgeu, dy, jfu, jf + jfu) - q4(UF, met23uedgef);
        } else if (kbot1) {
          af = F / dy + hx(U, met23edgeu, dy, jfu, jf + jfu) + hx(D, met23edged, dy, jfd, jf + jfd) -
               q4(UF, met23uedgef) + q1(DF, met23dedgef);
        } else if (ktop1) {
          af = F / dy + hx(U, met23edgeu, dy, jfu, jf + jfu) + hx(D, met23edged, dy, jfd, jf + jfd) -
               q1(UF, met23uedgef) + q4(DF, met23dedgef);
        } else if (ktop) {
          af = F * (1.0 / dy + 0.75 * met23edgef / dz) + hx(D, met23edged, dy, jfd, jf + jfd) + q4(DF, met23dedgef);
        } else {
          af = F / dy + hx(U, met23edgeu, dy, jfu, jf + jfu) + hx(D, met23edged, dy, jfd, jf + jfd) -
               q4(UF, met23uedgef) + q4(DF, met23dedgef);
        }
        // A(i,j-1,k)
        if (kbot) {
          ab = -B * (1.0 / dy + 0.75 * met23edgeb / dz) - hx(U, met23edgeu, dy, jbu, jb + jbu) - q4(UB, met23uedgeb);
        } else if (kbot1) {
          ab = -B / dy - hx(U, met23edgeu, dy, jbu, jb + jbu) - hx(D, met23edged, dy, jbd, jb + jbd) -
               q4(UB, met23uedgeb) + q1(DB, met23dedgeb);
        } else if (ktop1) {
          ab = -B / dy - hx(U, met23edgeu, dy, jbu, jb + jbu) - hx(D, met23edged, dy, jbd, jb + jbd) -
               q1(UB, met23uedgeb) + q4(DB, met23dedgeb);
        } else if (ktop) {
          ab = -B * (1.0 / dy - 0.75 * met23edgeb / dz) - hx(D, met23edged, dy, jbd, jb + jbd) + q4(DB, met23dedgeb);
        } else {
          ab = -B / dy - hx(U, met23edgeu, dy, jbu, jb + jbu) - hx(D, met23edged, dy, jbd, jb + jbd) -
               q4(UB, met23uedgeb) + q4(DB, met23dedgeb);
        }
        // A(i,j,k+1)
        if (kbot) {
          au = q1(R, met13edger) + q1(L, met13edgel) + q1(F, met23edgef) + q1(B, met23edgeb) + U * met33edgeu / dz -
               UR / dx + UL / dx - UF / dy + UB / dy;
        } else if (kbot1) {
          au = q4(R, met13edger) + q4(L, met13edgel) + q4(F, met23edgef) + q4(B, met23edgeb) + U * met33edgeu / dz -
               UR / dx + UL / dx - UF / dy + UB / dy - q4(DR, met13dedger) - q4(DL, met13dedgel) -
               q4(DF, met23dedgef) - q4(DB, met23dedgeb);
        } else if (ktop1) {
          au = q4(R, met13edger) + q4(L, met13edgel) + q4(F, met23edgef) + q4(B, met23edgeb) + U * met33edgeu / dz -
               UR * (1.0 / dx - 0.75 * met13uedger / dz) + UL * (1.0 / dx + 0.75 * met13uedgel / dz) -
               UF * (1.0 / dy - 0.75 * met23uedgef / dz) + UB * (1.0 / dy + 0.75 * met23uedgeb / dz);
        } else if (ktop) {
          au = 0.0;
        } else {
          au = q4(R, met13edger) + q4(L, met13edgel) + q4(F, met23edgef) + q4(B, met23edgeb) + U * met33edgeu / dz -
               UR / dx + UL / dx - UF / dy + UB / dy;
        }
        // A(i,j,k-1)
        if (kbot) {
          ad = 0.0;
        } else if (kbot1) {
          ad = -q4(R, met13edger) - q4(L, met13edgel) - q4(F, met23edgef) - q4(B, met23edgeb) - D * met33edged / dz -
               DR * (1.0 / dx + 0.75 * met13dedger / dz) + DL * (1.0 / dx - 0.75 * met13dedgel / dz) -
               DF * (1.0 / dy + 0.75 * met23dedgef / dz) + DB * (1.0 / dy - 0.75 * met23dedgeb / dz);
        } else if (ktop1) {
          ad = -q4(R, met13edger) - q4(L, met13edgel) - q4(F, met23edgef) - q4(B, met23edgeb) - D * met33edged / dz -
               DR / dx + DL / dx - DF / dy + DB / dy + q4(UR, met13uedger) + q4(UL, met13uedgel) +
               q4(UF, met23uedgef) + q4(UB, met23uedgeb);
        } else if (ktop) {
          ad = -q1(R, met13edger) - q1(L, met13edgel) - q1(F, met23edgef) - q1(B, met23edgeb) - D * met33edged / dz -
               DR / dx + DL / dx - DF / dy + DB / dy;
        } else {
          ad = -q4(R, met13edger) - q4(L, met13edgel) - q4(F, met23edgef) - q4(B, met23edgeb) - D * met33edged / dz -
               DR / dx + DL / dx - DF / dy + DB / dy;
        }
        // A(i+1,j,k+1)
        if (kbot) {
          aru = q1(R, met13edger) + hx(U, met13edgeu, dx, jr, jr + jru) + UR / dx;
        } else if (kbot1) {
          aru = q4(R, met13edger) + hx(U, met13edgeu, dx, jr, jr + jru) + UR / dx - q4(DR, met13dedger);
        } else if (ktop1) {
          aru = q4(R, met13edger) + hx(U, met13edgeu, dx, jr, jr + jru) + UR * (1.0 / dx + 0.75 * met13uedger / dz);
        } else if (ktop) {
          aru = 0.0;
        } else {
          aru = q4(R, met13edger) + hx(U, met13edgeu, dx, jr, jr + jru) + UR / dx;
        }
        // A(i+1,j,k-1)
        if (kbot) {
          ard = 0.0;
        } else if (kbot1) {
          ard = -q4(R, met13edger) + hx(D, met13edged, dx, jr, jr + jrd) + DR * (1.0 / dx - 0.75 * met13dedger / dz);
        } else if (ktop1) {
          ard = -q4(R, met13edger) + hx(D, met13edged, dx, jr, jr + jrd) + DR / dx + q4(UR, met13uedger);
        } else if (ktop) {
          ard = -q1(R, met13edger) + hx(D, met13edged, dx, jr, jr + jrd) + DR / dx;
        } else {
          ard = -q4(R, met13edger) + hx(D, met13edged, dx, jr, jr + jrd) + DR / dx;
        }
        // A(i-1,j,k+1)
        if (kbot) {
          alu = q1(L, met13edgel) - hx(U, met13edgeu, dx, jl, jl + jlu) - UL / dx;
        } else if (kbot1) {
          alu = q4(L, met13edgel) - hx(U, met13edgeu, dx, jl, jl + jlu) - UL / dx - q4(DL, met13dedgel);
        } else if (ktop1) {
          alu = q4(L, met13edgel) - hx(U, met13edgeu, dx, jl, jl + jlu) - UL * (1.0 / dx - 0.75 * met13uedgel / dz);
        } else if (ktop) {
          alu = 0.0;
        } else {
          alu = q4(L, met13edgel) - hx(U, met13edgeu, dx, jl, jl + jlu) - UL / dx;
        }
        // A(i-1,j,k-1)
        if (kbot) {
          ald = 0.0;
        } else if (kbot1) {
          ald = -q4(L, met13edgel) - hx(D, met13edged, dx, jl, jl + jld) - DL * (1.0 / dx + 0.75 * met13dedgel / dz);
        } else if (ktop1) {
          ald = -q4(L, met13edgel) - hx(D, met13edged, dx, jl, jl + jld) - DL / dx + q4(UL, met13uedgel);
        } else if (ktop) {
          ald = -q1(L, met13edgel) - hx(D, met13edged, dx, jl, jl + jld) - DL / dx;
        } else {
          ald = -q4(L, met13edgel) - hx(D, met13edged, dx, jl, jl + jld) - DL / dx;
        }
        // A(i,j+1,k+1)
        if (kbot) {
          afu = q1(F, met23edgef) + hx(U, met23edgeu, dy, jf, jf + jfu) + UF / dy;
        } else if (kbot1) {
          afu = q4(F, met23edgef) + hx(U, met23edgeu, dy, jf, jf + jfu) + UF / dy - q4(DF, met23dedgef);
        } else if (ktop1) {
          afu = q4(F, met23edgef) + hx(U, met23edgeu, dy, jf, jf + jfu) + UF * (1.0 / dy + 0.75 * met23uedgef / dz);
        } else if (ktop) {
          afu = 0.0;
        } else {
          afu = q4(F, met23edgef) + hx(U, met23edgeu, dy, jf, jf + jfu) + UF / dy;
        }
        // A(i,j+1,k-1)
        if (kbot) {
          afd = 0.0;
        } else if (kbot1) {
          afd = -q4(F, met23edgef) + hx(D, met23edged, dy, jf, jf + jfd) + DF * (1.0 / dy - 0.75 * met23dedgef / dz);
        } else if (ktop1) {
          afd = -q4(F, met23edgef) + hx(D, met23edged, dy, jf, jf + jfd) + DF / dy + q4(UF, met23uedgef);
        } else if (ktop) {
          afd = -q1(F, met23edgef) + hx(D, met23edged, dy, jf, jf + jfd) + DF / dy;
        } else {
          afd = -q4(F, met23edgef) + hx(D, met23edged, dy, jf, jf + jfd) + DF / dy;
        }
        // A(i,j-1,k+1)
        if (kbot) {
          abu = q1(B, met23edgeb) - hx(U, met23edgeu, dy, jb, jb + jbu) - UB / dy;
        } else if (kbot1) {
          abu = q4(B, met23edgeb) - hx(U, met23edgeu, dy, jb, jb + jbu) - UB / dy - q4(DB, met23dedgeb);
        } else if (ktop1) {
          abu = q4(B, met23edgeb) - hx(U, met23edgeu, dy, jb, jb + jbu) - UB * (1.0 / dy - 0.75 * met23uedgeb / dz);
        } else if (ktop) {
          abu = 0.0;
        } else {
          abu = q4(B, met23edgeb) - hx(U, met23edgeu, dy, jb, jb + jbu) - UB / dy;
        }
        // A(i,j-1,k-1)
        if (kbot) {
          abd = 0.0;
        } else if (kbot1) {
          abd = -q4(B, met23edgeb) - hx(D, met23edged, dy, jb, jb + jbd) - DB * (1.0 / dy + 0.75 * met23dedgeb / dz);
        } else if (ktop1) {
          abd = -q4(B, met23edgeb) - hx(D, met23edged, dy, jb, jb + jbd) - DB / dy + q4(UB, met23uedgeb);
        } else if (ktop) {
          abd = -q1(B, met23edgeb) - hx(D, met23edged, dy, jb, jb + jbd) - DB / dy;
        } else {
          abd = -q4(B, met23edgeb) - hx(D, met23edged, dy, jb, jb + jbd) - DB / dy;
        }
        // A(i,j,k+2)
        if (kbot) {
          auu = -q4(R, met13edger) - q4(L, met13edgel) - q4(F, met23edgef) - q4(B, met23edgeb) + q4(UR, met13uedger) +
                q4(UL, met13uedgel) + q4(UF, met23uedgef) + q4(UB, met23uedgeb);
        } else if (ko + k >= z_size + nbz - 1) {
          auu = 0.0;
        } else {
          auu = q4(UR, met13uedger) + q4(UL, met13uedgel) + q4(UF, met23uedgef) + q4(UB, met23uedgeb);
        }
        // A(i,j,k-2)
        if (ko + k <= k0 + 1) {
          add = 0.0;
        } else if (ktop) {
          add = q4(R, met13edger) + q4(L, met13edgel) + q4(F, met23edgef) + q4(B, met23edgeb) - q4(DR, met13dedger) -
                q4(DL, met13dedgel) - q4(DF, met23dedgef) - q4(DB, met23dedgeb);
        } else {
          add = -q4(DR, met13dedger) - q4(DL, met13dedgel) - q4(DF, met23dedgef) - q4(DB, met23dedgeb);
        }
        // A(i+1,j,k+2), A(i+1,j,k-2), A(i-1,j,k+2), A(i-1,j,k-2)
        if (kbot) aruu = -q4(R, met13edger) + q4(UR, met13uedger);
        else if (ko + k >= z_size + nbz - 1) aruu = 0.0;
        else aruu = q4(UR, met13uedger);
        if (ko + k <= k0 + 1) ardd = 0.0;
        else if (ktop) ardd = q4(R, met13edger) - q4(DR, met13dedger);
        else ardd = -q4(DR, met13dedger);
        if (kbot) aluu = -q4(L, met13edgel) + q4(UL, met13uedgel);
        else if (ko + k >= z_size + nbz - 1) aluu = 0.0;
        else aluu = q4(UL, met13uedgel);
        if (ko + k <= k0 + 1) aldd = 0.0;
        else if (ktop) aldd = q4(L, met13edgel) - q4(DL, met13dedgel);
        else aldd = -q4(DL, met13dedgel);
        // A(i,j+1,k+2), A(i,j+1,k-2), A(i,j-1,k+2), A(i,j-1,k-2)
        if (kbot) afuu = -q4(F, met23edgef) + q4(UF, met23uedgef);
        else if (ko + k >= z_size + nbz - 1) afuu = 0.0;
        else afuu = q4(UF, met23uedgef);
        if (ko + k <= k0 + 1) afdd = 0.0;
        else if (ktop) afdd = q4(F, met23edgef) - q4(DF, met23dedgef);
        else afdd = -q4(DF, met23dedgef);
        if (kbot) abuu = -q4(B, met23edgeb) + q4(UB, met23uedgeb);
        else if (ko + k >= z_size + nbz - 1) abuu = 0.0;
        else abuu = q4(UB, met23uedgeb);
        if (ko + k <= k0 + 1) abdd = 0.0;
        else if (ktop) abdd = q4(B, met23edgeb) - q4(DB, met23dedgeb);
        else abdd = -q4(DB, met23dedgeb);

        // scaling (compute_operator!.jl:1545-1570)
        const int ia = i - st.i0 + 1, ja = j - st.j0 + 1, ka = k - st.k0 + 1;
        auto put = [&](int slot, double val, int di, int dj, int dk) {
          st.tensor[slot](ia, ja, ka) = val / fcscal / fcs(st, i + di, j + dj, k + dk);
        };
        st.tensor[T_AC](ia, ja, ka) = ac / (fcscal * fcscal);
        put(T_AR, ar, 1, 0, 0);
        put(T_AL, al, -1, 0, 0);
        put(T_AF, af, 0, 1, 0);
        put(T_AB, ab, 0, -1, 0);
        put(T_AU, au, 0, 0, 1);
        put(T_AD, ad, 0, 0, -1);
        put(T_ARU, aru, 1, 0, 1);
        put(T_ARD, ard, 1, 0, -1);
        put(T_ALU, alu, -1, 0, 1);
        put(T_ALD, ald, -1, 0, -1);
        put(T_AFU, afu, 0, 1, 1);
        put(T_AFD, afd, 0, 1, -1);
        put(T_ABU, abu, 0, -1, 1);
        put(T_ABD, abd, 0, -1, -1);
        put(T_AUU, auu, 0, 0, 2);
        put(T_ADD, add, 0, 0, -2);
        put(T_ARUU, aruu, 1, 0, 2);
        put(T_ARDD, ardd, 1, 0, -2);
        put(T_ALUU, aluu, -1, 0, 2);
        put(T_ALDD, aldd, -1, 0, -2);
        put(T_AFUU, afuu, 0, 1, 2);
        put(T_AFDD, afdd, 0, 1, -2);
        put(T_ABUU, abuu, 0, -1, 2);
        put(T_ABDD, abdd, 0, -1, -2);
      }
}

// apply_operator!.jl:48-272 (Total) and 275-479 (Horizontal).  At npz = 1 only
// the one-layer zonal and meridional halos of `s` are set; its vertical
// halos keep the zeros they were allocated with.
void apply_operator(State& st, Arr3& sin, Arr3& ls, bool total) {
  Arr3& s = st.s_op;
  for (int k = 1; k <= st.nz; ++k)
    for (int j = 1; j <= st.ny; ++j)
      for (int i = 1; i <= st.nx; ++i) s(i + st.i0 - 1, j + st.j0 - 1, k + st.k0 - 1) = sin(i, j, k);
  zonal_bc(st, s, 1, 1, 1);
  meridional_bc(st, s, 1, 1, 1);
  Arr3* T = st.tensor;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= st.nz; ++k)
    for (int j = 1; j <= st.ny; ++j)
      for (int i = 1; i <= st.nx; ++i) {
        const int is = i + st.i0 - 1, js = j + st.j0 - 1, ks = k + st.k0 - 1;
        double v;
        if (total) {
          v = T[T_AL](i, j, k) * s(is - 1, js, ks) + T[T_AR](i, j, k) * s(is + 1, js, ks) +
              T[T_AF](i, j, k) * s(is, js + 1, ks) + T[T_AB](i, j, k) * s(is, js - 1, ks) +
              T[T_AU](i, j, k) * s(is, js, ks + 1) + T[T_AD](i, j, k) * s(is, js, ks - 1) +
              T[T_AC](i, j, k) * s(is, js, ks);
        } else {
          v = T[T_AL](i, j, k) * s(is - 1, js, ks) + T[T_AR](i, j, k) * s(is + 1, js, ks) +
              T[T_AF](i, j, k) * s(is, js + 1, ks) + T[T_AB](i, j, k) * s(is, js - 1, ks);
        }
        v += T[T_ARU](i, j, k) * s(is + 1, js, ks + 1) + T[T_ARD](i, j, k) * s(is + 1, js, ks - 1) +
             T[T_ALU](i, j, k) * s(is - 1, js, ks + 1) + T[T_ALD](i, j, k) * s(is - 1, js, ks - 1) +
             T[T_AFU](i, j, k) * s(is, js + 1, ks + 1) + T[T_AFD](i, j, k) * s(is, js + 1, ks - 1) +
             T[T_ABU](i, j, k) * s(is, js - 1, ks + 1) + T[T_ABD](i, j, k) * s(is, js - 1, ks - 1) +
             T[T_AUU](i, j, k) * s(is, js, ks + 2) + T[T_ADD](i, j, k) * s(is, js, ks - 2) +
             T[T_ARUU](i, j, k) * s(is + 1, js, ks + 2) + T[T_ARDD](i, j, k) * s(is + 1, js, ks - 2) +
             T[T_ALUU](i, j, k) * s(is - 1, js, ks + 2) + T[T_ALDD](i, j, k) * s(is - 1, js, ks - 2) +
             T[T_AFUU](i, j, k) * s(is, js + 1, ks + 2) + T[T_AFDD](i, j, k) * s(is, js + 1, ks - 2) +
             T[T_ABUU](i, j, k) * s(is, js - 1, ks + 2) + T[T_ABDD](i, j, k) * s(is, js - 1, ks - 2);
        ls(i, j, k) = v;
      }
}

// apply_preconditioner!.jl:40-124 (ko == 0 and ko + nz == z_size)
void apply_preconditioner(State& st, Arr3& sin, Arr3& sout) {
  const double dtau = st.cfg.dtau;
  const int nx = st.nx, ny = st.ny, nz = st.nz;
  Arr3 &s_pc = st.s_pc, &q_pc = st.q_pc;
  Arr3 &au = st.tensor[T_AU], &ac = st.tensor[T_AC], &ad = st.tensor[T_AD];
  s_pc.fill(0.0);
  q_pc.fill(0.0);
  const double deta = dtau / (2 * (1 / (st.dx * st.dx) + 1 / (st.dy * st.dy)));
  for (int it = 1; it <= st.cfg.preconditioner_iterations; ++it) {
    apply_operator(st, s_pc, q_pc, false);
    for (size_t q = 0; q < s_pc.a.size(); ++q) s_pc.a[q] += deta * (q_pc.a[q] - sin.a[q]);
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        q_pc(i, j, 1) = deta * au(i, j, 1) / (1 - deta * ac(i, j, 1));
        s_pc(i, j, 1) /= 1 - deta * ac(i, j, 1);
        for (int k = 2; k <= nz; ++k) {
          double p = 1 / (1 - deta * ac(i, j, k) - deta * ad(i, j, k) * q_pc(i, j, k - 1));
          q_pc(i, j, k) = deta * au(i, j, k) * p;
          s_pc(i, j, k) = (s_pc(i, j, k) + deta * ad(i, j, k) * s_pc(i, j, k - 1)) * p;
        }
        for (int k = nz - 1; k >= 1; --k) s_pc(i, j, k) += q_pc(i, j, k) * s_pc(i, j, k + 1);
      }
  }
  sout.a = s_pc.a;
}

// apply_bicgstab!.jl:24-150
int apply_bicgstab(State& st, double tolref, int& niter) {
  const pcf_config& n = st.cfg;
  const double N = (double)n.x_size * n.y_size * n.z_size;
  (void)N;
  Arr3 &x = st.solution, &b = st.rhs, &p = st.bp, &r0 = st.br0, &rold = st.brold, &r = st.br;
  Arr3 &s = st.bs, &t = st.bt, &v = st.bv, &matvec = st.bmatvec, &v_pc = st.bv_pc;
  x.fill(0.0);
  const int maxit = n.poisson_iterations;
  const double tol = n.tolerance_is_relative ? n.tolerance : n.tolerance / tolref;
  apply_operator(st, x, matvec, true);
  for (size_t q = 0; q < r0.a.size(); ++q) r0.a[q] = b.a[q] - matvec.a[q];
  p.a = r0.a;
  r.a = r0.a;
  auto rms = [&](double ss) { return std::sqrt(ss / n.x_size / n.y_size / n.z_size); };
  double res = rms(sumsq(r));
  const double b_norm = res;
  st.last_residual = 0.0;
  if (res == 0.0 || res / b_norm <= tol) {
    niter = 0;
    return PCF_OK;
  }
  int jb = 0;
  while (jb < maxit) {
    ++jb;
    if (n.preconditioner) apply_preconditioner(st, p, v_pc);
    else v_pc.a = p.a;
    apply_operator(st, v_pc, matvec, true);
    v.a = matvec.a;
    double alpha = dot(r, r0) / dot(v, r0);
    for (size_t q = 0; q < s.a.size(); ++q) s.a[q] = r.a[q] - alpha * v.a[q];
    if (n.preconditioner) apply_preconditioner(st, s, v_pc);
    else v_pc.a = s.a;
    apply_operator(st, v_pc, matvec, true);
    t.a = matvec.a;
    double omega = dot(t, s) / dot(t, t);
    for (size_t q = 0; q < x.a.size(); ++q) x.a[q] += alpha * p.a[q] + omega * s.a[q];
    rold.a = r.a;
    for (size_t q = 0; q < r.a.size(); ++q) r.a[q] = s.a[q] - omega * t.a[q];
    res = rms(sumsq(r));
    st.last_residual = res / b_norm;
    if (res / b_norm <= tol) {
      niter = jb;
      if (n.preconditioner) {
        s.a = x.a;
        apply_preconditioner(st, s, x);
      }
      return PCF_OK;
    }
    double beta = alpha / omega * dot(r, r0) / dot(rold, r0);
    for (size_t q = 0; q < p.a.size(); ++q) p.a[q] = r.a[q] + beta * (p.a[q] - omega * v.a[q]);
  }
  niter = jb;
  return PCF_E_NOCONV;
}

// apply_corrector!.jl:38-66 with solve_poisson!.jl:39-75
int apply_corrector(State& st, double dt, double rf, int& niter) {
  double tolref = 0.0;
  int rc = compute_lhs(st, tolref);
  if (rc) return rc;
  // solve_poisson!
  st.solution.fill(0.0);
  if (dt == 0.0) {
    st.err = "Error in solve_poisson!: dt = 0.0!";
    return PCF_E_DT_ZERO;
  }
  double dtinv = 1.0 / dt;
  compute_operator(st, dt, rf);
  rc = apply_bicgstab(st, tolref, niter);
  if (rc) return rc;
  solution_to_dpip(st, dtinv);
  correct(st, dt, rf);
  return PCF_OK;
}

// solve_poisson!.jl:69-72: dpip = x / (sqrt(pbar^2 / rhobar) dt), then the
// boundaries of dpip (apply_corrector!.jl:58-60)
void solution_to_dpip(State& st, double dtinv) {
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double& sol = st.solution(i - st.i0 + 1, j - st.j0 + 1, k - st.k0 + 1);
        sol /= std::sqrt(st.pbar(i, j, k) * st.pbar(i, j, k) / st.rhobar(i, j, k));
        st.dpip(i, j, k) = dtinv * sol;
      }
  zonal_bc(st, st.dpip);
  meridional_bc(st, st.dpip);
  vertical_bc(st, st.dpip, +1);
}

// correct!.jl:164-385
void correct(State& st, double dt, double rf) {
  const bool damp = st.cfg.damp_horizontal_wind_on_rhs != 0;
  Arr3 &u = st.pr.u, &v = st.pr.v, &w = st.pr.w, &rhop = st.pr.rhop, &rho = st.pr.rho;
  Arr3 &corx = st.corx, &cory = st.cory, &dpip = st.dpip, &betar = st.betar;
  Arr3 &jac = st.jac, &m13 = st.met13, &m23 = st.met23, &n2 = st.n2;
  const double dt2 = dt * dt;
  const int kmax_uv = st.ko + st.nz == st.cfg.z_size ? st.k1 : st.k1 + 1;
  // U (correct!.jl:177-215)
  for (int k = st.k0; k <= kmax_uv; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0 - 1; i <= st.i1; ++i) {
        double factor = 1.0;
        if (damp) factor += dt * 0.5 * (betar(i, j, k) + betar(i + 1, j, k)) * rf;
        double gradient = pressure_gradient_u(st, dpip, i, j, k);
        corx(i, j, k) = dt / factor * gradient;
        double jpedger = wind_factor_u(st, i, j, k);
        u(i, j, k) -= jpedger * corx(i, j, k);
      }
  // V (correct!.jl:217-255)
  for (int k = st.k0; k <= kmax_uv; ++k)
    for (int j = st.j0 - 1; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double factor = 1.0;
        if (damp) factor += dt * 0.5 * (betar(i, j, k) + betar(i, j + 1, k)) * rf;
        double gradient = pressure_gradient_v(st, dpip, i, j, k);
        cory(i, j, k) = dt / factor * gradient;
        double jpedgef = wind_factor_v(st, i, j, k);
        v(i, j, k) -= jpedgef * cory(i, j, k);
      }
  // W (correct!.jl:257-316)
  const int kmin_w = st.ko == 0 ? st.k0 : st.k0 - 1;
  const int kmax_w = st.ko + st.nz == st.cfg.z_size ? st.k1 - 1 : st.k1;
  for (int k = kmin_w; k <= kmax_w; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double factor = 1.0;
        factor += dt * (jac(i, j, k + 1) * betar(i, j, k) + jac(i, j, k) * betar(i, j, k + 1)) /
                  (jac(i, j, k) + jac(i, j, k + 1)) * rf;
        double n2edgeu = (jac(i, j, k + 1) * n2(i, j, k) + jac(i, j, k) * n2(i, j, k + 1)) /
                         (jac(i, j, k) + jac(i, j, k + 1));
        double gradient = pressure_gradient_w(st, dpip, i, j, k);
        double jpedgeu = wind_factor_w(st, i, j, k);
        double fw = buoyancy_factor_w(st, i, j, k);
        w(i, j, k) +=
            -dt / (factor + fw * n2edgeu * dt2) * jpedgeu * gradient -
            1.0 / (factor + fw * n2edgeu * dt2) * fw * n2edgeu * dt2 * jpedgeu * 0.5 *
                (jac(i, j, k + 1) * (m13(i, j, k) * (corx(i, j, k) + corx(i - 1, j, k)) +
                                     m23(i, j, k) * (cory(i, j, k) + cory(i, j - 1, k))) +
                 jac(i, j, k) * (m13(i, j, k + 1) * (corx(i, j, k + 1) + corx(i - 1, j, k + 1)) +
                                 m23(i, j, k + 1) * (cory(i, j, k + 1) + cory(i, j - 1, k + 1)))) /
                (jac(i, j, k) + jac(i, j, k + 1));
      }
  // RhoP (correct!.jl:318-371)
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double factor = 1.0;
        factor += dt * betar(i, j, k) * rf;
        double lower_gradient = pressure_gradient_w(st, dpip, i, j, k - 1);
        double upper_gradient = pressure_gradient_w(st, dpip, i, j, k);
        if (st.ko + k == st.k0) lower_gradient = 0.0;
        else if (st.ko + k == st.cfg.z_size + st.nbz) upper_gradient = 0.0;
        double gradient = 0.5 * (lower_gradient + upper_gradient);
        double fb = buoyancy_factor_rhop(st, i, j, k);
        double db = -1.0 / (factor + fb * n2(i, j, k) * dt2) *
                    (-fb * n2(i, j, k) * dt2 * jac(i, j, k) * gradient +
                     fb * n2(i, j, k) * dt * jac(i, j, k) * factor * 0.5 *
                         (m13(i, j, k) * (corx(i, j, k) + corx(i - 1, j, k)) +
                          m23(i, j, k) * (cory(i, j, k) + cory(i, j - 1, k))));
        rhop(i, j, k) -= (rho(i, j, k) + st.rhobar(i, j, k)) / st.c.g_ndim * db;
      }
  // PiP (correct!.jl:373-385)
  for (int k = st.k0 - 1; k <= st.k1 + 1; ++k)
    for (int j = st.j0 - 1; j <= st.j1 + 1; ++j)
      for (int i = st.i0 - 1; i <= st.i1 + 1; ++i) st.pr.pip(i, j, k) += dpip(i, j, k);
}

}  // namespace orc

extern "C" void orc_set_reduction_order(int block, int backwards) {
  orc::g_order_block = block > 0 ? (size_t)block : 0;
  orc::g_order_backwards = backwards != 0;
}
