// ORACLE -- test infrastructure only (see orc.hpp).
// State construction: Constants, Time, Domain, Grid, Atmosphere, Predictands
// (src/Types/State.jl:106-131 and the files cited per function).
#include <algorithm>
#include <cfloat>
#include <cstring>

#include "orc.hpp"

namespace orc {

static const double kEps = DBL_EPSILON;  // Julia eps()

// Closed set of the reference's function-valued namelist entries.
// resolved_topography (GridNamelist.jl:35-45); forms of periodic_hill.jl:29,
// mountain_wave.jl:34 and sums of Gaussian hills.
double hb_dim(const pcf_config& cfg, double x, double y) {
  switch (cfg.topography) {
    case PCF_TOPO_COSINE_HILL:
      return cfg.topo_height / 2 * (1 + std::cos(M_PI / cfg.topo_width * x));
    case PCF_TOPO_AGNESI:
      return cfg.topo_height / (1 + (x * x + y * y) / (cfg.topo_width * cfg.topo_width));
    case PCF_TOPO_GAUSSIANS: {
      double h = 0.0;
      for (int n = 0; n < cfg.n_hills; ++n) {
        const pcf_hill& q = cfg.hills[n];
        double ddx = x - q.x0, ddy = y - q.y0;
        h += q.height * std::exp(-(ddx * ddx + ddy * ddy) / (q.width * q.width));
      }
      return h;
    }
    default:
      return 0.0;
  }
}

// vertical_grid_stretching (GridNamelist.jl:35-45), dimensional in and out.
double stretch_dim(const pcf_config& cfg, double zhat) {
  if (cfg.stretching == PCF_STRETCH_SINH) {
    double b = cfg.stretching_parameter;
    return cfg.lz * std::sinh(b * zhat / cfg.lz) / std::sinh(b);
  }
  return zhat;
}

// Radius of the cosine bubble; a direction with radius <= 0 is dropped.
static double bubble_r(const pcf_config& cfg, double x, double y, double z) {
  double r2 = 0.0;
  if (cfg.pert_rx > 0) { double q = (x - cfg.pert_x0) / cfg.pert_rx; r2 += q * q; }
  if (cfg.pert_ry > 0) { double q = (y - cfg.pert_y0) / cfg.pert_ry; r2 += q * q; }
  if (cfg.pert_rz > 0) { double q = (z - cfg.pert_z0) / cfg.pert_rz; r2 += q * q; }
  return std::sqrt(r2);
}

// Constants.jl:108-182
static void make_constants(State& st) {
  Constants& c = st.c;
  const pcf_config& n = st.cfg;
  c.gamma = 1.4;
  c.gammainv = 1.0 / c.gamma;
  c.kappa = (c.gamma - 1.0) / c.gamma;
  c.kappainv = 1.0 / c.kappa;
  c.rsp = 287.0;
  c.g = 9.81;
  c.rhoref = 1.184;
  c.pref = 101325.0;
  c.aref = std::sqrt(c.pref / c.rhoref);
  c.uref = c.aref;
  c.lref = c.pref / c.rhoref / c.g;
  c.tref = c.lref / c.aref;
  c.thetaref = c.aref * c.aref / c.rsp;
  c.fref = c.rhoref * c.uref * c.uref / c.lref;
  c.g_ndim = c.g / (c.uref * c.uref / c.lref);
  if (n.specify_reynolds_number) {
    c.re = n.inverse_reynolds_number < kEps ? 1 / kEps : 1.0 / n.inverse_reynolds_number;
  } else {
    c.re = n.kinematic_viscosity / c.uref / c.lref < kEps ? 1 / kEps
                                                           : c.uref * c.lref / n.kinematic_viscosity;
  }
  c.ma = c.uref / c.aref;
  c.mainv2 = 1.0 / (c.ma * c.ma);
  c.ma2 = c.ma * c.ma;
  c.fr = c.uref / std::sqrt(c.g * c.lref);
  c.frinv2 = 1.0 / (c.fr * c.fr);
  c.fr2 = c.fr * c.fr;
  c.sig = c.ma * c.ma / (c.fr * c.fr);
}

// Grid.jl:173-338 and compute_topography.jl:145-176 (NoWKB)
static int make_grid(State& st) {
  const pcf_config& n = st.cfg;
  const double lref = st.c.lref;
  st.lx = n.lx / lref;
  st.ly = n.ly / lref;
  st.lz = n.lz / lref;
  st.dx = st.lx / n.x_size;
  st.dy = st.ly / n.y_size;
  st.dz = st.lz / n.z_size;
  const double lx = st.lx, ly = st.ly, lz = st.lz, dx = st.dx, dy = st.dy, dz = st.dz;
  st.x.assign(st.nxx + 1, 0.0);
  st.y.assign(st.nyy + 1, 0.0);
  for (int i = 1; i <= st.nxx; ++i) st.x[i] = -lx / 2 + (i + st.io - st.i0) * dx + dx / 2;
  for (int j = 1; j <= st.nyy; ++j) st.y[j] = -ly / 2 + (j + st.jo - st.j0) * dy + dy / 2;
  const int nzt = n.z_size + 2 * st.nbz;
  std::vector<double> z(nzt + 1, 0.0);
  for (int k = 1; k <= nzt; ++k) z[k] = (k - st.k0) * dz + dz / 2;
  st.ztildes.assign(nzt + 1, 0.0);
  st.zs.assign(nzt + 1, 0.0);
  for (int k = 1; k <= nzt; ++k) {
    double level = z[k] + 0.5 * dz;
    if (level < 0)
      st.ztildes[k] = -stretch_dim(n, -level * lref) / lref;
    else if (level > lz)
      st.ztildes[k] = 2 * lz - stretch_dim(n, (2 * lz - level) * lref) / lref;
    else
      st.ztildes[k] = stretch_dim(n, level * lref) / lref;
  }
  for (int k = st.k0; k < st.k0 + n.z_size - 1; ++k) {
    if (st.ztildes[k + 1] <= st.ztildes[k]) {
      st.err = "Error in Grid: Impossible vertical grid stretching!";
      return PCF_E_CONFIG;
    }
  }
  for (int k = 2; k <= nzt; ++k) st.zs[k] = 0.5 * (st.ztildes[k] + st.ztildes[k - 1]);
  st.zs[1] = st.ztildes[1] - 0.5 * (st.ztildes[2 * st.nbz] - st.ztildes[2 * st.nbz - 1]);
  const std::vector<double>& zt = st.ztildes;
  const std::vector<double>& zs = st.zs;

  // compute_topography (NoWKB)
  st.hb.assign((size_t)st.nxx * st.nyy, 0.0);
  auto HB = [&](int i, int j) -> double& { return st.hb[(size_t)(i - 1) + (size_t)st.nxx * (j - 1)]; };
  for (int j = st.j0; j <= st.j1; ++j)
    for (int i = st.i0; i <= st.i1; ++i) HB(i, j) = hb_dim(n, st.x[i] * lref, st.y[j] * lref) / lref;
  zonal_bc_2d(st, st.hb);
  meridional_bc_2d(st, st.hb);

  st.jac.init(st.nxx, st.nyy, st.nzz);
  st.met13.init(st.nxx, st.nyy, st.nzz);
  st.met23.init(st.nxx, st.nyy, st.nzz);
  st.met33.init(st.nxx, st.nyy, st.nzz);
  const int kmin = st.ko == 0 ? 2 : 1, kmax = st.nzz;
  for (int k = kmin; k <= kmax; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i)
        st.jac(i, j, k) = (lz - HB(i, j)) / lz * (zt[st.ko + k] - zt[st.ko + k - 1]) / dz;
  if (st.ko == 0)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i) st.jac(i, j, 1) = st.jac(i, j, 2 * st.nbz);

  for (int k = kmin; k <= kmax; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = st.i0; i <= st.i1; ++i)
        st.met13(i, j, k) = (HB(i + 1, j) - HB(i - 1, j)) / (2.0 * dx) * (zs[st.ko + k] - lz) /
                            (lz - HB(i, j)) * dz / (zt[st.ko + k] - zt[st.ko + k - 1]);
  zonal_bc(st, st.met13);
  if (st.ko == 0)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i)
        st.met13(i, j, 1) = st.met13(i, j, 2 * st.nbz) * (zs[1] - lz) / (zs[2 * st.nbz] - lz);

  for (int k = 2; k <= st.nzz; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = 1; i <= st.nxx; ++i)
        st.met23(i, j, k) = (HB(i, j + 1) - HB(i, j - 1)) / (2.0 * dy) * (zs[st.ko + k] - lz) /
                            (lz - HB(i, j)) * dz / (zt[st.ko + k] - zt[st.ko + k - 1]);
  meridional_bc(st, st.met23);
  if (st.ko == 0)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i)
        st.met23(i, j, 1) = st.met23(i, j, 2 * st.nbz) * (zs[1] - lz) / (zs[2 * st.nbz] - lz);

  auto sq = [](double a) { return a * a; };
  for (int k = kmin; k <= kmax; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i)
        st.met33(i, j, k) =
            (sq(lz / (lz - HB(i, j))) +
             sq((zs[st.ko + k] - lz) / (lz - HB(i, j))) *
                 (sq((HB(i + 1, j) - HB(i - 1, j)) / (2.0 * dx)) +
                  sq((HB(i, j + 1) - HB(i, j - 1)) / (2.0 * dy)))) *
            sq(dz / (zt[st.ko + k] - zt[st.ko + k - 1]));
  if (st.ko == 0)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i)
        st.met33(i, j, 1) =
            (sq(lz / (lz - HB(i, j))) +
             sq((zs[1] - lz) / (lz - HB(i, j))) *
                 (sq((HB(i + 1, j) - HB(i - 1, j)) / (2.0 * dx)) +
                  sq((HB(i, j + 1) - HB(i, j - 1)) / (2.0 * dy)))) *
            sq(dz / (zt[2 * st.nbz] - zt[2 * st.nbz - 1]));
  zonal_bc(st, st.met33);
  meridional_bc(st, st.met33);

  st.zctilde.init(st.nxx, st.nyy, st.nzz);
  st.zc.init(st.nxx, st.nyy, st.nzz);
  for (int k = 1; k <= st.nzz; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i) {
        st.zctilde(i, j, k) = (lz - HB(i, j)) / lz * zt[st.ko + k] + HB(i, j);
        st.zc(i, j, k) = (lz - HB(i, j)) / lz * zs[st.ko + k] + HB(i, j);
      }
  double m = INFINITY;
  for (int k = 1; k < st.nzz; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i) m = std::min(m, st.zctilde(i, j, k + 1) - st.zctilde(i, j, k));
  st.dzcmin = m;
  return PCF_OK;
}

// compute_n2!.jl:43-82
static void compute_n2(State& st) {
  Arr3& n2 = st.n2;
  n2.fill(0.0);
  const double g = st.c.g_ndim, dz = st.dz;
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i)
        n2(i, j, k) = g / st.thetabar(i, j, k) / st.jac(i, j, k) * 0.5 *
                      (st.thetabar(i, j, k + 1) - st.thetabar(i, j, k - 1)) / dz;
  vertical_bc(st, n2, +1, false);
  if (st.ko == 0)
    for (int k = 1; k <= st.nbz; ++k)
      for (int j = 1; j <= st.nyy; ++j)
        for (int i = 1; i <= st.nxx; ++i)
          n2(i, j, k) = g / st.thetabar(i, j, st.k0 - 1) / st.jac(i, j, st.k0 - 1) *
                        (st.thetabar(i, j, st.k0) - st.thetabar(i, j, st.k0 - 1)) / dz;
  if (st.ko + st.nz == st.cfg.z_size)
    for (int k = 1; k <= st.nbz; ++k)
      for (int j = 1; j <= st.nyy; ++j)
        for (int i = 1; i <= st.nxx; ++i)
          n2(i, j, st.k1 + k) = g / st.thetabar(i, j, st.k1 + 1) / st.jac(i, j, st.k1 + 1) *
                                (st.thetabar(i, j, st.k1 + 1) - st.thetabar(i, j, st.k1)) / dz;
}

// Atmosphere.jl:241-488
static int make_atmosphere(State& st) {
  const pcf_config& n = st.cfg;
  const Constants& c = st.c;
  st.pbar.init(st.nxx, st.nyy, st.nzz);
  st.thetabar.init(st.nxx, st.nyy, st.nzz);
  st.rhobar.init(st.nxx, st.nyy, st.nzz);
  st.n2.init(st.nxx, st.nyy, st.nzz);
  const size_t N = st.pbar.a.size();
  if (n.model == PCF_BOUSSINESQ) {
    if (n.background != PCF_NEUTRAL_STRATIFICATION && n.background != PCF_STABLE_STRATIFICATION) {
      st.err = "Invalid background for the Boussinesq model";
      return PCF_E_OPTION;
    }
    double th = n.potential_temperature / c.thetaref * 1.0;
    double nn = n.background == PCF_STABLE_STRATIFICATION
                    ? (n.buoyancy_frequency * c.tref) * (n.buoyancy_frequency * c.tref) * 1.0
                    : 0.0;
    for (size_t q = 0; q < N; ++q) {
      st.rhobar.a[q] = 1.0;
      st.thetabar.a[q] = th;
      st.pbar.a[q] = 1.0 * th;
      st.n2.a[q] = nn;
    }
    return PCF_OK;
  }
  if (n.model != PCF_PSEUDO_INCOMPRESSIBLE && n.model != PCF_COMPRESSIBLE) {
    st.err = "Invalid model";
    return PCF_E_OPTION;
  }
  const std::vector<double>& zc = st.zc.a;
  switch (n.background) {
    case PCF_ISOTHERMAL: {
      double t0 = n.temperature / c.thetaref, p0 = n.ground_pressure / c.pref;
      for (size_t q = 0; q < N; ++q) {
        st.pbar.a[q] = p0 * std::exp(-c.sig * zc[q] / c.gamma / t0);
        st.thetabar.a[q] = t0 * std::exp(c.kappa * c.sig / t0 * zc[q]);
        st.rhobar.a[q] = st.pbar.a[q] / st.thetabar.a[q];
      }
      compute_n2(st);
      return PCF_OK;
    }
    case PCF_ISENTROPIC: {
      double minpt = c.kappa * c.g / c.rsp * n.lz;
      if (n.potential_temperature < minpt) {
        st.err = "Potential temperature too low for given configuration";
        return PCF_E_CONFIG;
      }
      double theta0 = n.potential_temperature / c.thetaref, p0 = n.ground_pressure / c.pref;
      for (size_t q = 0; q < N; ++q) {
        st.n2.a[q] = 0.0;
        st.thetabar.a[q] = theta0;
        st.pbar.a[q] = p0 * std::pow(1.0 - c.kappa * c.sig / theta0 * zc[q], 1.0 / (c.gamma - 1.0));
        st.rhobar.a[q] = st.pbar.a[q] / st.thetabar.a[q];
      }
      return PCF_OK;
    }
    case PCF_REALISTIC: {
      double p0 = n.ground_pressure / c.pref, ztrop = n.tropopause_height / c.lref;
      double theta0 = n.potential_temperature / c.thetaref;
      double ptrop = p0 * std::pow(1.0 - c.kappa * c.sig / theta0 * ztrop, 1.0 / (c.gamma - 1.0));
      double ttrop = theta0 * std::pow(ptrop / p0, c.kappa);
      double minpt = c.kappa * c.g / c.rsp * n.tropopause_height;
      if (n.potential_temperature < minpt) {
        st.err = "Potential temperature in isentropic troposphere too low";
        return PCF_E_CONFIG;
      }
      for (size_t q = 0; q < N; ++q) {
        if (zc[q] <= ztrop) {
          st.thetabar.a[q] = theta0;
          st.pbar.a[q] = p0 * std::pow(1.0 - c.kappa * c.sig / theta0 * zc[q], 1.0 / (c.gamma - 1.0));
        } else {
          st.thetabar.a[q] = theta0 * std::exp(c.kappa * c.sig / ttrop * (zc[q] - ztrop));
          st.pbar.a[q] = std::pow(p0, c.kappa) * std::pow(ptrop, c.gammainv) *
                         std::exp(-c.sig * c.gammainv / ttrop * (zc[q] - ztrop));
        }
        st.rhobar.a[q] = st.pbar.a[q] / st.thetabar.a[q];
      }
      compute_n2(st);
      return PCF_OK;
    }
    case PCF_LAPSE_RATES: {
      double gamma_t = n.troposphere_lapse_rate / c.thetaref * c.lref;
      double gamma_s = n.stratosphere_lapse_rate / c.thetaref * c.lref;
      double p0 = n.ground_pressure / c.pref, t0 = n.temperature / c.thetaref;
      double ztrop = n.tropopause_height / c.lref, ttrop = t0 - gamma_t * ztrop;
      double power_t = 0, power_s = 0, ptrop;
      if (gamma_t != 0.0) {
        power_t = c.g / c.rsp / n.troposphere_lapse_rate / c.gamma;
        ptrop = p0 * std::pow(ttrop / t0, power_t);
      } else {
        ptrop = p0 * std::exp(-ztrop * c.sig / c.gamma / t0);
      }
      if (gamma_s != 0.0) power_s = c.g / c.rsp / n.stratosphere_lapse_rate / c.gamma;
      for (size_t q = 0; q < N; ++q) {
        double tbar;
        if (zc[q] <= ztrop) {
          tbar = t0 - gamma_t * zc[q];
          st.pbar.a[q] = gamma_t != 0.0 ? p0 * std::pow(tbar / t0, power_t)
                                        : p0 * std::exp(-zc[q] * c.sig / c.gamma / t0);
        } else {
          tbar = ttrop - gamma_s * (zc[q] - ztrop);
          st.pbar.a[q] = gamma_s != 0.0 ? ptrop * std::pow(tbar / ttrop, power_s)
                                        : ptrop * std::exp(-(zc[q] - ztrop) * c.sig / c.gamma / ttrop);
        }
        st.thetabar.a[q] = tbar * std::pow(p0 / st.pbar.a[q], c.kappa * c.gamma);
        st.rhobar.a[q] = st.pbar.a[q] / st.thetabar.a[q];
      }
      compute_n2(st);
      return PCF_OK;
    }
    default:
      st.err = "Invalid background for the pseudo-incompressible model";
      return PCF_E_OPTION;
  }
}

// initial_tracer of TracerNamelist.jl as parametrised by pcf_initial_tracer
// (DIPOLE: vortex.jl:50-58)
static double initial_tracer(const pcf_config& n, double x, double y, double z) {
  (void)z;
  if (n.initial_tracer == PCF_TRACER_DIPOLE) {
    const double rx = n.tracer_rx, ry = n.tracer_ry;
    const double ax = (std::fabs(x) - rx) / rx, ay = y / ry;
    const double r = std::sqrt(ax * ax + ay * ay);
    if (r <= 1) {
      const double sg = x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0);
      return sg * (1 + std::cos(M_PI * r)) / 2;
    }
  }
  return 0.0;
}

// TracerPredictands.jl:113-139 (TracerOn) and TracerAuxiliaries.jl:18-28
static void make_tracer(State& st) {
  const pcf_config& n = st.cfg;
  st.tracer = n.tracer_setup == PCF_TRACER_ON;
  if (!st.tracer) return;
  const double lref = st.c.lref;
  st.chi.init(st.nxx, st.nyy, st.nzz);
  for (int k = 1; k <= st.nzz; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i)
        st.chi(i, j, k) = initial_tracer(n, st.x[i] * lref, st.y[j] * lref, st.zc(i, j, k) * lref);
  zonal_bc(st, st.chi);
  meridional_bc(st, st.chi);
  for (size_t q = 0; q < st.chi.a.size(); ++q) st.chi.a[q] *= st.pr.rho.a[q] + st.rhobar.a[q];
  st.initialtracer = st.chi;
  st.dchi.init(st.nxx, st.nyy, st.nzz);
  st.chi0.init(st.nxx, st.nyy, st.nzz);
  st.chitilde.init(st.nxx, st.nyy, st.nzz);
  st.phichi.init(st.nxx, st.nyy, st.nzz);
}

// Predictands.jl:70-157
static int make_predictands(State& st) {
  const pcf_config& n = st.cfg;
  const Constants& c = st.c;
  Predictands& p = st.pr;
  for (Arr3* a : {&p.rho, &p.rhop, &p.u, &p.v, &p.w, &p.pip}) a->init(st.nxx, st.nyy, st.nzz);
  for (int k = 1; k <= st.nzz; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double xd = st.x[i] * c.lref, yd = st.y[j] * c.lref, zd = st.zc(i, j, k) * c.lref;
        double pert = 0.0, ui = n.initial_u, vi = n.initial_v;
        if (n.perturbation == PCF_PERT_COSINE_BUBBLE) {
          double r = bubble_r(n, xd, yd, zd);
          pert = r <= 1 ? n.pert_amplitude * (1 + std::cos(M_PI * r)) : 0.0;
        } else if (n.perturbation == PCF_PERT_VORTEX) {
          // vortex.jl:23-39 (winds replace the uniform parts inside r <= 1)
          double rx = n.pert_rx, ry = n.pert_ry;
          double r = std::sqrt((xd / rx) * (xd / rx) + (yd / ry) * (yd / ry));
          ui = r <= 1 ? -n.pert_amplitude * yd / ry * (1 + std::cos(M_PI * r)) / 2 : 0.0;
          vi = r <= 1 ? n.pert_amplitude * xd / rx * (1 + std::cos(M_PI * r)) / 2 : 0.0;
          ui += n.initial_u;
          vi += n.initial_v;
        }
        if (n.buoyancy_initialization == PCF_INIT_RHOP) {
          p.rhop(i, j, k) = pert / c.rhoref;
        } else if (n.buoyancy_initialization == PCF_INIT_THETAP) {
          p.rhop(i, j, k) = st.rhobar(i, j, k) *
                            (1 / (1 + pert / c.thetaref / st.thetabar(i, j, k)) - 1);
        } else {
          st.err = "Unknown buoyancy-initialization mode!";
          return PCF_E_OPTION;
        }
        p.u(i, j, k) = ui / c.uref;
        p.v(i, j, k) = vi / c.uref;
        p.w(i, j, k) = n.initial_w / c.uref;
        p.pip(i, j, k) = 0.0;
      }
  for (Arr3* a : {&p.rhop, &p.u, &p.v, &p.w, &p.pip}) zonal_bc(st, *a);
  for (Arr3* a : {&p.rhop, &p.u, &p.v, &p.w, &p.pip}) meridional_bc(st, *a);
  if (n.model != PCF_BOUSSINESQ) p.rho.a = p.rhop.a;
  for (size_t q = 0; q < p.w.a.size(); ++q)
    p.w.a[q] = st.met13.a[q] * p.u.a[q] + st.met23.a[q] * p.v.a[q] + p.w.a[q] / st.jac.a[q];
  for (int k = 1; k <= st.nzz; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = st.i0; i <= st.i1; ++i) p.u(i, j, k) = (p.u(i, j, k) + p.u(i + 1, j, k)) / 2;
  zonal_bc(st, p.u);
  for (int k = 1; k <= st.nzz; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = 1; i <= st.nxx; ++i) p.v(i, j, k) = (p.v(i, j, k) + p.v(i, j + 1, k)) / 2;
  meridional_bc(st, p.v);
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = 1; j <= st.nyy; ++j)
      for (int i = 1; i <= st.nxx; ++i)
        p.w(i, j, k) = (st.jac(i, j, k + 1) * p.w(i, j, k) + st.jac(i, j, k) * p.w(i, j, k + 1)) /
                       (st.jac(i, j, k) + st.jac(i, j, k + 1));
  vertical_bc(st, p.w, -1, true);
  if (n.model == PCF_COMPRESSIBLE) p.p = st.pbar;  // set_p.jl:36-41
  return PCF_OK;
}

int setup(State& st, const pcf_config& cfg) {
  st.cfg = cfg;
  const pcf_config& n = st.cfg;
  if (n.nbx < 3 || n.nby < 3 || n.nbz < 3) {  // Domain.jl:146-154
    st.err = "Error in Domain: nb < 3!";
    return PCF_E_CONFIG;
  }
  if ((n.x_size > 1 && n.nbx > n.x_size) || (n.y_size > 1 && n.nby > n.y_size) ||
      (n.z_size > 1 && n.nbz > n.z_size)) {  // Domain.jl:160-168
    st.err = "Error in Domain: nb > size!";
    return PCF_E_CONFIG;
  }
  if (n.x_size < 1 || n.y_size < 1 || n.z_size < 1) {
    st.err = "Error in Domain: size < 1!";
    return PCF_E_CONFIG;
  }
  make_constants(st);
  st.compressible = n.model == PCF_COMPRESSIBLE;
  // Time.jl:31-38
  st.alphark[0] = 0.0; st.alphark[1] = -5.0 / 9.0; st.alphark[2] = -153.0 / 128.0;
  st.betark[0] = 1.0 / 3.0; st.betark[1] = 15.0 / 16.0; st.betark[2] = 8.0 / 15.0;
  st.stepfrac[0] = 1.0 / 3.0; st.stepfrac[1] = 5.0 / 12.0; st.stepfrac[2] = 1.0 / 4.0;
  // Domain.jl:178-210 (one process)
  st.nbx = n.nbx; st.nby = n.nby; st.nbz = n.nbz;
  st.nx = n.x_size; st.ny = n.y_size; st.nz = n.z_size;
  st.nxx = st.nx + 2 * st.nbx; st.nyy = st.ny + 2 * st.nby; st.nzz = st.nz + 2 * st.nbz;
  st.i0 = st.nbx + 1; st.i1 = st.i0 + st.nx - 1;
  st.j0 = st.nby + 1; st.j1 = st.j0 + st.ny - 1;
  st.k0 = st.nbz + 1; st.k1 = st.k0 + st.nz - 1;
  int rc = make_grid(st);
  if (rc) return rc;
  rc = make_atmosphere(st);
  if (rc) return rc;
  st.alphar.init(st.nxx, st.nyy, st.nzz);
  st.betar.init(st.nxx, st.nyy, st.nzz);
  st.horizontal_mean.assign(st.nz, 0.0);
  rc = make_predictands(st);
  if (rc) return rc;
  make_tracer(st);
  for (Arr3* a : {&st.drho, &st.drhop, &st.du, &st.dv, &st.dw, &st.dpip, &st.rhoold, &st.rhopold,
                  &st.uold, &st.vold, &st.wold, &st.phi, &st.s_op, &st.corx, &st.cory})
    a->init(st.nxx, st.nyy, st.nzz);
  for (Arr5* a : {&st.rhotilde, &st.rhoptilde, &st.utilde, &st.vtilde, &st.wtilde})
    a->init(st.nxx, st.nyy, st.nzz);
  for (Arr4* a : {&st.phirho, &st.phirhop, &st.phiu, &st.phiv, &st.phiw}) a->init(st.nxx, st.nyy, st.nzz);
  for (Arr3* a : {&st.rhs, &st.solution, &st.s_pc, &st.q_pc, &st.bp, &st.br0, &st.brold, &st.br,
                  &st.bs, &st.bt, &st.bv, &st.bmatvec, &st.bv_pc})
    a->init(st.nx, st.ny, st.nz);
  for (auto& t : st.tensor) t.init(st.nx, st.ny, st.nz);
  st.p_pc.assign((size_t)st.nx * st.ny, 0.0);
  for (Predictands* q : {&st.p0, &st.p1})
    for (Arr3* a : {&q->rho, &q->rhop, &q->u, &q->v, &q->w, &q->pip}) a->init(st.nxx, st.nyy, st.nzz);
  if (st.compressible) {  // Increments.jl:91-94, Fluxes.jl:89-92
    st.dp.init(st.nxx, st.nyy, st.nzz);
    st.phip.init(st.nxx, st.nyy, st.nzz);
    st.phitheta.init(st.nxx, st.nyy, st.nzz);
    st.p0.p.init(st.nxx, st.nyy, st.nzz);
    st.p1.p.init(st.nxx, st.nyy, st.nzz);
  }
  return PCF_OK;
}

}  // namespace orc
