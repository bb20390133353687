// One-time host setup of the model state in FP64: Constants, Grid,
// Atmosphere, initial Predictands and the static sponge profiles
// (State.jl:106-131).  The arrays are built on the host once and uploaded;
// from then on every field lives in HBM and only kernels touch it.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "pcf_internal.hpp"

namespace pcf {

namespace {

struct Host3 {
  int nx, ny, nz;
  std::vector<double> a;
  Host3(int x, int y, int z) : nx(x), ny(y), nz(z), a((size_t)x * y * z, 0.0) {}
  double& operator()(int i, int j, int k) {
    return a[(size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1))];
  }
};

// set_zonal/meridional_boundaries_of_field! on a host array (one process).
void periodic_xy(const Geo& g, Host3& f) {
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nbx; ++i) {
        f(g.i0 - i, j, k) = f(g.i1 - i + 1, j, k);
        f(g.i1 + i, j, k) = f(g.i0 + i - 1, j, k);
      }
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nby; ++j)
      for (int i = 1; i <= g.nxx; ++i) {
        f(i, g.j0 - j, k) = f(i, g.j1 - j + 1, k);
        f(i, g.j1 + j, k) = f(i, g.j0 + j - 1, k);
      }
}
void periodic_x(const Geo& g, Host3& f) {
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nbx; ++i) {
        f(g.i0 - i, j, k) = f(g.i1 - i + 1, j, k);
        f(g.i1 + i, j, k) = f(g.i0 + i - 1, j, k);
      }
}
void periodic_y(const Geo& g, Host3& f) {
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nby; ++j)
      for (int i = 1; i <= g.nxx; ++i) {
        f(i, g.j0 - j, k) = f(i, g.j1 - j + 1, k);
        f(i, g.j1 + j, k) = f(i, g.j0 + j - 1, k);
      }
}

// set_vertical_boundaries_of_field! (walls at both ends on one process)
void walls_z(const Geo& g, Host3& f, double sign, bool staggered) {
  for (int j = 1; j <= g.nyy; ++j)
    for (int i = 1; i <= g.nxx; ++i) {
      if (staggered) {
        f(i, j, g.k0 - 1) = 0.0;
        for (int k = 1; k <= g.nbz; ++k) f(i, j, g.k0 - k) = sign * f(i, j, g.k0 + k - 2);
        f(i, j, g.k1) = 0.0;
        for (int k = 1; k <= g.nbz; ++k) f(i, j, g.k1 + k) = sign * f(i, j, g.k1 - k);
      } else {
        for (int k = 1; k <= g.nbz; ++k) f(i, j, g.k0 - k) = sign * f(i, j, g.k0 + k - 1);
        for (int k = 1; k <= g.nbz; ++k) f(i, j, g.k1 + k) = sign * f(i, j, g.k1 - k + 1);
      }
    }
}

// The parametric forms of the reference's function-valued namelist
// entries (GridNamelist.jl, AtmosphereNamelist.jl, SpongeNamelist.jl),
// evaluated dimensionally as the closures are.
double topography(const pcf_config& c, double x, double y) {
  switch (c.topography) {
    case PCF_TOPO_COSINE_HILL:  // periodic_hill.jl:29
      return c.topo_height / 2 * (1 + std::cos(M_PI / c.topo_width * x));
    case PCF_TOPO_AGNESI:  // mountain_wave.jl:34
      return c.topo_height / (1 + (x * x + y * y) / (c.topo_width * c.topo_width));
    case PCF_TOPO_GAUSSIANS: {
      double h = 0.0;
      for (int n = 0; n < c.n_hills; ++n) {
        const pcf_hill& q = c.hills[n];
        double ddx = x - q.x0, ddy = y - q.y0;
        h += q.height * std::exp(-(ddx * ddx + ddy * ddy) / (q.width * q.width));
      }
      return h;
    }
    default:
      return 0.0;
  }
}

double stretching(const pcf_config& c, double zhat) {
  if (c.stretching == PCF_STRETCH_SINH) {
    double b = c.stretching_parameter;
    return c.lz * std::sinh(b * zhat / c.lz) / std::sinh(b);
  }
  return zhat;
}

// initial_tracer of TracerNamelist.jl as parametrised by pcf_initial_tracer
// (DIPOLE: vortex.jl:50-58)
double initial_tracer(const pcf_config& c, double x, double y, double z) {
  (void)z;
  if (c.initial_tracer == PCF_TRACER_DIPOLE) {
    const double rx = c.tracer_rx, ry = c.tracer_ry;
    const double ax = (std::fabs(x) - rx) / rx, ay = y / ry;
    const double r = std::sqrt(ax * ax + ay * ay);
    if (r <= 1) {
      const double sg = x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0);
      return sg * (1 + std::cos(M_PI * r)) / 2;
    }
  }
  return 0.0;
}

double bubble_radius(const pcf_config& c, double x, double y, double z) {
  double r2 = 0.0;
  if (c.pert_rx > 0) { double q = (x - c.pert_x0) / c.pert_rx; r2 += q * q; }
  if (c.pert_ry > 0) { double q = (y - c.pert_y0) / c.pert_ry; r2 += q * q; }
  if (c.pert_rz > 0) { double q = (z - c.pert_z0) / c.pert_rz; r2 += q * q; }
  return std::sqrt(r2);
}

double lhs_sponge(const pcf_config& n, double x, double y, double z) {  // mountain_wave.jl:40-53
  if (n.lhs_sponge != PCF_LHS_SPONGE_SIN2) return 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (n.lhs_dxr > 0 && std::fabs(x) >= (n.lx - n.lhs_dxr) / 2) {
    double s = std::sin(M_PI * (std::fabs(x) - (n.lx - n.lhs_dxr) / 2) / n.lhs_dxr);
    ax = s * s;
  }
  if (n.lhs_dyr > 0 && std::fabs(y) >= (n.ly - n.lhs_dyr) / 2) {
    double s = std::sin(M_PI * (std::fabs(y) - (n.ly - n.lhs_dyr) / 2) / n.lhs_dyr);
    ay = s * s;
  }
  if (n.lhs_dzr > 0 && z >= n.lz - n.lhs_dzr) {
    double s = std::sin(M_PI / 2 * (z - (n.lz - n.lhs_dzr)) / n.lhs_dzr);
    az = s * s;
  }
  return n.lhs_alpharmax * (ax + ay + az) / n.lhs_divisor;
}

// rhs_sponge (periodic_hill.jl:33-36) without its 1/dt factor, which the
// per-step kernel applies.
double rhs_sponge_numerator(const pcf_config& n, double z) {
  if (n.rhs_sponge != PCF_RHS_SPONGE_SIN2_TOP || z < n.rhs_zr) return 0.0;
  double s = std::sin(M_PI / 2 * (z - n.rhs_zr) / (n.lz - n.rhs_zr));
  return n.rhs_amplitude * (s * s);
}

int upload(pcf_state* st, double** dst, const std::vector<double>& src) {
  size_t bytes = src.size() * sizeof(double);
  int rc = device_alloc(st, dst, src.size());
  if (rc) return rc;
  PCF_HIP(hipMemcpy(*dst, src.data(), bytes, hipMemcpyHostToDevice));
  return PCF_OK;
}

}  // namespace

int host_setup(pcf_state* st, std::string& err) {
  const pcf_config& n = st->cfg;
  if (n.nbx < 3 || n.nby < 3 || n.nbz < 3) { err = "Error in Domain: nb < 3!"; return PCF_E_CONFIG; }
  if (n.x_size < 1 || n.y_size < 1 || n.z_size < 1) { err = "Error in Domain: size < 1!"; return PCF_E_CONFIG; }
  if ((n.x_size > 1 && n.nbx > n.x_size) || (n.y_size > 1 && n.nby > n.y_size) ||
      (n.z_size > 1 && n.nbz > n.z_size)) {
    err = "Error in Domain: nb > size!";
    return PCF_E_CONFIG;
  }
  if (n.model != PCF_BOUSSINESQ && n.model != PCF_PSEUDO_INCOMPRESSIBLE && n.model != PCF_COMPRESSIBLE) {
    err = "Invalid model";
    return PCF_E_OPTION;
  }
  if (!(n.thermal_conductivity >= 0.0)) { err = "thermal_conductivity < 0"; return PCF_E_CONFIG; }
  if (n.preconditioner_iterations < 1 && n.preconditioner) { err = "preconditioner_iterations < 1"; return PCF_E_CONFIG; }
  if (n.precision != PCF_PRECISION_FP64 && n.precision != PCF_PRECISION_MIXED) { err = "Invalid precision"; return PCF_E_OPTION; }
  if (n.flux_precision != PCF_PRECISION_FP64 && n.flux_precision != PCF_PRECISION_MIXED) {
    err = "Invalid flux_precision";
    return PCF_E_OPTION;
  }
  if (n.initial_guess != PCF_GUESS_ZERO && n.initial_guess != PCF_GUESS_PREVIOUS && n.initial_guess != PCF_GUESS_LEARNED &&
      n.initial_guess != PCF_GUESS_EXTRAPOLATED) {
    err = "Invalid initial_guess";
    return PCF_E_OPTION;
  }
  if (n.preconditioner_type != PCF_PC_LINE_RELAXATION && n.preconditioner_type != PCF_PC_MULTIGRID) {
    err = "Invalid preconditioner_type";
    return PCF_E_OPTION;
  }
  if (n.preconditioner_type != PCF_PC_LINE_RELAXATION && (n.precision != PCF_PRECISION_MIXED || !n.preconditioner)) {
    err = "the multigrid preconditioner needs precision = mixed and preconditioner = true";
    return PCF_E_CONFIG;
  }
  if (n.preconditioner_type == PCF_PC_MULTIGRID && (n.mg_smoothing_steps < 1 || n.mg_coarse_steps < 1)) {
    err = "mg_smoothing_steps and mg_coarse_steps must be at least 1";
    return PCF_E_CONFIG;
  }
  if (n.preconditioner_storage != PCF_STORAGE_FP32 && n.preconditioner_storage != PCF_STORAGE_BF16) {
    err = "Invalid preconditioner_storage";
    return PCF_E_OPTION;
  }
  if (n.preconditioner_storage != PCF_STORAGE_FP32 && (n.precision != PCF_PRECISION_MIXED || !n.preconditioner)) {
    err = "preconditioner_storage = bf16 needs precision = mixed and preconditioner = true";
    return PCF_E_CONFIG;
  }
  if (n.tracer_setup != PCF_NO_TRACER && n.tracer_setup != PCF_TRACER_ON) { err = "Invalid tracer_setup"; return PCF_E_OPTION; }
  if (n.initial_tracer != PCF_TRACER_ZERO && n.initial_tracer != PCF_TRACER_DIPOLE) {
    err = "Invalid initial_tracer";
    return PCF_E_OPTION;
  }
  if (n.tracer_setup == PCF_TRACER_ON && n.initial_tracer == PCF_TRACER_DIPOLE &&
      !(n.tracer_rx > 0.0 && n.tracer_ry > 0.0)) {
    err = "initial_tracer: the dipole radii must be positive";
    return PCF_E_CONFIG;
  }
  if (n.initial_guess != PCF_GUESS_ZERO && n.precision != PCF_PRECISION_MIXED) {
    err = "initial_guess other than zero needs precision = mixed (the FP64 solve starts from zero)";
    return PCF_E_CONFIG;
  }


  // Constants.jl:108-182
  const double gamma = 1.4, gammainv = 1.0 / gamma, kappa = (gamma - 1.0) / gamma, kappainv = 1.0 / kappa;
  const double rsp = 287.0, gacc = 9.81, rhoref = 1.184, pref = 101325.0;
  const double aref = std::sqrt(pref / rhoref), uref = aref, lref = pref / rhoref / gacc, tref = lref / aref;
  const double thetaref = aref * aref / rsp;
  const double g_ndim = gacc / (uref * uref / lref);
  double re;
  if (n.specify_reynolds_number)
    re = n.inverse_reynolds_number < DBL_EPSILON ? 1 / DBL_EPSILON : 1.0 / n.inverse_reynolds_number;
  else
    re = n.kinematic_viscosity / uref / lref < DBL_EPSILON ? 1 / DBL_EPSILON : uref * lref / n.kinematic_viscosity;
  const double ma = uref / aref, fr = uref / std::sqrt(gacc * lref);
  const double sig = ma * ma / (fr * fr);
  st->tref = tref; st->lref = lref; st->uref = uref; st->rhoref = rhoref; st->thetaref = thetaref;
  Phys& ph = st->ph;
  ph.kappa = kappa; ph.kappainv = kappainv; ph.mainv2 = 1.0 / (ma * ma); ph.ma2 = ma * ma;
  ph.g_ndim = g_ndim; ph.re = re; ph.fc = n.coriolis_frequency * tref;
  ph.tref = tref; ph.lref = lref; ph.uref = uref; ph.model = n.model;
  ph.damp = n.damp_horizontal_wind_on_rhs != 0;
  ph.viscous = !(1 / re <= DBL_EPSILON);
  ph.gamma = gamma;
  ph.rsp = rsp;
  ph.dpdpi_c = 1 / (gamma - 1) * std::pow(rsp / pref, 1 - gamma);
  st->compressible = n.model == PCF_COMPRESSIBLE;
  // Time.jl:31-38
  st->alphark[0] = 0.0; st->alphark[1] = -5.0 / 9.0; st->alphark[2] = -153.0 / 128.0;
  st->betark[0] = 1.0 / 3.0; st->betark[1] = 15.0 / 16.0; st->betark[2] = 8.0 / 15.0;
  st->stepfrac[0] = 1.0 / 3.0; st->stepfrac[1] = 5.0 / 12.0; st->stepfrac[2] = 1.0 / 4.0;

  // Domain.jl (one process)
  Geo& g = st->g;
  g.nbx = n.nbx; g.nby = n.nby; g.nbz = n.nbz;
  g.nx = n.x_size; g.ny = n.y_size; g.nz = n.z_size;
  g.nxx = g.nx + 2 * g.nbx; g.nyy = g.ny + 2 * g.nby; g.nzz = g.nz + 2 * g.nbz;
  g.i0 = g.nbx + 1; g.i1 = g.i0 + g.nx - 1;
  g.j0 = g.nby + 1; g.j1 = g.j0 + g.ny - 1;
  g.k0 = g.nbz + 1; g.k1 = g.k0 + g.nz - 1;
  g.z_size = n.z_size;

  // Grid.jl:173-338
  const double lx = n.lx / lref, ly = n.ly / lref, lz = n.lz / lref;
  const double dx = lx / n.x_size, dy = ly / n.y_size, dz = lz / n.z_size;
  g.dx = dx; g.dy = dy; g.dz = dz;
  std::vector<double> x(g.nxx + 1), y(g.nyy + 1);
  for (int i = 1; i <= g.nxx; ++i) x[i] = -lx / 2 + (i - g.i0) * dx + dx / 2;
  for (int j = 1; j <= g.nyy; ++j) y[j] = -ly / 2 + (j - g.j0) * dy + dy / 2;
  const int nzt = n.z_size + 2 * g.nbz;
  std::vector<double> zt(nzt + 1), zs(nzt + 1, 0.0);
  for (int k = 1; k <= nzt; ++k) {
    double level = ((k - g.k0) * dz + dz / 2) + 0.5 * dz;
    if (level < 0) zt[k] = -stretching(n, -level * lref) / lref;
    else if (level > lz) zt[k] = 2 * lz - stretching(n, (2 * lz - level) * lref) / lref;
    else zt[k] = stretching(n, level * lref) / lref;
  }
  for (int k = g.k0; k < g.k0 + n.z_size - 1; ++k)
    if (zt[k + 1] <= zt[k]) { err = "Error in Grid: Impossible vertical grid stretching!"; return PCF_E_CONFIG; }
  for (int k = 2; k <= nzt; ++k) zs[k] = 0.5 * (zt[k] + zt[k - 1]);
  zs[1] = zt[1] - 0.5 * (zt[2 * g.nbz] - zt[2 * g.nbz - 1]);

  std::vector<double> hbv((size_t)g.nxx * g.nyy, 0.0);
  auto hb = [&](int i, int j) -> double& { return hbv[(size_t)(i - 1) + (size_t)g.nxx * (j - 1)]; };
  for (int j = g.j0; j <= g.j1; ++j)
    for (int i = g.i0; i <= g.i1; ++i) hb(i, j) = topography(n, x[i] * lref, y[j] * lref) / lref;
  for (int j = 1; j <= g.nyy; ++j)
    for (int i = 1; i <= g.nbx; ++i) {
      hb(g.i0 - i, j) = hb(g.i1 - i + 1, j);
      hb(g.i1 + i, j) = hb(g.i0 + i - 1, j);
    }
  for (int j = 1; j <= g.nby; ++j)
    for (int i = 1; i <= g.nxx; ++i) {
      hb(i, g.j0 - j) = hb(i, g.j1 - j + 1);
      hb(i, g.j1 + j) = hb(i, g.j0 + j - 1);
    }

  Host3 jac(g.nxx, g.nyy, g.nzz), m13(g.nxx, g.nyy, g.nzz), m23(g.nxx, g.nyy, g.nzz), m33(g.nxx, g.nyy, g.nzz);
  auto sq = [](double a) { return a * a; };
  for (int k = 2; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nxx; ++i) jac(i, j, k) = (lz - hb(i, j)) / lz * (zt[k] - zt[k - 1]) / dz;
  for (int j = 1; j <= g.nyy; ++j)
    for (int i = 1; i <= g.nxx; ++i) jac(i, j, 1) = jac(i, j, 2 * g.nbz);
  for (int k = 2; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = g.i0; i <= g.i1; ++i)
        m13(i, j, k) = (hb(i + 1, j) - hb(i - 1, j)) / (2.0 * dx) * (zs[k] - lz) / (lz - hb(i, j)) * dz /
                       (zt[k] - zt[k - 1]);
  periodic_x(g, m13);
  for (int j = 1; j <= g.nyy; ++j)
    for (int i = 1; i <= g.nxx; ++i) m13(i, j, 1) = m13(i, j, 2 * g.nbz) * (zs[1] - lz) / (zs[2 * g.nbz] - lz);
  for (int k = 2; k <= g.nzz; ++k)
    for (int j = g.j0; j <= g.j1; ++j)
      for (int i = 1; i <= g.nxx; ++i)
        m23(i, j, k) = (hb(i, j + 1) - hb(i, j - 1)) / (2.0 * dy) * (zs[k] - lz) / (lz - hb(i, j)) * dz /
                       (zt[k] - zt[k - 1]);
  periodic_y(g, m23);
  for (int j = 1; j <= g.nyy; ++j)
    for (int i = 1; i <= g.nxx; ++i) m23(i, j, 1) = m23(i, j, 2 * g.nbz) * (zs[1] - lz) / (zs[2 * g.nbz] - lz);
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = g.j0; j <= g.j1; ++j)
      for (int i = g.i0; i <= g.i1; ++i) {
        double zsk = k == 1 ? zs[1] : zs[k];
        double dzt = k == 1 ? zt[2 * g.nbz] - zt[2 * g.nbz - 1] : zt[k] - zt[k - 1];
        m33(i, j, k) = (sq(lz / (lz - hb(i, j))) +
                        sq((zsk - lz) / (lz - hb(i, j))) *
                            (sq((hb(i + 1, j) - hb(i - 1, j)) / (2.0 * dx)) + sq((hb(i, j + 1) - hb(i, j - 1)) / (2.0 * dy)))) *
                       sq(dz / dzt);
      }
  periodic_xy(g, m33);
  Host3 zc(g.nxx, g.nyy, g.nzz), zct(g.nxx, g.nyy, g.nzz);
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nxx; ++i) {
        zct(i, j, k) = (lz - hb(i, j)) / lz * zt[k] + hb(i, j);
        zc(i, j, k) = (lz - hb(i, j)) / lz * zs[k] + hb(i, j);
      }
  double dzcmin = INFINITY;
  for (int k = 1; k < g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nxx; ++i) dzcmin = std::min(dzcmin, zct(i, j, k + 1) - zct(i, j, k));
  st->dzcmin = dzcmin;

  // Atmosphere.jl:241-488
  Host3 pbar(g.nxx, g.nyy, g.nzz), thb(g.nxx, g.nyy, g.nzz), rhb(g.nxx, g.nyy, g.nzz), n2(g.nxx, g.nyy, g.nzz);
  const size_t N = pbar.a.size();
  bool need_n2 = false;
  if (n.model == PCF_BOUSSINESQ) {
    if (n.background != PCF_NEUTRAL_STRATIFICATION && n.background != PCF_STABLE_STRATIFICATION) {
      err = "Invalid background for the Boussinesq model";
      return PCF_E_OPTION;
    }
    double th = n.potential_temperature / thetaref * 1.0;
    double nn = n.background == PCF_STABLE_STRATIFICATION
                    ? (n.buoyancy_frequency * tref) * (n.buoyancy_frequency * tref) * 1.0 : 0.0;
    for (size_t q = 0; q < N; ++q) { rhb.a[q] = 1.0; thb.a[q] = th; pbar.a[q] = 1.0 * th; n2.a[q] = nn; }
  } else if (n.background == PCF_ISOTHERMAL) {
    double t0 = n.temperature / thetaref, p0 = n.ground_pressure / pref;
    for (size_t q = 0; q < N; ++q) {
      pbar.a[q] = p0 * std::exp(-sig * zc.a[q] / gamma / t0);
      thb.a[q] = t0 * std::exp(kappa * sig / t0 * zc.a[q]);
      rhb.a[q] = pbar.a[q] / thb.a[q];
    }
    need_n2 = true;
  } else if (n.background == PCF_ISENTROPIC) {
    if (n.potential_temperature < kappa * gacc / rsp * n.lz) {
      err = "Potential temperature too low for given configuration";
      return PCF_E_CONFIG;
    }
    double theta0 = n.potential_temperature / thetaref, p0 = n.ground_pressure / pref;
    for (size_t q = 0; q < N; ++q) {
      n2.a[q] = 0.0;
      thb.a[q] = theta0;
      pbar.a[q] = p0 * std::pow(1.0 - kappa * sig / theta0 * zc.a[q], 1.0 / (gamma - 1.0));
      rhb.a[q] = pbar.a[q] / thb.a[q];
    }
  } else if (n.background == PCF_REALISTIC) {
    double p0 = n.ground_pressure / pref, ztrop = n.tropopause_height / lref;
    double theta0 = n.potential_temperature / thetaref;
    double ptrop = p0 * std::pow(1.0 - kappa * sig / theta0 * ztrop, 1.0 / (gamma - 1.0));
    double ttrop = theta0 * std::pow(ptrop / p0, kappa);
    if (n.potential_temperature < kappa * gacc / rsp * n.tropopause_height) {
      err = "Potential temperature in isentropic troposphere too low";
      return PCF_E_CONFIG;
    }
    for (size_t q = 0; q < N; ++q) {
      if (zc.a[q] <= ztrop) {
        thb.a[q] = theta0;
        pbar.a[q] = p0 * std::pow(1.0 - kappa * sig / theta0 * zc.a[q], 1.0 / (gamma - 1.0));
      } else {
        thb.a[q] = theta0 * std::exp(kappa * sig / ttrop * (zc.a[q] - ztrop));
        pbar.a[q] = std::pow(p0, kappa) * std::pow(ptrop, gammainv) * std::exp(-sig * gammainv / ttrop * (zc.a[q] - ztrop));
      }
      rhb.a[q] = pbar.a[q] / thb.a[q];
    }
    need_n2 = true;
  } else if (n.background == PCF_LAPSE_RATES) {
    double gamma_t = n.troposphere_lapse_rate / thetaref * lref, gamma_s = n.stratosphere_lapse_rate / thetaref * lref;
    double p0 = n.ground_pressure / pref, t0 = n.temperature / thetaref;
    double ztrop = n.tropopause_height / lref, ttrop = t0 - gamma_t * ztrop;
    double power_t = 0, power_s = 0, ptrop;
    if (gamma_t != 0.0) { power_t = gacc / rsp / n.troposphere_lapse_rate / gamma; ptrop = p0 * std::pow(ttrop / t0, power_t); }
    else ptrop = p0 * std::exp(-ztrop * sig / gamma / t0);
    if (gamma_s != 0.0) power_s = gacc / rsp / n.stratosphere_lapse_rate / gamma;
    for (size_t q = 0; q < N; ++q) {
      double tbar;
      if (zc.a[q] <= ztrop) {
        tbar = t0 - gamma_t * zc.a[q];
        pbar.a[q] = gamma_t != 0.0 ? p0 * std::pow(tbar / t0, power_t) : p0 * std::exp(-zc.a[q] * sig / gamma / t0);
      } else {
        tbar = ttrop - gamma_s * (zc.a[q] - ztrop);
        pbar.a[q] = gamma_s != 0.0 ? ptrop * std::pow(tbar / ttrop, power_s)
                                   : ptrop * std::exp(-(zc.a[q] - ztrop) * sig / gamma / ttrop);
      }
      thb.a[q] = tbar * std::pow(p0 / pbar.a[q], kappa * gamma);
      rhb.a[q] = pbar.a[q] / thb.a[q];
    }
    need_n2 = true;
  } else {
    err = "Invalid background for the pseudo-incompressible model";
    return PCF_E_OPTION;
  }
  if (need_n2) {  // compute_n2!.jl:43-82
    for (size_t q = 0; q < N; ++q) n2.a[q] = 0.0;
    for (int k = g.k0; k <= g.k1; ++k)
      for (int j = 1; j <= g.nyy; ++j)
        for (int i = 1; i <= g.nxx; ++i)
          n2(i, j, k) = g_ndim / thb(i, j, k) / jac(i, j, k) * 0.5 * (thb(i, j, k + 1) - thb(i, j, k - 1)) / dz;
    walls_z(g, n2, 1.0, false);
    for (int k = 1; k <= g.nbz; ++k)
      for (int j = 1; j <= g.nyy; ++j)
        for (int i = 1; i <= g.nxx; ++i) {
          n2(i, j, k) = g_ndim / thb(i, j, g.k0 - 1) / jac(i, j, g.k0 - 1) * (thb(i, j, g.k0) - thb(i, j, g.k0 - 1)) / dz;
        }
    for (int k = 1; k <= g.nbz; ++k)
      for (int j = 1; j <= g.nyy; ++j)
        for (int i = 1; i <= g.nxx; ++i)
          n2(i, j, g.k1 + k) = g_ndim / thb(i, j, g.k1 + 1) / jac(i, j, g.k1 + 1) *
                               (thb(i, j, g.k1 + 1) - thb(i, j, g.k1)) / dz;
  }
  Host3 fcs(g.nxx, g.nyy, g.nzz);
  for (size_t q = 0; q < N; ++q) fcs.a[q] = std::sqrt(pbar.a[q] * pbar.a[q] / rhb.a[q]);

  // Predictands.jl:70-157
  Host3 rho(g.nxx, g.nyy, g.nzz), rhop(g.nxx, g.nyy, g.nzz), u(g.nxx, g.nyy, g.nzz), v(g.nxx, g.nyy, g.nzz),
      w(g.nxx, g.nyy, g.nzz), pip(g.nxx, g.nyy, g.nzz);
  if (n.buoyancy_initialization != PCF_INIT_RHOP && n.buoyancy_initialization != PCF_INIT_THETAP) {
    err = "Unknown buoyancy-initialization mode!";
    return PCF_E_OPTION;
  }
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = g.j0; j <= g.j1; ++j)
      for (int i = g.i0; i <= g.i1; ++i) {
        double xd = x[i] * lref, yd = y[j] * lref, zd = zc(i, j, k) * lref;
        double pert = 0.0, ui = n.initial_u, vi = n.initial_v;
        if (n.perturbation == PCF_PERT_COSINE_BUBBLE) {
          double r = bubble_radius(n, xd, yd, zd);
          pert = r <= 1 ? n.pert_amplitude * (1 + std::cos(M_PI * r)) : 0.0;
        } else if (n.perturbation == PCF_PERT_VORTEX) {  // vortex.jl:23-39
          double rx = n.pert_rx, ry = n.pert_ry;
          double r = std::sqrt((xd / rx) * (xd / rx) + (yd / ry) * (yd / ry));
          ui = r <= 1 ? -n.pert_amplitude * yd / ry * (1 + std::cos(M_PI * r)) / 2 : 0.0;
          vi = r <= 1 ? n.pert_amplitude * xd / rx * (1 + std::cos(M_PI * r)) / 2 : 0.0;
          ui += n.initial_u;
          vi += n.initial_v;
        }
        if (n.buoyancy_initialization == PCF_INIT_RHOP) rhop(i, j, k) = pert / rhoref;
        else rhop(i, j, k) = rhb(i, j, k) * (1 / (1 + pert / thetaref / thb(i, j, k)) - 1);
        u(i, j, k) = ui / uref;
        v(i, j, k) = vi / uref;
        w(i, j, k) = n.initial_w / uref;
      }
  for (Host3* f : {&rhop, &u, &v, &w, &pip}) periodic_x(g, *f);
  for (Host3* f : {&rhop, &u, &v, &w, &pip}) periodic_y(g, *f);
  if (n.model != PCF_BOUSSINESQ) rho.a = rhop.a;
  for (size_t q = 0; q < N; ++q) w.a[q] = m13.a[q] * u.a[q] + m23.a[q] * v.a[q] + w.a[q] / jac.a[q];
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = g.i0; i <= g.i1; ++i) u(i, j, k) = (u(i, j, k) + u(i + 1, j, k)) / 2;
  periodic_x(g, u);
  for (int k = 1; k <= g.nzz; ++k)
    for (int j = g.j0; j <= g.j1; ++j)
      for (int i = 1; i <= g.nxx; ++i) v(i, j, k) = (v(i, j, k) + v(i, j + 1, k)) / 2;
  periodic_y(g, v);
  for (int k = g.k0; k <= g.k1; ++k)
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nxx; ++i)
        w(i, j, k) = (jac(i, j, k + 1) * w(i, j, k) + jac(i, j, k) * w(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
  walls_z(g, w, -1.0, true);

  // TracerPredictands.jl:113-139 (TracerOn): the initial tracer times the
  // density, kept also as the sponge's background (TracerAuxiliaries.jl:18-28)
  st->tracer = n.tracer_setup == PCF_TRACER_ON;
  Host3 chi(st->tracer ? g.nxx : 1, st->tracer ? g.nyy : 1, st->tracer ? g.nzz : 1);
  if (st->tracer) {
    for (int k = 1; k <= g.nzz; ++k)
      for (int j = g.j0; j <= g.j1; ++j)
        for (int i = g.i0; i <= g.i1; ++i)
          chi(i, j, k) = initial_tracer(n, x[i] * lref, y[j] * lref, zc(i, j, k) * lref);
    periodic_x(g, chi);
    periodic_y(g, chi);
    for (size_t q = 0; q < N; ++q) chi.a[q] *= rho.a[q] + rhb.a[q];
  }

  // Sponge profiles (compute_sponges!.jl:26-65): alphar does not depend on
  // time or dt for the supported forms; betar = (numerator / dt) tref.
  Host3 alphar(g.nxx, g.nyy, g.nzz), bnum(g.nxx, g.nyy, g.nzz);
  for (int k = g.k0; k <= g.k1; ++k)
    for (int j = g.j0; j <= g.j1; ++j)
      for (int i = g.i0; i <= g.i1; ++i) {
        double xd = x[i] * lref, yd = y[j] * lref, zd = zc(i, j, k) * lref;
        alphar(i, j, k) = lhs_sponge(n, xd, yd, zd) * tref;
        bnum(i, j, k) = rhs_sponge_numerator(n, zd);
      }
  for (Host3* f : {&alphar, &bnum}) {
    periodic_x(g, *f);
  }
  for (Host3* f : {&alphar, &bnum}) {
    periodic_y(g, *f);
    for (int j = 1; j <= g.nyy; ++j)
      for (int i = 1; i <= g.nxx; ++i) {
        (*f)(i, j, g.k0 - 1) = (*f)(i, j, g.k0);
        (*f)(i, j, g.k1 + 1) = (*f)(i, j, g.k1);
      }
  }

  st->deta = n.dtau / (2 * (1 / (dx * dx) + 1 / (dy * dy)));

  int rc;
  if ((rc = upload(st, &st->jac, jac.a))) return rc;
  if ((rc = upload(st, &st->met13, m13.a))) return rc;
  if ((rc = upload(st, &st->met23, m23.a))) return rc;
  if ((rc = upload(st, &st->met33, m33.a))) return rc;
  if ((rc = upload(st, &st->zc, zc.a))) return rc;
  if ((rc = upload(st, &st->zctilde, zct.a))) return rc;
  if ((rc = upload(st, &st->pbar, pbar.a))) return rc;
  if ((rc = upload(st, &st->thetabar, thb.a))) return rc;
  if ((rc = upload(st, &st->rhobar, rhb.a))) return rc;
  if ((rc = upload(st, &st->n2, n2.a))) return rc;
  if ((rc = upload(st, &st->fcscal, fcs.a))) return rc;
  if ((rc = upload(st, &st->rho, rho.a))) return rc;
  if ((rc = upload(st, &st->rhop, rhop.a))) return rc;
  if ((rc = upload(st, &st->u, u.a))) return rc;
  if ((rc = upload(st, &st->v, v.a))) return rc;
  if ((rc = upload(st, &st->w, w.a))) return rc;
  if ((rc = upload(st, &st->pip, pip.a))) return rc;
  if (st->compressible && (rc = upload(st, &st->p, pbar.a))) return rc;  // set_p.jl:36-41
  if (st->tracer) {
    if ((rc = upload(st, &st->chi, chi.a))) return rc;
    if ((rc = upload(st, &st->chi_init, chi.a))) return rc;
  }
  if ((rc = upload(st, &st->alphar, alphar.a))) return rc;
  if ((rc = upload(st, &st->betar_s, bnum.a))) return rc;
  return PCF_OK;
}

}  // namespace pcf
