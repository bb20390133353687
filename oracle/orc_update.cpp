// ORACLE -- test infrastructure only (see orc.hpp).
// src/Update/*.jl for the Boussinesq, pseudo-incompressible and compressible
// models with wkb_mode = :NoWKB (compute_volume_force = 0 for the winds and
// the conductive heating for P).
#include "orc.hpp"

namespace orc {

// compute_compressible_wind_factor.jl:102-157: 1 outside the compressible
// model, else J P interpolated to the staggered wind points.
double wind_factor_u(State& st, int i, int j, int k) {
  if (!st.compressible) return 1.0;
  Arr3 &jac = st.jac, &p = st.pr.p;
  return (jac(i, j, k) * p(i, j, k) + jac(i + 1, j, k) * p(i + 1, j, k)) / 2;
}
double wind_factor_v(State& st, int i, int j, int k) {
  if (!st.compressible) return 1.0;
  Arr3 &jac = st.jac, &p = st.pr.p;
  return (jac(i, j, k) * p(i, j, k) + jac(i, j + 1, k) * p(i, j + 1, k)) / 2;
}
double wind_factor_w(State& st, int i, int j, int k) {
  if (!st.compressible) return 1.0;
  Arr3 &jac = st.jac, &p = st.pr.p;
  return jac(i, j, k) * jac(i, j, k + 1) * (p(i, j, k) + p(i, j, k + 1)) / (jac(i, j, k) + jac(i, j, k + 1));
}

// conductive_heating.jl:67-117 (0 outside the compressible model)
double conductive_heating(State& st, int i, int j, int k) {
  if (!st.compressible) return 0.0;
  Arr4& f = st.phitheta;
  double rhotot = (st.pr.rho(i, j, k) + st.rhobar(i, j, k)) / st.jac(i, j, k);
  return -rhotot * ((f(i, j, k, 1) - f(i - 1, j, k, 1)) / st.dx + (f(i, j, k, 2) - f(i, j - 1, k, 2)) / st.dy +
                    (f(i, j, k, 3) - f(i, j, k - 1, 3)) / st.dz);
}

// dP/dpi of update!.jl:1385-1386, apply_lhs_sponge!.jl:556-557 and
// compute_operator!.jl:855-858, with Constants.jl's dimensional gamma, rsp, pref
double dpdpi(State& st, double p) {
  const double gamma = st.c.gamma;
  return 1 / (gamma - 1) * std::pow(st.c.rsp / st.c.pref, 1 - gamma) * std::pow(p, 2 - gamma);
}

// compute_buoyancy_factor.jl:108-133 (RhoP)
double buoyancy_factor_rhop(State& st, int i, int j, int k) {
  if (st.compressible)
    return st.pbar(i, j, k) / st.thetabar(i, j, k) / (st.pr.rho(i, j, k) + st.rhobar(i, j, k));
  return st.rhobar(i, j, k) / (st.pr.rho(i, j, k) + st.rhobar(i, j, k));
}

// compute_buoyancy_factor.jl:135-172 (W)
double buoyancy_factor_w(State& st, int i, int j, int k) {
  Arr3 &jac = st.jac, &rb = st.rhobar, &rho = st.pr.rho;
  if (st.compressible) {
    Arr3 &pb = st.pbar, &tb = st.thetabar;
    return (jac(i, j, k + 1) * pb(i, j, k) / tb(i, j, k) + jac(i, j, k) * pb(i, j, k + 1) / tb(i, j, k + 1)) /
           (jac(i, j, k + 1) * (rho(i, j, k) + rb(i, j, k)) + jac(i, j, k) * (rho(i, j, k + 1) + rb(i, j, k + 1)));
  }
  return (jac(i, j, k + 1) * rb(i, j, k) + jac(i, j, k) * rb(i, j, k + 1)) /
         (jac(i, j, k + 1) * (rho(i, j, k) + rb(i, j, k)) + jac(i, j, k) * (rho(i, j, k + 1) + rb(i, j, k + 1)));
}

// compute_pressure_gradient.jl:106-156 (U)
double pressure_gradient_u(State& st, Arr3& pip, int i, int j, int k) {
  const double kinv = st.c.kappainv, mainv2 = st.c.mainv2, dx = st.dx, dz = st.dz;
  double rhoedger = 0.5 * (st.pr.rho(i, j, k) + st.pr.rho(i + 1, j, k));
  double rhobaredger = 0.5 * (st.rhobar(i, j, k) + st.rhobar(i + 1, j, k));
  rhoedger += rhobaredger;
  double pedger = 0.5 * (st.pbar(i, j, k) + st.pbar(i + 1, j, k));
  double met13edger = 0.5 * (st.met13(i, j, k) + st.met13(i + 1, j, k));
  if (st.ko + k == st.k0) {
    double pipuuedger = 0.5 * (pip(i, j, k + 2) + pip(i + 1, j, k + 2));
    double pipuedger = 0.5 * (pip(i, j, k + 1) + pip(i + 1, j, k + 1));
    double pipedger = 0.5 * (pip(i, j, k) + pip(i + 1, j, k));
    return kinv * mainv2 * pedger / rhoedger *
           ((pip(i + 1, j, k) - pip(i, j, k)) / dx +
            met13edger * (-pipuuedger + 4.0 * pipuedger - 3.0 * pipedger) * 0.5 / dz);
  } else if (st.ko + k == st.cfg.z_size + st.nbz) {
    double pipddedger = 0.5 * (pip(i, j, k - 2) + pip(i + 1, j, k - 2));
    double pipdedger = 0.5 * (pip(i, j, k - 1) + pip(i + 1, j, k - 1));
    double pipedger = 0.5 * (pip(i, j, k) + pip(i + 1, j, k));
    return kinv * mainv2 * pedger / rhoedger *
           ((pip(i + 1, j, k) - pip(i, j, k)) / dx +
            met13edger * (pipddedger - 4.0 * pipdedger + 3.0 * pipedger) * 0.5 / dz);
  }
  double pipuedger = 0.5 * (pip(i, j, k + 1) + pip(i + 1, j, k + 1));
  double pipdedger = 0.5 * (pip(i, j, k - 1) + pip(i + 1, j, k - 1));
  return kinv * mainv2 * pedger / rhoedger *
         ((pip(i + 1, j, k) - pip(i, j, k)) / dx + met13edger * (pipuedger - pipdedger) * 0.5 / dz);
}

// compute_pressure_gradient.jl:158-209 (V)
double pressure_gradient_v(State& st, Arr3& pip, int i, int j, int k) {
  const double kinv = st.c.kappainv, mainv2 = st.c.mainv2, dy = st.dy, dz = st.dz;
  double rhoedgef = 0.5 * (st.pr.rho(i, j, k) + st.pr.rho(i, j + 1, k));
  double rhobaredgef = 0.5 * (st.rhobar(i, j, k) + st.rhobar(i, j + 1, k));
  rhoedgef += rhobaredgef;
  double pedgef = 0.5 * (st.pbar(i, j, k) + st.pbar(i, j + 1, k));
  double met23edgef = 0.5 * (st.met23(i, j, k) + st.met23(i, j + 1, k));
  if (st.ko + k == st.k0) {
    double pipuuedgef = 0.5 * (pip(i, j, k + 2) + pip(i, j + 1, k + 2));
    double pipuedgef = 0.5 * (pip(i, j, k + 1) + pip(i, j + 1, k + 1));
    double pipedgef = 0.5 * (pip(i, j, k) + pip(i, j + 1, k));
    return kinv * mainv2 * pedgef / rhoedgef *
           ((pip(i, j + 1, k) - pip(i, j, k)) / dy +
            met23edgef * (-pipuuedgef + 4.0 * pipuedgef - 3.0 * pipedgef) * 0.5 / dz);
  } else if (st.ko + k == st.cfg.z_size + st.nbz) {
    double pipddedgef = 0.5 * (pip(i, j, k - 2) + pip(i, j + 1, k - 2));
    double pipdedgef = 0.5 * (pip(i, j, k - 1) + pip(i, j + 1, k - 1));
    double pipedgef = 0.5 * (pip(i, j, k) + pip(i, j + 1, k));
    return kinv * mainv2 * pedgef / rhoedgef *
           ((pip(i, j + 1, k) - pip(i, j, k)) / dy +
            met23edgef * (pipddedgef - 4.0 * pipdedgef + 3.0 * pipedgef) * 0.5 / dz);
  }
  double pipuedgef = 0.5 * (pip(i, j, k + 1) + pip(i, j + 1, k + 1));
  double pipdedgef = 0.5 * (pip(i, j, k - 1) + pip(i, j + 1, k - 1));
  return kinv * mainv2 * pedgef / rhoedgef *
         ((pip(i, j + 1, k) - pip(i, j, k)) / dy + met23edgef * (pipuedgef - pipdedgef) * 0.5 / dz);
}

// compute_pressure_gradient.jl:211-297 (W)
double pressure_gradient_w(State& st, Arr3& pip, int i, int j, int k) {
  const double kinv = st.c.kappainv, mainv2 = st.c.mainv2, dx = st.dx, dy = st.dy, dz = st.dz;
  Arr3 &jac = st.jac, &rho = st.pr.rho, &rb = st.rhobar, &pb = st.pbar;
  Arr3 &m13 = st.met13, &m23 = st.met23, &m33 = st.met33;
  const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
  double rhoedgeu = (ju * rho(i, j, k) + jc * rho(i, j, k + 1)) / (jc + ju);
  rhoedgeu += (ju * rb(i, j, k) + jc * rb(i, j, k + 1)) / (jc + ju);
  double pedgeu = (ju * pb(i, j, k) + jc * pb(i, j, k + 1)) / (jc + ju);
  double met13edgeu = (ju * m13(i, j, k) + jc * m13(i, j, k + 1)) / (jc + ju);
  double met23edgeu = (ju * m23(i, j, k) + jc * m23(i, j, k + 1)) / (jc + ju);
  double met33edgeu = (ju * m33(i, j, k) + jc * m33(i, j, k + 1)) / (jc + ju);
  double pipredgeu = (jac(i + 1, j, k + 1) * pip(i + 1, j, k) + jac(i + 1, j, k) * pip(i + 1, j, k + 1)) /
                     (jac(i + 1, j, k) + jac(i + 1, j, k + 1));
  double pipledgeu = (jac(i - 1, j, k + 1) * pip(i - 1, j, k) + jac(i - 1, j, k) * pip(i - 1, j, k + 1)) /
                     (jac(i - 1, j, k) + jac(i - 1, j, k + 1));
  double pipfedgeu = (jac(i, j + 1, k + 1) * pip(i, j + 1, k) + jac(i, j + 1, k) * pip(i, j + 1, k + 1)) /
                     (jac(i, j + 1, k) + jac(i, j + 1, k + 1));
  double pipbedgeu = (jac(i, j - 1, k + 1) * pip(i, j - 1, k) + jac(i, j - 1, k) * pip(i, j - 1, k + 1)) /
                     (jac(i, j - 1, k) + jac(i, j - 1, k + 1));
  return kinv * mainv2 * pedgeu / rhoedgeu *
         (met13edgeu * (pipredgeu - pipledgeu) * 0.5 / dx + met23edgeu * (pipfedgeu - pipbedgeu) * 0.5 / dy +
          met33edgeu * (pip(i, j, k + 1) - pip(i, j, k)) / dz);
}

// transform.jl:86-122 (Cartesian)
double transform_cartesian(State& st, int i, int j, int k, double uedger, double uuedger, double uedgel,
                           double uuedgel, double vedgef, double vuedgef, double vedgeb, double vuedgeb,
                           double wedgeu) {
  Arr3 &jac = st.jac, &m13 = st.met13, &m23 = st.met23;
  const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
  double jacedgeu = 2.0 * jc * ju / (jc + ju);
  double uc = 0.5 * (uedger + uedgel), uu = 0.5 * (uuedger + uuedgel);
  double vc = 0.5 * (vedgef + vedgeb), vu = 0.5 * (vuedgef + vuedgeb);
  return jacedgeu *
         (-(ju * (m13(i, j, k) * uc + m23(i, j, k) * vc) + jc * (m13(i, j, k + 1) * uu + m23(i, j, k + 1) * vu)) /
              (jc + ju) +
          wedgeu);
}

// transform.jl:124-157 (Transformed)
double transform_transformed(State& st, int i, int j, int k, double uedger, double uuedger, double uedgel,
                             double uuedgel, double vedgef, double vuedgef, double vedgeb, double vuedgeb,
                             double wedgeu) {
  Arr3 &jac = st.jac, &m13 = st.met13, &m23 = st.met23;
  const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
  double jacedgeu = 2.0 * jc * ju / (jc + ju);
  double uc = 0.5 * (uedger + uedgel), uu = 0.5 * (uuedger + uuedgel);
  double vc = 0.5 * (vedgef + vedgeb), vu = 0.5 * (vuedgef + vuedgeb);
  return (ju * (m13(i, j, k) * uc + m23(i, j, k) * vc) + jc * (m13(i, j, k + 1) * uu + m23(i, j, k + 1) * vu)) /
             (jc + ju) +
         wedgeu / jacedgeu;
}

// compute_vertical_wind.jl:27-63
double vertical_wind(State& st, int i, int j, int k) {
  Arr3 &u = st.pr.u, &v = st.pr.v, &w = st.pr.w;
  return transform_cartesian(st, i, j, k, u(i, j, k), u(i, j, k + 1), u(i - 1, j, k), u(i - 1, j, k + 1),
                             v(i, j, k), v(i, j, k + 1), v(i, j - 1, k), v(i, j - 1, k + 1), w(i, j, k));
}

// compute_stress_tensor.jl:55-164
double stress_tensor(State& st, int i, int j, int k, int mu, int nu) {
  Arr3 &u = st.pr.u, &v = st.pr.v, &w = st.pr.w, &jac = st.jac;
  const double dx = st.dx, dy = st.dy, dz = st.dz;
  double jacedger = 0.5 * (jac(i, j, k) + jac(i + 1, j, k));
  double jacedgel = 0.5 * (jac(i, j, k) + jac(i - 1, j, k));
  double jacedgef = 0.5 * (jac(i, j, k) + jac(i, j + 1, k));
  double jacedgeb = 0.5 * (jac(i, j, k) + jac(i, j - 1, k));
  double jacedgeu = 2.0 * jac(i, j, k) * jac(i, j, k + 1) / (jac(i, j, k) + jac(i, j, k + 1));
  double jacedged = 2.0 * jac(i, j, k) * jac(i, j, k - 1) / (jac(i, j, k) + jac(i, j, k - 1));
  double uf = 0.5 * (u(i, j + 1, k) + u(i - 1, j + 1, k));
  double ub = 0.5 * (u(i, j - 1, k) + u(i - 1, j - 1, k));
  double uu = 0.5 * (u(i, j, k + 1) + u(i - 1, j, k + 1));
  double ud = 0.5 * (u(i, j, k - 1) + u(i - 1, j, k - 1));
  double vr = 0.5 * (v(i + 1, j, k) + v(i + 1, j - 1, k));
  double vl = 0.5 * (v(i - 1, j, k) + v(i - 1, j - 1, k));
  double vu = 0.5 * (v(i, j, k + 1) + v(i, j - 1, k + 1));
  double vd = 0.5 * (v(i, j, k - 1) + v(i, j - 1, k - 1));
  double wr = 0.5 * (vertical_wind(st, i + 1, j, k) + vertical_wind(st, i + 1, j, k - 1));
  double wl = 0.5 * (vertical_wind(st, i - 1, j, k) + vertical_wind(st, i - 1, j, k - 1));
  double wf = 0.5 * (vertical_wind(st, i, j + 1, k) + vertical_wind(st, i, j + 1, k - 1));
  double wb = 0.5 * (vertical_wind(st, i, j - 1, k) + vertical_wind(st, i, j - 1, k - 1));
  // compute_stress_tensor.jl:112-118: 2/3 * (divergence) / jac, left to right
  auto div = [&]() {
    return 2.0 / 3.0 *
           ((jacedger * u(i, j, k) - jacedgel * u(i - 1, j, k)) / dx +
            (jacedgef * v(i, j, k) - jacedgeb * v(i, j - 1, k)) / dy +
            (jacedgeu * w(i, j, k) - jacedged * w(i, j, k - 1)) / dz) /
           jac(i, j, k);
  };
  if (mu == 1 && nu == 1)
    return 2.0 * (u(i, j, k) - u(i - 1, j, k)) / dx + st.met13(i, j, k) * (uu - ud) / dz - div();
  if ((mu == 1 && nu == 2) || (mu == 2 && nu == 1))
    return 0.5 * (uf - ub) / dy + 0.5 * st.met23(i, j, k) * (uu - ud) / dz + 0.5 * (vr - vl) / dx +
           0.5 * st.met13(i, j, k) * (vu - vd) / dz;
  if ((mu == 1 && nu == 3) || (mu == 3 && nu == 1))
    return 0.5 * (uu - ud) / dz / jac(i, j, k) + 0.5 * (wr - wl) / dx +
           st.met13(i, j, k) * (vertical_wind(st, i, j, k) - vertical_wind(st, i, j, k - 1)) / dz;
  if (mu == 2 && nu == 2)
    return 2.0 * (v(i, j, k) - v(i, j - 1, k)) / dy + st.met23(i, j, k) * (vu - vd) / dz - div();
  if ((mu == 2 && nu == 3) || (mu == 3 && nu == 2))
    return 0.5 * (vu - vd) / dz / jac(i, j, k) + 0.5 * (wf - wb) / dy +
           st.met23(i, j, k) * (vertical_wind(st, i, j, k) - vertical_wind(st, i, j, k - 1)) / dz;
  // mu == 3 && nu == 3
  return 2.0 * (vertical_wind(st, i, j, k) - vertical_wind(st, i, j, k - 1)) / dz / jac(i, j, k) -
         div();
}

// Sponge closures of SpongeNamelist.jl:48-66 in the parametric forms of
// mountain_wave.jl:40-53 (lhs) and periodic_hill.jl:33-36 (rhs); dimensional.
static double lhs_sponge_dim(const pcf_config& n, double x, double y, double z) {
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

static double rhs_sponge_dim(const pcf_config& n, double z, double dt) {
  if (n.rhs_sponge != PCF_RHS_SPONGE_SIN2_TOP) return 0.0;
  if (z >= n.rhs_zr) {
    double s = std::sin(M_PI / 2 * (z - n.rhs_zr) / (n.lz - n.rhs_zr));
    return n.rhs_amplitude * (s * s) / dt;
  }
  return 0.0;
}

// compute_sponges!.jl:26-65
void compute_sponges(State& st, double dt, double time) {
  const int kmin = st.ko == 0 ? st.k0 : st.k0 - 1;
  const int kmax = st.ko + st.nz == st.cfg.z_size ? st.k1 : st.k1 + 1;
  const double lref = st.c.lref, tref = st.c.tref;
  for (int k = kmin; k <= kmax; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double xd = st.x[i] * lref, yd = st.y[j] * lref, zd = st.zc(i, j, k) * lref;
        double dtd = dt * tref;
        (void)time;
        st.alphar(i, j, k) = lhs_sponge_dim(st.cfg, xd, yd, zd) * tref;
        st.betar(i, j, k) = rhs_sponge_dim(st.cfg, zd, dtd) * tref;
      }
  zonal_bc(st, st.alphar);
  zonal_bc(st, st.betar);
  meridional_bc(st, st.alphar);
  meridional_bc(st, st.betar);
  for (int j = 1; j <= st.nyy; ++j)
    for (int i = 1; i <= st.nxx; ++i) {
      if (st.ko == 0) {
        st.alphar(i, j, st.k0 - 1) = st.alphar(i, j, st.k0);
        st.betar(i, j, st.k0 - 1) = st.betar(i, j, st.k0);
      }
      if (st.ko + st.nz == st.cfg.z_size) {
        st.alphar(i, j, st.k1 + 1) = st.alphar(i, j, st.k1);
        st.betar(i, j, st.k1 + 1) = st.betar(i, j, st.k1);
      }
    }
}

// update!(state, dt, m, var[, LHS()]) (update!.jl:505-1192)
void update_lhs(State& st, double dt, int m, int var) {
  const double dx = st.dx, dy = st.dy, dz = st.dz;
  const double am = st.alphark[m - 1], bm = st.betark[m - 1];
  const double fc = st.cfg.coriolis_frequency * st.c.tref;
  Arr3 &jac = st.jac, &rb = st.rhobar;
  Arr3 &rho = st.pr.rho, &rhoold = st.rhoold;
  const bool pi = st.cfg.model != PCF_BOUSSINESQ;
  auto divergence = [&](Arr4& f, int i, int j, int k) {
    return (f(i, j, k, 1) - f(i - 1, j, k, 1)) / dx + (f(i, j, k, 2) - f(i, j - 1, k, 2)) / dy +
           (f(i, j, k, 3) - f(i, j, k - 1, 3)) / dz;
  };
  if (var == PCF_VAR_CHI) {  // update!.jl:1460-1500
    if (!st.tracer) return;
    if (m == 1) st.dchi.fill(0.0);
    Arr3 &q = st.chi, &dq = st.dchi;
    Arr4& f = st.phichi;
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double fluxdiff = (f(i, j, k, 1) - f(i - 1, j, k, 1)) / dx + (f(i, j, k, 2) - f(i, j - 1, k, 2)) / dy +
                            (f(i, j, k, 3) - f(i, j, k - 1, 3)) / dz;
          fluxdiff /= jac(i, j, k);
          // compute_volume_force(..., Chi()) is 0 without MS-GWaM
          double fr = -fluxdiff + 0.0;
          dq(i, j, k) = dt * fr + am * dq(i, j, k);
          q(i, j, k) += bm * dq(i, j, k);
        }
    return;
  }
  if (var == PCF_VAR_RHO || var == PCF_VAR_RHOP || var == PCF_VAR_P) {
    if (var == PCF_VAR_RHO && !pi) return;            // update!.jl:511-519
    if (var == PCF_VAR_P && !st.compressible) return;  // update!.jl:1400-1408
    Arr3& q = var == PCF_VAR_RHO ? st.pr.rho : var == PCF_VAR_RHOP ? st.pr.rhop : st.pr.p;
    Arr3& dq = var == PCF_VAR_RHO ? st.drho : var == PCF_VAR_RHOP ? st.drhop : st.dp;
    Arr4& f = var == PCF_VAR_RHO ? st.phirho : var == PCF_VAR_RHOP ? st.phirhop : st.phip;
    if (m == 1) dq.fill(0.0);
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double fluxdiff = (f(i, j, k, 1) - f(i - 1, j, k, 1)) / dx + (f(i, j, k, 2) - f(i, j - 1, k, 2)) / dy +
                            (f(i, j, k, 3) - f(i, j, k - 1, 3)) / dz;
          fluxdiff /= jac(i, j, k);
          double fr;
          if (var == PCF_VAR_RHO) {
            fr = -fluxdiff;
          } else if (var == PCF_VAR_RHOP) {  // update!.jl:586-594
            double heating = conductive_heating(st, i, j, k);
            fr = -fluxdiff - heating / st.thetabar(i, j, k);
          } else {  // update!.jl:1410-1448
            double heating = conductive_heating(st, i, j, k);
            fr = -fluxdiff + heating;
          }
          dq(i, j, k) = dt * fr + am * dq(i, j, k);
          q(i, j, k) += bm * dq(i, j, k);
        }
    return;
  }
  if (var == PCF_VAR_U) {  // update!.jl:730-806
    Arr3 &u = st.pr.u, &v = st.pr.v;
    if (m == 1) st.du.fill(0.0);
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0 - 1; i <= st.i1; ++i) {
          double fluxdiff = divergence(st.phiu, i, j, k);
          double jacedger = 0.5 * (jac(i, j, k) + jac(i + 1, j, k));
          fluxdiff /= jacedger;
          double vc = 0.5 * (v(i, j, k) + v(i, j - 1, k));
          double vr = 0.5 * (v(i + 1, j, k) + v(i + 1, j - 1, k));
          double volforce = 0.5 * fc * ((rhoold(i, j, k) + rb(i, j, k)) * vc + (rhoold(i + 1, j, k) + rb(i + 1, j, k)) * vr);
          double force = -fluxdiff + volforce;
          double rhom_1 = 0.5 * (rhoold(i, j, k) + rhoold(i + 1, j, k));
          double rhom = 0.5 * (rho(i, j, k) + rho(i + 1, j, k));
          double rhobaredger = 0.5 * (rb(i, j, k) + rb(i + 1, j, k));
          rhom_1 += rhobaredger;
          rhom += rhobaredger;
          double um_1 = u(i, j, k);
          double momm_1 = rhom_1 * um_1;
          st.du(i, j, k) = dt * force + am * st.du(i, j, k);
          double momm = momm_1 + bm * st.du(i, j, k);
          u(i, j, k) = momm / rhom;
        }
    return;
  }
  if (var == PCF_VAR_V) {  // update!.jl:884-957
    Arr3 &v = st.pr.v, &uold = st.uold;
    if (m == 1) st.dv.fill(0.0);
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0 - 1; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double fluxdiff = divergence(st.phiv, i, j, k);
          double jacedgef = 0.5 * (jac(i, j, k) + jac(i, j + 1, k));
          fluxdiff /= jacedgef;
          double uc = 0.5 * (uold(i, j, k) + uold(i - 1, j, k));
          double uf = 0.5 * (uold(i, j + 1, k) + uold(i - 1, j + 1, k));
          double volforce = -0.5 * fc * ((rhoold(i, j, k) + rb(i, j, k)) * uc + (rhoold(i, j + 1, k) + rb(i, j + 1, k)) * uf);
          double force = -fluxdiff + volforce;
          double rhom_1 = 0.5 * (rhoold(i, j, k) + rhoold(i, j + 1, k));
          double rhom = 0.5 * (rho(i, j, k) + rho(i, j + 1, k));
          double rhobaredgef = 0.5 * (rb(i, j, k) + rb(i, j + 1, k));
          rhom_1 += rhobaredgef;
          rhom += rhobaredgef;
          double vm_1 = v(i, j, k);
          double momm_1 = rhom_1 * vm_1;
          st.dv(i, j, k) = dt * force + am * st.dv(i, j, k);
          double momm = momm_1 + bm * st.dv(i, j, k);
          v(i, j, k) = momm / rhom;
        }
    return;
  }
  if (var == PCF_VAR_W) {  // update!.jl:1035-1192
    Arr3 &w = st.pr.w, &uold = st.uold, &vold = st.vold, &m13 = st.met13, &m23 = st.met23;
    if (m == 1) st.dw.fill(0.0);
    const int kmin = st.ko == 0 ? st.k0 : st.k0 - 1;
    const int kmax = st.ko + st.nz == st.cfg.z_size ? st.k1 - 1 : st.k1;
#pragma omp parallel for schedule(static)
    for (int k = kmin; k <= kmax; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double fluxdiff = divergence(st.phiw, i, j, k);
          double jacedgeu = 2.0 * jac(i, j, k) * jac(i, j, k + 1) / (jac(i, j, k) + jac(i, j, k + 1));
          fluxdiff /= jacedgeu;
          double fdu[2][2], fdv[2][2];
          for (int ll = 0; ll <= 1; ++ll)
            for (int mm = 0; mm <= 1; ++mm) {
              fdu[ll][mm] = divergence(st.phiu, i - ll, j, k + mm);
              double jacedger = 0.5 * (jac(i - ll, j, k + mm) + jac(i + 1 - ll, j, k + mm));
              fdu[ll][mm] /= jacedger;
            }
          for (int ll = 0; ll <= 1; ++ll)
            for (int mm = 0; mm <= 1; ++mm) {
              fdv[ll][mm] = divergence(st.phiv, i, j - ll, k + mm);
              double jacedgef = 0.5 * (jac(i, j - ll, k + mm) + jac(i, j + 1 - ll, k + mm));
              fdv[ll][mm] /= jacedgef;
            }
          fluxdiff = transform_transformed(st, i, j, k, fdu[0][0], fdu[0][1], fdu[1][0], fdu[1][1], fdv[0][0],
                                           fdv[0][1], fdv[1][0], fdv[1][1], fluxdiff);
          double vc = 0.5 * (vold(i, j, k) + vold(i, j - 1, k));
          double vu = 0.5 * (vold(i, j, k + 1) + vold(i, j - 1, k + 1));
          double uc = 0.5 * (uold(i, j, k) + uold(i - 1, j, k));
          double uu = 0.5 * (uold(i, j, k + 1) + uold(i - 1, j, k + 1));
          const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
          double volforce =
              fc * (ju * m13(i, j, k) * (rhoold(i, j, k) + rb(i, j, k)) * vc +
                    jc * m13(i, j, k + 1) * (rhoold(i, j, k + 1) + rb(i, j, k + 1)) * vu) / (jc + ju) -
              fc * (ju * m23(i, j, k) * (rhoold(i, j, k) + rb(i, j, k)) * uc +
                    jc * m23(i, j, k + 1) * (rhoold(i, j, k + 1) + rb(i, j, k + 1)) * uu) / (jc + ju);
          double force = -fluxdiff + volforce;
          double rhom_1 = (ju * rhoold(i, j, k) + jc * rhoold(i, j, k + 1)) / (jc + ju);
          double rhom = (ju * rho(i, j, k) + jc * rho(i, j, k + 1)) / (jc + ju);
          double rhobaredgeu = (ju * rb(i, j, k) + jc * rb(i, j, k + 1)) / (jc + ju);
          rhom_1 += rhobaredgeu;
          rhom += rhobaredgeu;
          double wm_1 = w(i, j, k);
          double momm_1 = rhom_1 * wm_1;
          st.dw(i, j, k) = dt * force + am * st.dw(i, j, k);
          double momm = momm_1 + bm * st.dw(i, j, k);
          w(i, j, k) = momm / rhom;
        }
  }
}

// update!(state, dt, var, RHS(), Explicit()|Implicit()[, rf])
void update_rhs(State& st, double dt, int var, bool implicit, double rf) {
  const bool damp = st.cfg.damp_horizontal_wind_on_rhs != 0;
  const double g = st.c.g_ndim, dt2 = dt * dt;
  Arr3 &jac = st.jac, &rb = st.rhobar, &n2 = st.n2, &betar = st.betar, &m13 = st.met13, &m23 = st.met23;
  Arr3 &rho = st.pr.rho, &rhop = st.pr.rhop, &u = st.pr.u, &v = st.pr.v, &w = st.pr.w, &pip = st.pr.pip;
  const int kmax_uv = implicit && st.ko + st.nz != st.cfg.z_size ? st.k1 + 1 : st.k1;
  if (var == PCF_VAR_PIP) {  // update!.jl:1360-1392 (explicit; compressible only)
    if (!st.compressible) return;
    Arr3 &uold = st.uold, &vold = st.vold, &wold = st.wold, &p = st.pr.p;
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double fluxdiff = (uold(i, j, k) - uold(i - 1, j, k)) / st.dx + (vold(i, j, k) - vold(i, j - 1, k)) / st.dy +
                            (wold(i, j, k) - wold(i, j, k - 1)) / st.dz;
          fluxdiff /= jac(i, j, k);
          double heating = conductive_heating(st, i, j, k);
          double dpdpi_ = dpdpi(st, p(i, j, k));
          pip(i, j, k) -= dt * (fluxdiff - heating) / dpdpi_;
        }
    return;
  }
  if (var == PCF_VAR_U) {  // update!.jl:809-881
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= kmax_uv; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0 - 1; i <= st.i1; ++i) {
          double rhoedger = 0.5 * (rho(i, j, k) + rho(i + 1, j, k));
          double rhobaredger = 0.5 * (rb(i, j, k) + rb(i + 1, j, k));
          rhoedger += rhobaredger;
          double gradient = pressure_gradient_u(st, pip, i, j, k);
          double force = 0.0;
          double jpedger = wind_factor_u(st, i, j, k);
          if (!implicit) {
            u(i, j, k) += dt * (-gradient + force / rhoedger) * jpedger;
          } else {
            double factor = 1.0;
            if (damp) factor += dt * 0.5 * (betar(i, j, k) + betar(i + 1, j, k)) * rf;
            u(i, j, k) = 1.0 / factor * (u(i, j, k) + dt * (-gradient + force / rhoedger) * jpedger);
          }
        }
    return;
  }
  if (var == PCF_VAR_V) {  // update!.jl:960-1032
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= kmax_uv; ++k)
      for (int j = st.j0 - 1; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double rhoedgef = 0.5 * (rho(i, j, k) + rho(i, j + 1, k));
          double rhobaredgef = 0.5 * (rb(i, j, k) + rb(i, j + 1, k));
          rhoedgef += rhobaredgef;
          double gradient = pressure_gradient_v(st, pip, i, j, k);
          double force = 0.0;
          double jpedgef = wind_factor_v(st, i, j, k);
          if (!implicit) {
            v(i, j, k) += dt * (-gradient + force / rhoedgef) * jpedgef;
          } else {
            double factor = 1.0;
            if (damp) factor += dt * 0.5 * (betar(i, j, k) + betar(i, j + 1, k)) * rf;
            v(i, j, k) = 1.0 / factor * (v(i, j, k) + dt * (-gradient + force / rhoedgef) * jpedgef);
          }
        }
    return;
  }
  if (var == PCF_VAR_W) {  // update!.jl:1195-1342
    const int kmin = st.ko == 0 ? st.k0 : st.k0 - 1;
    const int kmax = st.ko + st.nz == st.cfg.z_size ? st.k1 - 1 : st.k1;
    Arr3& rhopold = st.rhopold;
#pragma omp parallel for schedule(static)
    for (int k = kmin; k <= kmax; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          const double jc = jac(i, j, k), ju = jac(i, j, k + 1);
          double rhoc = rho(i, j, k);
          double rhou = rho(i, j, k + 1);
          double rhoedgeu = (ju * rho(i, j, k) + jc * rho(i, j, k + 1)) / (jc + ju);
          rhoc += rb(i, j, k);
          rhou += rb(i, j, k + 1);
          rhoedgeu += (ju * rb(i, j, k) + jc * rb(i, j, k + 1)) / (jc + ju);
          double gradient = pressure_gradient_w(st, pip, i, j, k);
          double force = 0.0;
          if (!implicit) {
            double b = -g * (ju * rhopold(i, j, k) / rhoc / jc + jc * rhopold(i, j, k + 1) / rhou / ju) / (jc + ju);
            double jpedgeu = wind_factor_w(st, i, j, k);
            w(i, j, k) += dt * (b - gradient + force / rhoedgeu) * jpedgeu;
          } else {
            double n2edgeu = (ju * n2(i, j, k) + jc * n2(i, j, k + 1)) / (jc + ju);
            double factor = 1.0;
            factor += dt * (ju * betar(i, j, k) + jc * betar(i, j, k + 1)) / (jc + ju) * rf;
            double b = -g * (ju * rhop(i, j, k) / rhoc / jc + jc * rhop(i, j, k + 1) / rhou / ju) / (jc + ju);
            double jpedger = wind_factor_u(st, i, j, k), jpedgel = wind_factor_u(st, i - 1, j, k);
            double jpedgef = wind_factor_v(st, i, j, k), jpedgeb = wind_factor_v(st, i, j - 1, k);
            double jpuedger = wind_factor_u(st, i, j, k + 1), jpuedgel = wind_factor_u(st, i - 1, j, k + 1);
            double jpuedgef = wind_factor_v(st, i, j, k + 1), jpuedgeb = wind_factor_v(st, i, j - 1, k + 1);
            double uc = 0.5 * (u(i, j, k) / jpedger + u(i - 1, j, k) / jpedgel);
            double uu = 0.5 * (u(i, j, k + 1) / jpuedger + u(i - 1, j, k + 1) / jpuedgel);
            double vc = 0.5 * (v(i, j, k) / jpedgef + v(i, j - 1, k) / jpedgeb);
            double vu = 0.5 * (v(i, j, k + 1) / jpuedgef + v(i, j - 1, k + 1) / jpuedgeb);
            double jpedgeu = wind_factor_w(st, i, j, k);
            double fw = buoyancy_factor_w(st, i, j, k);
            w(i, j, k) = 1.0 / (factor + fw * n2edgeu * dt2) *
                         (w(i, j, k) - dt * gradient * jpedgeu + dt * b * jpedgeu + dt * force / rhoedgeu * jpedgeu +
                          jpedgeu * fw * n2edgeu * dt2 *
                              (ju * (m13(i, j, k) * uc + m23(i, j, k) * vc) + jc * (m13(i, j, k + 1) * uu + m23(i, j, k + 1) * vu)) /
                              (jc + ju));
          }
        }
    return;
  }
  if (var == PCF_VAR_RHOP) {
    if (!implicit) {  // update!.jl:600-629
#pragma omp parallel for schedule(static)
      for (int k = st.k0; k <= st.k1; ++k)
        for (int j = st.j0; j <= st.j1; ++j)
          for (int i = st.i0; i <= st.i1; ++i) {
            double jpu = wind_factor_w(st, i, j, k), jpd = wind_factor_w(st, i, j, k - 1);
            double wvrt = 0.5 * (vertical_wind(st, i, j, k) / jpu + vertical_wind(st, i, j, k - 1) / jpd);
            double buoy = -g * rhop(i, j, k) / (rho(i, j, k) + rb(i, j, k));
            double fb = buoyancy_factor_rhop(st, i, j, k);
            buoy -= dt * fb * n2(i, j, k) * wvrt;
            rhop(i, j, k) = -buoy * (rho(i, j, k) + rb(i, j, k)) / g;
          }
      return;
    }
    // update!.jl:632-727
    Arr3& wold = st.wold;
#pragma omp parallel for schedule(static)
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double rhoc = rho(i, j, k) + rb(i, j, k);
          const double jc = jac(i, j, k), ju = jac(i, j, k + 1), jd = jac(i, j, k - 1);
          double rhoedgeu = (ju * rho(i, j, k) + jc * rho(i, j, k + 1)) / (jc + ju);
          rhoedgeu += (ju * rb(i, j, k) + jc * rb(i, j, k + 1)) / (jc + ju);
          double rhoedged = (jd * rho(i, j, k) + jc * rho(i, j, k - 1)) / (jc + jd);
          rhoedged += (jd * rb(i, j, k) + jc * rb(i, j, k - 1)) / (jc + jd);
          double jpedgeu = wind_factor_w(st, i, j, k), jpedged = wind_factor_w(st, i, j, k - 1);
          double wv = 0.5 * (wold(i, j, k) / jpedgeu + wold(i, j, k - 1) / jpedged);
          double lower_gradient = pressure_gradient_w(st, pip, i, j, k - 1);
          double lower_force = 0.0;
          double upper_gradient = pressure_gradient_w(st, pip, i, j, k);
          double upper_force = 0.0;
          if (st.ko + k == st.k0) {
            lower_gradient = 0.0;
            lower_force = 0.0;
          } else if (st.ko + k == st.cfg.z_size + st.nbz) {
            upper_gradient = 0.0;
            upper_force = 0.0;
          }
          double gradient = 0.5 * (lower_gradient + upper_gradient);
          double force = 0.5 * (lower_force / rhoedged + upper_force / rhoedgeu) * rhoc;
          double factor = 1.0;
          factor += dt * betar(i, j, k) * rf;
          double b = -g * rhop(i, j, k) / (rho(i, j, k) + rb(i, j, k));
          double jpedger = wind_factor_u(st, i, j, k), jpedgel = wind_factor_u(st, i - 1, j, k);
          double jpedgef = wind_factor_v(st, i, j, k), jpedgeb = wind_factor_v(st, i, j - 1, k);
          double fb = buoyancy_factor_rhop(st, i, j, k);
          b = 1.0 / (factor + fb * n2(i, j, k) * dt2) *
              (-fb * n2(i, j, k) * dt * jac(i, j, k) * (wv + dt * (-gradient + force / rhoc)) + factor * b +
               fb * n2(i, j, k) * dt * jac(i, j, k) * factor * 0.5 *
                   (m13(i, j, k) * (u(i, j, k) / jpedger + u(i - 1, j, k) / jpedgel) +
                    m23(i, j, k) * (v(i, j, k) / jpedgef + v(i, j - 1, k) / jpedgeb)));
          rhop(i, j, k) = -b * (rho(i, j, k) + rb(i, j, k)) / g;
        }
  }
}

// apply_lhs_sponge! (apply_lhs_sponge!.jl:255-636), relax_to_mean = false
// or true (horizontal means over the interior, no MPI).
void apply_lhs_sponge(State& st, double dt, double time, int var) {
  (void)time;
  Arr3& alphar = st.alphar;
  const double lref = st.c.lref, uref = st.c.uref;
  const pcf_config& n = st.cfg;
  const bool pi = n.model != PCF_BOUSSINESQ;
  if (var == PCF_VAR_RHOP && st.compressible) {  // apply_lhs_sponge!.jl:301-328
    Arr3 &rho = st.pr.rho, &rhop = st.pr.rhop, &p = st.pr.p, &rb = st.rhobar, &tb = st.thetabar;
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double rhopbg = rb(i, j, k) * (1.0 - p(i, j, k) / tb(i, j, k) / (rho(i, j, k) + rb(i, j, k)));
          double alpha = alphar(i, j, k);
          double rhopold = rhop(i, j, k);
          double beta = 1.0 / (1.0 + alpha * dt);
          rhop(i, j, k) = (1.0 - beta) * rhopbg + beta * rhopold;
        }
    return;
  }
  if (var == PCF_VAR_P) {  // apply_lhs_sponge!.jl:580-602 (compressible only)
    if (!st.compressible) return;
    Arr3 &rho = st.pr.rho, &p = st.pr.p, &rb = st.rhobar;
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double pb = rb(i, j, k) * p(i, j, k) / (rho(i, j, k) + rb(i, j, k));
          double alpha = alphar(i, j, k);
          double pold = p(i, j, k);
          double beta = 1 / (1 + alpha * dt);
          p(i, j, k) = (1 - beta) * pb + beta * pold;
        }
    return;
  }
  if (var == PCF_VAR_PIP) {  // apply_lhs_sponge!.jl:542-568 (compressible only)
    if (!st.compressible) return;
    Arr3 &rho = st.pr.rho, &pip = st.pr.pip, &p = st.pr.p, &rb = st.rhobar;
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double dpdpi_ = dpdpi(st, p(i, j, k));
          double pib = rb(i, j, k) * p(i, j, k) / (rho(i, j, k) + rb(i, j, k)) / dpdpi_;
          double alpha = alphar(i, j, k);
          double pipold = pip(i, j, k);
          pip(i, j, k) = pipold - alpha * dt * (p(i, j, k) / dpdpi_ - pib);
        }
    return;
  }
  if (var == PCF_VAR_RHO || var == PCF_VAR_RHOP) {
    if (var == PCF_VAR_RHO && !pi) return;
    Arr3& q = var == PCF_VAR_RHO ? st.pr.rho : st.pr.rhop;
    const double bg = 0.0;
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double alpha = alphar(i, j, k);
          double old = q(i, j, k);
          double beta = 1.0 / (1.0 + alpha * dt);
          q(i, j, k) = (1.0 - beta) * bg + beta * old;
        }
    return;
  }
  if (var == PCF_VAR_CHI) {  // apply_lhs_sponge!.jl:615-636
    if (!st.tracer) return;
    for (int k = st.k0; k <= st.k1; ++k)
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) {
          double alpha = alphar(i, j, k);
          double old = st.chi(i, j, k);
          double beta = 1.0 / (1.0 + alpha * dt);
          st.chi(i, j, k) = (1.0 - beta) * st.initialtracer(i, j, k) + beta * old;
        }
    return;
  }
  Arr3& f = var == PCF_VAR_U ? st.pr.u : var == PCF_VAR_V ? st.pr.v : st.pr.w;
  if (n.relax_to_mean) {
    for (int k = st.k0; k <= st.k1; ++k) {
      double s = 0.0;
      for (int j = st.j0; j <= st.j1; ++j)
        for (int i = st.i0; i <= st.i1; ++i) s += f(i, j, k) / n.x_size / n.y_size;
      st.horizontal_mean[k - st.k0] = s;
    }
  }
  (void)lref;
  for (int k = st.k0; k <= st.k1; ++k)
    for (int j = st.j0; j <= st.j1; ++j)
      for (int i = st.i0; i <= st.i1; ++i) {
        double bgv, alpha;
        if (var == PCF_VAR_U) {
          bgv = n.relax_to_mean ? st.horizontal_mean[k - st.k0] : (n.relaxed_u + n.relaxed_u) / uref / 2;
          alpha = 0.5 * (alphar(i, j, k) + alphar(i + 1, j, k));
        } else if (var == PCF_VAR_V) {
          bgv = n.relax_to_mean ? st.horizontal_mean[k - st.k0] : (n.relaxed_v + n.relaxed_v) / uref / 2;
          alpha = 0.5 * (alphar(i, j, k) + alphar(i, j + 1, k));
        } else {
          const double jc = st.jac(i, j, k), ju = st.jac(i, j, k + 1);
          bgv = n.relax_to_mean
                    ? st.horizontal_mean[k - st.k0]
                    : (jc * (st.met13(i, j, k + 1) * n.relaxed_u + st.met23(i, j, k + 1) * n.relaxed_v + n.relaxed_w / ju) +
                       ju * (st.met13(i, j, k) * n.relaxed_u + st.met23(i, j, k) * n.relaxed_v + n.relaxed_w / jc)) /
                          (jc + ju) / uref;
          alpha = (ju * alphar(i, j, k) + jc * alphar(i, j, k + 1)) / (jc + ju);
        }
        double old = f(i, j, k);
        double beta = 1.0 / (1.0 + alpha * dt);
        f(i, j, k) = (1.0 - beta) * bgv + beta * old;
      }
}

}  // namespace orc
