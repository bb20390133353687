// ORACLE -- test infrastructure only.
//
// CPU FP64 restatement of PinCFlow.jl's semi-implicit dynamical-core step
// (Boussinesq, pseudo-incompressible and compressible, npx = npy = npz = 1).  It exists to
// check the HIP library and to time the CPU baseline; nothing in the product
// (pincflow.jl_amd/) includes, links or calls it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Indices are 1-based and arrays are column-major (x fastest) so every loop
// reads like the reference loop it restates; each function cites the
// reference file:line it follows (paths relative to the reference root).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../include/pincflow_hip.h"

namespace orc {

// 3-D array with Julia indexing a(i, j, k), i fastest.
struct Arr3 {
  int nx = 0, ny = 0, nz = 0;
  std::vector<double> a;
  void init(int x, int y, int z) {
    nx = x; ny = y; nz = z;
    a.assign((size_t)x * y * z, 0.0);
  }
  double& operator()(int i, int j, int k) {
    return a[(size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1))];
  }
  double operator()(int i, int j, int k) const {
    return a[(size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1))];
  }
  void fill(double v) { for (auto& x : a) x = v; }
};

// 4-D flux array phi(i, j, k, d), d in 1..3 (Fluxes.jl:58-93).
struct Arr4 {
  Arr3 c[3];
  void init(int x, int y, int z) { for (auto& q : c) q.init(x, y, z); }
  double& operator()(int i, int j, int k, int d) { return c[d - 1](i, j, k); }
};

// 5-D reconstruction array tilde(i, j, k, d, s), d in 1..3, s in 1..2
// (Reconstructions.jl:51-84).
struct Arr5 {
  Arr3 c[3][2];
  void init(int x, int y, int z) { for (auto& d : c) for (auto& q : d) q.init(x, y, z); }
  double& operator()(int i, int j, int k, int d, int s) { return c[d - 1][s - 1](i, j, k); }
};

struct Predictands {  // Predictands.jl:60-68 (p is empty outside compressible, set_p.jl)
  Arr3 rho, rhop, u, v, w, pip, p;
};

struct Constants {  // Constants.jl:108-182
  double gamma, gammainv, kappa, kappainv, rsp, g;
  double rhoref, pref, aref, uref, lref, tref, thetaref, fref;
  double g_ndim, re, ma, mainv2, ma2, fr, frinv2, fr2, sig;
};

struct State {
  pcf_config cfg;
  Constants c;
  // Time.jl:31-38
  int nstages = 3;
  double alphark[3], betark[3], stepfrac[3];
  // Domain.jl (npx = npy = npz = 1: io = jo = ko = 0)
  int nx, ny, nz, nxx, nyy, nzz, nbx, nby, nbz;
  int io = 0, jo = 0, ko = 0;
  int i0, i1, j0, j1, k0, k1;
  // Grid.jl
  double lx, ly, lz, dx, dy, dz, dzcmin;
  std::vector<double> x, y, ztildes, zs;
  std::vector<double> hb;  // (nxx, nyy)
  Arr3 jac, met13, met23, met33, zc, zctilde;
  // Atmosphere.jl
  Arr3 pbar, thetabar, rhobar, n2;
  // Sponge.jl
  Arr3 alphar, betar;
  std::vector<double> horizontal_mean;
  // Variables.jl
  Predictands pr, p0, p1;
  Arr3 drho, drhop, du, dv, dw, dpip;        // Increments.jl
  Arr3 rhoold, rhopold, uold, vold, wold;    // Backups.jl
  Arr3 phi;                                  // Auxiliaries.jl
  Arr5 rhotilde, rhoptilde, utilde, vtilde, wtilde;
  Arr4 phirho, phirhop, phiu, phiv, phiw;
  // compressible model: P increment and the P / heat-conduction fluxes
  bool compressible = false;
  Arr3 dp;
  Arr4 phip, phitheta;
  // Poisson.jl
  Arr3 rhs, solution;                        // interior
  Arr3 tensor[25];                           // interior, order of enum pcf_vector - 2
  Arr3 s_op;                                 // Operator.jl (haloed)
  Arr3 s_pc, q_pc;                           // Preconditioner.jl (interior)
  std::vector<double> p_pc;                  // (nx, ny)
  Arr3 bp, br0, brold, br, bs, bt, bv, bmatvec, bv_pc;  // BiCGSTAB.jl
  Arr3 corx, cory;                           // Correction.jl (haloed)
  // Tracer.jl (tracer_setup = TracerOn): chi = rho chi, its increment,
  // initial state (TracerAuxiliaries), the chi0 backup of integrate.jl:276,
  // reconstructions and fluxes
  bool tracer = false;
  Arr3 chi, dchi, initialtracer, chi0;
  Arr5 chitilde;
  Arr4 phichi;
  double last_residual = 0.0;
  // integrate bookkeeping (integrate.jl:68-115)
  double time = 0.0;
  double nextoutputtime = 0.0;
  int itime = 0;
  std::string err;
  bool pbar_flat_bous = false;
};

// tensor slots, in the order of enum pcf_vector starting at PCF_V_AC
enum {
  T_AC = 0, T_AL, T_AR, T_AB, T_AF, T_AD, T_AU,
  T_ALD, T_ALU, T_ARD, T_ARU, T_ABD, T_ABU, T_AFD, T_AFU,
  T_ADD, T_AUU, T_ALDD, T_ALUU, T_ARDD, T_ARUU, T_ABDD, T_ABUU, T_AFDD, T_AFUU
};

// ---- setup (orc_setup.cpp) ----
int setup(State& st, const pcf_config& cfg);
double hb_dim(const pcf_config& cfg, double xdim, double ydim);
double stretch_dim(const pcf_config& cfg, double zhat);

// ---- boundaries (orc_boundaries.cpp) ----
void zonal_bc(State& st, Arr3& f, int nbx = -1, int nby = -1, int nbz = -1);
void meridional_bc(State& st, Arr3& f, int nbx = -1, int nby = -1, int nbz = -1);
void vertical_bc(State& st, Arr3& f, int mode, bool staggered = false);
void zonal_bc_2d(State& st, std::vector<double>& f);
void meridional_bc_2d(State& st, std::vector<double>& f);
void set_boundaries_predictands(State& st);
void set_boundaries_reconstructions(State& st);
void set_boundaries_fluxes(State& st);

// ---- Poisson (orc_poisson.cpp) ----
int compute_lhs(State& st, double& tolref);
void compute_operator(State& st, double dt, double rf);
void apply_operator(State& st, Arr3& sin, Arr3& ls, bool total);
void apply_preconditioner(State& st, Arr3& sin, Arr3& sout);
int apply_bicgstab(State& st, double tolref, int& niter);
int apply_corrector(State& st, double dt, double rf, int& niter);
void correct(State& st, double dt, double rf);
void solution_to_dpip(State& st, double dtinv);
double dot(const Arr3& a, const Arr3& b);
double sumsq(const Arr3& a);

// ---- pointwise helpers (orc_update.cpp) ----
double wind_factor_u(State& st, int i, int j, int k);
double wind_factor_v(State& st, int i, int j, int k);
double wind_factor_w(State& st, int i, int j, int k);
double conductive_heating(State& st, int i, int j, int k);
double dpdpi(State& st, double p);
double buoyancy_factor_rhop(State& st, int i, int j, int k);
double buoyancy_factor_w(State& st, int i, int j, int k);
double pressure_gradient_u(State& st, Arr3& pip, int i, int j, int k);
double pressure_gradient_v(State& st, Arr3& pip, int i, int j, int k);
double pressure_gradient_w(State& st, Arr3& pip, int i, int j, int k);
double vertical_wind(State& st, int i, int j, int k);
double transform_cartesian(State& st, int i, int j, int k, double uedger, double uuedger,
                           double uedgel, double uuedgel, double vedgef, double vuedgef,
                           double vedgeb, double vuedgeb, double wedgeu);
double transform_transformed(State& st, int i, int j, int k, double uedger, double uuedger,
                             double uedgel, double uuedgel, double vedgef, double vuedgef,
                             double vedgeb, double vuedgeb, double wedgeu);
double stress_tensor(State& st, int i, int j, int k, int mu, int nu);
void compute_sponges(State& st, double dt, double time);
void update_lhs(State& st, double dt, int m, int var);
void update_rhs(State& st, double dt, int var, bool implicit, double rf);
void apply_lhs_sponge(State& st, double dt, double time, int var);

// ---- advection (orc_flux.cpp) ----
void reconstruct(State& st);
void compute_fluxes(State& st, Predictands& p0);
void compute_theta_fluxes(State& st, Predictands& p0);

// ---- integration (orc_integrate.cpp) ----
int compute_time_step(State& st, double& dt);
int explicit_integration_lhs(State& st, Predictands& p0, double dt, double time);
int explicit_integration_rhs(State& st, double dt, double time);
int implicit_integration(State& st, double dt, double time, double rf, int& niter);
void copy_predictands(Predictands& dst, const Predictands& src);
void modify_compressible_wind(State& st, bool multiply);
void synchronize_compressible_atmosphere(State& st, const Predictands& p);

}  // namespace orc
