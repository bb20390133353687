// ORACLE -- test infrastructure only (see orc.hpp).
// Boundary conditions, src/Boundaries/*.jl with the MPI halo exchange of
// src/MPIOperations replaced by its one-process equivalent (periodic copy in
// x and y; PROC_NULL, i.e. nothing, in z).
#include "orc.hpp"

namespace orc {

// set_zonal_boundaries_of_field!.jl:78-101 (3-D; x_size > 1 goes through the
// self Sendrecv! of set_zonal_halos_of_field!.jl:48-73, which is the same copy)
void zonal_bc(State& st, Arr3& f, int nbx, int nby, int nbz) {
  if (nbx < 0) nbx = st.nbx;
  if (nby < 0) nby = st.nby;
  if (nbz < 0) nbz = st.nbz;
  for (int k = st.k0 - nbz; k <= st.k1 + nbz; ++k)
    for (int j = st.j0 - nby; j <= st.j1 + nby; ++j)
      for (int i = 1; i <= nbx; ++i) {
        f(st.i0 - i, j, k) = f(st.i1 - i + 1, j, k);
        f(st.i1 + i, j, k) = f(st.i0 + i - 1, j, k);
      }
}

// set_meridional_boundaries_of_field!.jl (3-D)
void meridional_bc(State& st, Arr3& f, int nbx, int nby, int nbz) {
  if (nbx < 0) nbx = st.nbx;
  if (nby < 0) nby = st.nby;
  if (nbz < 0) nbz = st.nbz;
  for (int k = st.k0 - nbz; k <= st.k1 + nbz; ++k)
    for (int j = 1; j <= nby; ++j)
      for (int i = st.i0 - nbx; i <= st.i1 + nbx; ++i) {
        f(i, st.j0 - j, k) = f(i, st.j1 - j + 1, k);
        f(i, st.j1 + j, k) = f(i, st.j0 + j - 1, k);
      }
}

// set_zonal_boundaries_of_field!.jl:56-76 (2-D)
void zonal_bc_2d(State& st, std::vector<double>& f) {
  auto F = [&](int i, int j) -> double& { return f[(size_t)(i - 1) + (size_t)st.nxx * (j - 1)]; };
  for (int j = 1; j <= st.nyy; ++j)
    for (int i = 1; i <= st.nbx; ++i) {
      F(st.i0 - i, j) = F(st.i1 - i + 1, j);
      F(st.i1 + i, j) = F(st.i0 + i - 1, j);
    }
}

void meridional_bc_2d(State& st, std::vector<double>& f) {
  auto F = [&](int i, int j) -> double& { return f[(size_t)(i - 1) + (size_t)st.nxx * (j - 1)]; };
  for (int j = 1; j <= st.nby; ++j)
    for (int i = 1; i <= st.nxx; ++i) {
      F(i, st.j0 - j) = F(i, st.j1 - j + 1);
      F(i, st.j1 + j) = F(i, st.j0 + j - 1);
    }
}

// set_vertical_boundaries_of_field!.jl:50-101 (3-D, solid walls at ko == 0
// and at the top); mode = +1 or -1 is the reflection sign.
void vertical_bc(State& st, Arr3& f, int mode, bool staggered) {
  const double s = mode > 0 ? 1.0 : -1.0;
  auto md = [&](double v) { return mode > 0 ? +v : -v; };
  (void)s;
  for (int j = st.j0 - st.nby; j <= st.j1 + st.nby; ++j)
    for (int i = st.i0 - st.nbx; i <= st.i1 + st.nbx; ++i) {
      if (st.ko == 0) {
        if (staggered) {
          f(i, j, st.k0 - 1) = 0.0;
          for (int k = 1; k <= st.nbz; ++k) f(i, j, st.k0 - k) = md(f(i, j, st.k0 + k - 2));
        } else {
          for (int k = 1; k <= st.nbz; ++k) f(i, j, st.k0 - k) = md(f(i, j, st.k0 + k - 1));
        }
      }
      if (st.ko + st.nz == st.cfg.z_size) {
        if (staggered) {
          f(i, j, st.k1) = 0.0;
          for (int k = 1; k <= st.nbz; ++k) f(i, j, st.k1 + k) = md(f(i, j, st.k1 - k));
        } else {
          for (int k = 1; k <= st.nbz; ++k) f(i, j, st.k1 + k) = md(f(i, j, st.k1 - k + 1));
        }
      }
    }
}

// set_boundaries!(state, BoundaryPredictands()) (set_boundaries!.jl:43-56):
// zonal, meridional, vertical (set_*_boundaries!.jl, BoundaryPredictands).
void set_boundaries_predictands(State& st) {
  Predictands& p = st.pr;
  const bool pi = st.cfg.model != PCF_BOUSSINESQ;
  std::vector<Arr3*> fields;
  if (pi) fields.push_back(&p.rho);
  for (Arr3* a : {&p.rhop, &p.u, &p.v, &p.w, &p.pip}) fields.push_back(a);
  if (st.compressible) fields.push_back(&p.p);
  for (Arr3* a : fields) zonal_bc(st, *a);
  for (Arr3* a : fields) meridional_bc(st, *a);
  if (pi) vertical_bc(st, p.rho, -1);
  vertical_bc(st, p.rhop, -1);
  vertical_bc(st, p.u, +1);
  vertical_bc(st, p.v, +1);
  vertical_bc(st, p.w, -1, true);
  vertical_bc(st, p.pip, +1);
  if (st.compressible) vertical_bc(st, p.p, +1);  // set_vertical_boundaries!.jl:217-234
  if (st.tracer) {  // set_tracer_*_boundaries!.jl, BoundaryPredictands
    zonal_bc(st, st.chi);
    meridional_bc(st, st.chi);
    vertical_bc(st, st.chi, -1);
  }
}

// set_boundaries!(state, BoundaryReconstructions()): periodic copies of the
// 5-D arrays (set_zonal_boundaries_of_field!.jl:104-132); z is a no-op at npz = 1.
void set_boundaries_reconstructions(State& st) {
  const bool pi = st.cfg.model != PCF_BOUSSINESQ;
  std::vector<Arr5*> rec;
  if (pi) rec.push_back(&st.rhotilde);
  for (Arr5* a : {&st.rhoptilde, &st.utilde, &st.vtilde, &st.wtilde}) rec.push_back(a);
  if (st.tracer) rec.push_back(&st.chitilde);
  for (Arr5* a : rec)
    for (auto& d : a->c)
      for (auto& q : d) zonal_bc(st, q);
  for (Arr5* a : rec)
    for (auto& d : a->c)
      for (auto& q : d) meridional_bc(st, q);
}

// set_boundaries!(state, BoundaryFluxes()) (set_vertical_boundaries!.jl:274-347)
void set_boundaries_fluxes(State& st) {
  const bool pi = st.cfg.model != PCF_BOUSSINESQ;
  for (int j = 1; j <= st.nyy; ++j)
    for (int i = 1; i <= st.nxx; ++i) {
      if (st.ko == 0) {
        if (pi) st.phirho(i, j, st.k0 - 1, 3) = 0.0;
        st.phirhop(i, j, st.k0 - 1, 3) = 0.0;
        st.phiu(i, j, st.k0 - 1, 3) = 0.0;
        st.phiv(i, j, st.k0 - 1, 3) = 0.0;
        st.phiw(i, j, st.k0 - 2, 3) = 0.0;
        if (st.tracer) st.phichi(i, j, st.k0 - 1, 3) = 0.0;
        if (st.compressible) {  // set_vertical_boundaries!.jl:324-347
          st.phitheta(i, j, st.k0 - 1, 3) = 0.0;
          st.phip(i, j, st.k0 - 1, 3) = 0.0;
        }
      }
      if (st.ko + st.nz == st.cfg.z_size) {
        if (pi) st.phirho(i, j, st.k1, 3) = 0.0;
        st.phirhop(i, j, st.k1, 3) = 0.0;
        st.phiu(i, j, st.k1, 3) = 0.0;
        st.phiv(i, j, st.k1, 3) = 0.0;
        st.phiw(i, j, st.k1, 3) = 0.0;
        if (st.tracer) st.phichi(i, j, st.k1, 3) = 0.0;
        if (st.compressible) {
          st.phitheta(i, j, st.k1, 3) = 0.0;
          st.phip(i, j, st.k1, 3) = 0.0;
        }
      }
    }
}

}  // namespace orc
