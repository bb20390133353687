// On-device boundary conditions, replacing src/Boundaries with the MPI halo
// exchange of src/MPIOperations collapsed to one GPU: periodic copies in x and
// y, solid walls (reflection) in z.
#include "pcf_internal.hpp"

namespace pcf {

namespace {

constexpr int kMaxBcFields = 8;
struct FieldList {
  double* f[kMaxBcFields];
  int n;
};

// Periodic halos in x and y for several fields in one launch
// (set_zonal_boundaries_of_field!.jl:78-101 followed by
// set_meridional_boundaries_of_field!.jl).  Every halo cell receives the value
// the zonal-then-meridional copy leaves there: the interior cell at the
// wrapped (i, j).  Sources are interior cells and destinations halo cells, so
// one launch reproduces the two sequential passes.
__global__ void k_bc_xy(FieldList fl, Geo g, int nbx, int nby, int kb, int ke) {
  // halo cells of one z layer: x-halos for j in j0-nby..j1+nby, plus
  // y-halos for i in i0..i1 (x-halo corners are covered by the first set)
  const int nxh = 2 * nbx, nyh = 2 * nby;
  const int ncol = g.ny + nyh;                  // j range of the x-halos
  const long long per_layer = (long long)nxh * ncol + (long long)g.nx * nyh;
  const long long total = per_layer * (ke - kb + 1);
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int k = kb + (int)(t / per_layer);
    long long r = t % per_layer;
    int i, j;
    if (r < (long long)nxh * ncol) {
      int hx = (int)(r % nxh);
      j = g.j0 - nby + (int)(r / nxh);
      i = hx < nbx ? g.i0 - nbx + hx : g.i1 + 1 + (hx - nbx);
    } else {
      r -= (long long)nxh * ncol;
      int hy = (int)(r / g.nx);
      i = g.i0 + (int)(r % g.nx);
      j = hy < nby ? g.j0 - nby + hy : g.j1 + 1 + (hy - nby);
    }
    // wrapped interior source
    int is = i, js = j;
    while (is < g.i0) is += g.nx;
    while (is > g.i1) is -= g.nx;
    while (js < g.j0) js += g.ny;
    while (js > g.j1) js -= g.ny;
    size_t dst = g.h(i, j, k), src = g.h(is, js, k);
    for (int q = 0; q < fl.n; ++q) fl.f[q][dst] = fl.f[q][src];
  }
}

// Solid walls: set_vertical_boundaries_of_field!.jl:50-101, one thread per
// column, bottom then top, in the reference's sequential order.
__global__ void k_bc_z(double* f, Geo g, double sign, int staggered) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 1 + blockIdx.y;
  if (i > g.nxx) return;
  if (staggered) {
    f[g.h(i, j, g.k0 - 1)] = 0.0;
    for (int k = 1; k <= g.nbz; ++k) f[g.h(i, j, g.k0 - k)] = sign * f[g.h(i, j, g.k0 + k - 2)];
    f[g.h(i, j, g.k1)] = 0.0;
    for (int k = 1; k <= g.nbz; ++k) f[g.h(i, j, g.k1 + k)] = sign * f[g.h(i, j, g.k1 - k)];
  } else {
    for (int k = 1; k <= g.nbz; ++k) f[g.h(i, j, g.k0 - k)] = sign * f[g.h(i, j, g.k0 + k - 1)];
    for (int k = 1; k <= g.nbz; ++k) f[g.h(i, j, g.k1 + k)] = sign * f[g.h(i, j, g.k1 - k + 1)];
  }
}

// set_vertical_boundaries!(state, BoundaryFluxes()) (set_vertical_boundaries!.jl:274-347)
__global__ void k_bc_fluxes(double* chi3, Geo g, double* rho3, double* rhop3, double* u3, double* v3,
                            double* w3, double* p3, double* theta3) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 1 + blockIdx.y;
  if (i > g.nxx) return;
  size_t b = g.h(i, j, g.k0 - 1), t = g.h(i, j, g.k1);
  if (rho3) { rho3[b] = 0.0; rho3[t] = 0.0; }
  rhop3[b] = 0.0; rhop3[t] = 0.0;
  u3[b] = 0.0; u3[t] = 0.0;
  v3[b] = 0.0; v3[t] = 0.0;
  w3[g.h(i, j, g.k0 - 2)] = 0.0; w3[t] = 0.0;
  if (chi3) { chi3[b] = 0.0; chi3[t] = 0.0; }  // set_tracer_vertical_boundaries!.jl, BoundaryFluxes
  if (p3) { p3[b] = 0.0; p3[t] = 0.0; }        // compressible: set_vertical_boundaries!.jl:324-347
  if (theta3) { theta3[b] = 0.0; theta3[t] = 0.0; }
}

}  // namespace

int bc_xy(pcf_state* st, double* const* fields, int n, int nbx, int nby, int nbz) {
  const Geo& g = st->g;
  if (nbx < 0) nbx = g.nbx;
  if (nby < 0) nby = g.nby;
  if (nbz < 0) nbz = g.nbz;
  for (int off = 0; off < n; off += kMaxBcFields) {
    FieldList fl;
    fl.n = n - off < kMaxBcFields ? n - off : kMaxBcFields;
    for (int q = 0; q < fl.n; ++q) fl.f[q] = fields[off + q];
    int kb = g.k0 - nbz, ke = g.k1 + nbz;
    long long total = ((long long)2 * nbx * (g.ny + 2 * nby) + (long long)g.nx * 2 * nby) * (ke - kb + 1);
    unsigned blocks = blocks_for(total, 256);
    if (blocks > 65535u * 8) blocks = 65535u * 8;
    hipLaunchKernelGGL(k_bc_xy, dim3(blocks), dim3(256), 0, st->stream, fl, g, nbx, nby, kb, ke);
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int bc_z(pcf_state* st, double* f, int mode, bool staggered) {
  const Geo& g = st->g;
  hipLaunchKernelGGL(k_bc_z, dim3(blocks_for(g.nxx, 64), g.nyy), dim3(64), 0, st->stream, f, g,
                     mode > 0 ? 1.0 : -1.0, staggered ? 1 : 0);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// set_boundaries!(state, BoundaryPredictands()) (set_boundaries!.jl:43-56)
int set_boundaries_predictands(pcf_state* st) {
  ProfScope prof(st, PCF_K_UPDATES);
  const bool pi = st->cfg.model != PCF_BOUSSINESQ;
  double* f[7];
  int n = 0;
  if (pi) f[n++] = st->rho;
  f[n++] = st->rhop; f[n++] = st->u; f[n++] = st->v; f[n++] = st->w; f[n++] = st->pip;
  if (st->compressible) f[n++] = st->p;
  int rc = bc_xy(st, f, n);
  if (rc) return rc;
  if (pi && (rc = bc_z(st, st->rho, -1, false))) return rc;
  if ((rc = bc_z(st, st->rhop, -1, false))) return rc;
  if ((rc = bc_z(st, st->u, +1, false))) return rc;
  if ((rc = bc_z(st, st->v, +1, false))) return rc;
  if ((rc = bc_z(st, st->w, -1, true))) return rc;
  if ((rc = bc_z(st, st->pip, +1, false))) return rc;
  if (st->compressible && (rc = bc_z(st, st->p, +1, false))) return rc;  // set_vertical_boundaries!.jl:217-234
  if (st->tracer) {  // set_tracer_*_boundaries!.jl, BoundaryPredictands
    if ((rc = bc_xy(st, &st->chi, 1))) return rc;
    if ((rc = bc_z(st, st->chi, -1, false))) return rc;
  }
  return PCF_OK;
}

int set_boundaries_fluxes(pcf_state* st) {
  const Geo& g = st->g;
  const bool pi = st->cfg.model != PCF_BOUSSINESQ;
  hipLaunchKernelGGL(k_bc_fluxes, dim3(blocks_for(g.nxx, 64), g.nyy), dim3(64), 0, st->stream,
                     st->tracer ? st->flux[5][2] : nullptr, g,
                     pi ? st->flux[0][2] : nullptr, st->flux[1][2], st->flux[2][2], st->flux[3][2], st->flux[4][2],
                     st->compressible ? st->flux[6][2] : nullptr, st->compressible ? st->phitheta[2] : nullptr);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

}  // namespace pcf
