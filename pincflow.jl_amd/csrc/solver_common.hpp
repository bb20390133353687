// Deterministic reductions and the BiCGSTAB scalar updates shared by the
// FP64 solve (poisson.hip) and the mixed-precision solve (mixed.hip).
// Each translation unit gets its own copies (anonymous namespace).
#pragma once

#include "pcf_internal.hpp"

namespace pcf {
namespace {

// ---------------------------------------------------------------------------
// block reduction of up to 3 values (blockDim.x * blockDim.y == 256)
// ---------------------------------------------------------------------------
template <int ND>
__device__ __forceinline__ void block_partials(double (&val)[ND], double* partials, int block) {
  __shared__ double red[ND][kReduceBlock / 64];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    double x = val[d];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[d][wid] = x;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double s = red[d][0];
      for (int q = 1; q < kReduceBlock / 64; ++q) s += red[d][q];
      partials[(size_t)d * kPartialStride + block] = s;
    }
  }
}

// Sum nparts partials of dot slot d with one block of 1024 threads, in a
// fixed order.
__device__ double finalize_sum(const double* partials, int d, int nparts) {
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int stride = blockDim.x;
  const double* pp = partials + (size_t)d * kPartialStride;
  double s = 0.0;
  // the same order as one partial at a time, with 8 loads in flight
  int q = tid;
  for (; q + 7 * stride < nparts; q += 8 * stride) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = pp[q + u * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; q < nparts; q += stride) s += pp[q];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int q = 0; q < (int)(blockDim.x / 64); ++q) r += red[q];
  __syncthreads();
  return r;  // valid in thread 0
}

// Initial residual and convergence test (apply_bicgstab!.jl:56-69).  Partial
// slot 0 holds <r0, r0>; with a nonzero initial guess (guessed = 1) slot 1
// holds <b, b> and the criterion stays |r| / |b|, as for x0 = 0.
__global__ void k_bicg_start(const double* partials, int nparts, SolverScalars* sc, double nxyz, double tol,
                             int guessed) {
  double rr = finalize_sum(partials, 0, nparts);
  double bb = guessed ? finalize_sum(partials, 1, nparts) : rr;
  if (threadIdx.x == 0) {
    double res = sqrt(rr / nxyz);
    double b_norm = guessed ? sqrt(bb / nxyz) : res;
    sc->b_norm = b_norm;
    sc->res = res;
    sc->tol = tol;
    sc->niter = 0;
    sc->rr0 = rr;  // <r0, r0> = <r, r0> for r = r0
    sc->done = (res == 0.0 || res / b_norm <= tol) ? 1 : 0;
  }
}

// The scalar steps below are one block of 1024 threads per solve: the
// single-solve kernels run them on their state's scalars, the ensemble
// kernels (poisson.hip) on each member's, one block per member.

// alpha = <r, r0> / <v, r0>, from the partials of the A p kernel
__device__ __forceinline__ void bicg_alpha(const double* partials, int nparts, SolverScalars* sc) {
  if (sc->done) return;
  double vr0 = finalize_sum(partials, 0, nparts);
  if (threadIdx.x == 0) {
    sc->vr0 = vr0;
    sc->alpha = sc->rr0 / vr0;
  }
}
__global__ void k_bicg_alpha(const double* partials, int nparts, SolverScalars* sc) {
  bicg_alpha(partials, nparts, sc);
}

// omega = <t, s> / <t, t>
__device__ __forceinline__ void bicg_omega(const double* partials, int nparts, SolverScalars* sc) {
  if (sc->done) return;
  double ts = finalize_sum(partials, 0, nparts);
  double tt = finalize_sum(partials, 1, nparts);
  if (threadIdx.x == 0) {
    sc->ts = ts;
    sc->tt = tt;
    sc->omega = ts / tt;
  }
}
__global__ void k_bicg_omega(const double* partials, int nparts, SolverScalars* sc) {
  bicg_omega(partials, nparts, sc);
}

// residual, convergence and beta (apply_bicgstab!.jl:116-142).  maxit > 0
// (the ensemble solve) stops a solve that reaches maxit iterations without
// converging with done = 2; the single solve stops by its host loop.
__device__ __forceinline__ void bicg_check(const double* partials, int nparts, SolverScalars* sc, double nxyz,
                                           int maxit) {
  if (sc->done) return;
  double rr = finalize_sum(partials, 0, nparts);
  double rr0n = finalize_sum(partials, 1, nparts);
  if (threadIdx.x == 0) {
    double res = sqrt(rr / nxyz);
    sc->rr = rr;
    sc->res = res;
    sc->niter = sc->niter + 1;
    if (res / sc->b_norm <= sc->tol) {
      sc->done = 1;
    } else {
      sc->beta = sc->alpha / sc->omega * rr0n / sc->rr0;
      sc->rr0 = rr0n;
      if (maxit > 0 && sc->niter >= maxit) sc->done = 2;
    }
    sc->rr0_new = rr0n;
  }
}
__global__ void k_bicg_check(const double* partials, int nparts, SolverScalars* sc, double nxyz) {
  bicg_check(partials, nparts, sc, nxyz, 0);
}

}  // namespace
}  // namespace pcf
