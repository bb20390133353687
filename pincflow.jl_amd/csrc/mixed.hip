// Mixed-precision pressure solve (precision = PCF_PRECISION_MIXED).
//
// BiCGSTAB keeps the reference's FP64 Krylov iteration (apply_bicgstab!.jl:
// 24-150): the operator applies, the residual recurrence, the dot products,
// norms and scalars and the solution update stay FP64.  The preconditioner
// (apply_preconditioner!.jl:40-124) runs in FP32 on FP32 copies of the 22
// horizontal operator coefficients and of the line-solve factors, which
// halves its traffic.  Because an FP32 M^-1 is not exactly linear, the
// solution is accumulated in the textbook form x += alpha M^-1 p + omega M^-1 s
// (bicgstab_core, k_bicg_xr_yz), which keeps x consistent with the FP64
// residual recurrence; the stopping test is the reference criterion.
//
// Optionally (initial_guess = PCF_GUESS_PREVIOUS) each implicit stage starts
// from the same stage's solution of the previous time step, rescaled by the
// ratio of the time steps (the solution is proportional to dt times the
// pressure correction); the criterion stays |r| / |b|.
#include <cmath>

#include "pcf_device.hpp"
#include "solver_common.hpp"

namespace pcf {

namespace {

typedef float flt4 __attribute__((ext_vector_type(4)));

struct PcArgs32 {
  const float* a[T_COUNT];  // the 22 horizontal coefficients are used
  const float* s;
  float* out;
  int nx, ny, nz;
};

// W consecutive floats: W = 4 as one 16-byte access, W = 1 as a scalar.
template <int W>
__device__ __forceinline__ void ld(const float* p, float (&o)[W]) {
  if constexpr (W == 4) {
    const flt4 v = *(const flt4*)p;
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
    o[0] = *p;
  }
}
template <int W>
__device__ __forceinline__ void ld_stream(const float* p, float (&o)[W]) {
  if constexpr (W == 4) {
    const flt4 v = __builtin_nontemporal_load((const flt4*)p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
    o[0] = __builtin_nontemporal_load(p);
  }
}
template <int W>
__device__ __forceinline__ void st_stream(float* p, const float (&v)[W]) {
  if constexpr (W == 4) {
    flt4 x;
    x.x = v[0];
    x.y = v[1];
    x.z = v[2];
    x.w = v[3];
    __builtin_nontemporal_store(x, (flt4*)p);
  } else {
    __builtin_nontemporal_store(v[0], p);
  }
}

// The later pseudo-time steps of the preconditioner in FP32:
// out = s + deta (H s - sin) with H the horizontal 22-point part of the
// operator (apply_preconditioner!.jl:62-63); sin is the FP64 input vector.
// W x-adjacent cells per thread, block = (64 / W) x (4 W) threads covering
// 64 x (4 W) cells of one z-level.
template <int W>
__global__ void __launch_bounds__(256) k_pc_rhs32(PcArgs32 A, const double* __restrict__ sin, float deta,
                                                  const SolverScalars* sc) {
  if (sc && sc->done) return;
  constexpr int BX = 64 / W;
  const int i0 = (blockIdx.x * BX + threadIdx.x) * W;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (i0 >= nx || j >= ny) return;
  const size_t sxy = (size_t)nx * ny;
  const size_t o = (size_t)i0 + (size_t)nx * j + sxy * k;
  const int im = i0 == 0 ? nx - 1 : i0 - 1, ip = i0 + W == nx ? 0 : i0 + W;
  const int jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
  const float* s = A.s;
  float C[5][W], F[5][W], B[5][W], L[5], R[5];
#pragma unroll
  for (int w = 0; w < 5; ++w) {
    const int kk = k - 2 + w;
    if (kk < 0 || kk >= nz) {
#pragma unroll
      for (int c = 0; c < W; ++c) C[w][c] = F[w][c] = B[w][c] = 0.0f;
      L[w] = R[w] = 0.0f;
      continue;
    }
    const size_t pl = sxy * kk;
    ld<W>(s + pl + (size_t)nx * j + i0, C[w]);
    ld<W>(s + pl + (size_t)nx * jp + i0, F[w]);
    ld<W>(s + pl + (size_t)nx * jm + i0, B[w]);
    L[w] = s[pl + (size_t)nx * j + im];
    R[w] = s[pl + (size_t)nx * j + ip];
  }
  float a[T_COUNT][W];
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) {
    if (t == T_AC || t == T_AD || t == T_AU) continue;
    ld_stream<W>(A.a[t] + o, a[t]);
  }
  float out[W];
#pragma unroll
  for (int c = 0; c < W; ++c) {
    auto XL = [&](int w) -> float { return c == 0 ? L[w] : C[w][c - 1]; };
    auto XR = [&](int w) -> float { return c == W - 1 ? R[w] : C[w][c + 1]; };
    auto X0 = [&](int w) -> float { return C[w][c]; };
    auto YF = [&](int w) -> float { return F[w][c]; };
    auto YB = [&](int w) -> float { return B[w][c]; };
    float v = a[T_AL][c] * XL(2) + a[T_AR][c] * XR(2) + a[T_AF][c] * YF(2) + a[T_AB][c] * YB(2);
    v += a[T_ARU][c] * XR(3) + a[T_ARD][c] * XR(1) + a[T_ALU][c] * XL(3) + a[T_ALD][c] * XL(1) +
         a[T_AFU][c] * YF(3) + a[T_AFD][c] * YF(1) + a[T_ABU][c] * YB(3) + a[T_ABD][c] * YB(1) +
         a[T_AUU][c] * X0(4) + a[T_ADD][c] * X0(0) + a[T_ARUU][c] * XR(4) + a[T_ARDD][c] * XR(0) +
         a[T_ALUU][c] * XL(4) + a[T_ALDD][c] * XL(0) + a[T_AFUU][c] * YF(4) + a[T_AFDD][c] * YF(0) +
         a[T_ABUU][c] * YB(4) + a[T_ABDD][c] * YB(0);
    out[c] = X0(2) + deta * (v - (float)sin[o + c]);
  }
  st_stream<W>(A.out + o, out);
}

// The preconditioner's line solve in z with FP32 factors (the recurrence of
// the FP64 k_pc_thomas); SrcT / OutT: FP64 or FP32 input and output vectors.
// The upward sweep keeps its values in the FP32 scratch `col`.
// FIRST: rhs = deta (0 - src).  Not __restrict__: keeps each batch of loads
// ahead of the batch's stores.
template <bool FIRST, typename SrcT, typename OutT>
__global__ void __launch_bounds__(64) k_pc_thomas32(const float* P, const float* E, const float* Q, const SrcT* src,
                                                    OutT* out, float* col, int nx, int ny, int nz, float deta,
                                                    const SolverScalars* sc) {
  if (sc && sc->done) return;
  constexpr int U = 8;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const size_t sxy = (size_t)nx * ny;
  const size_t c = (size_t)i + (size_t)nx * j;
  auto rhs_of = [&](SrcT v) -> float { return FIRST ? deta * (0.0f - (float)v) : (float)v; };
  float sk = rhs_of(src[c]) / P[c];
  col[c] = sk;
  int k = 1;
  for (; k + U <= nz; k += U) {
    float r[U], e[U], p[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t o = c + sxy * (k + u);
      r[u] = rhs_of(src[o]);
      e[u] = E[o];
      p[u] = P[o];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sk = (r[u] + e[u] * sk) * p[u];
      col[c + sxy * (k + u)] = sk;
    }
  }
  for (; k < nz; ++k) {
    const size_t o = c + sxy * k;
    sk = (rhs_of(src[o]) + E[o] * sk) * P[o];
    col[o] = sk;
  }
  out[c + sxy * (nz - 1)] = (OutT)sk;
  k = nz - 2;
  for (; k - (U - 1) >= 0; k -= U) {
    float r[U], q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t o = c + sxy * (k - u);
      r[u] = col[o];
      q[u] = Q[o];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sk = r[u] + q[u] * sk;
      out[c + sxy * (k - u)] = (OutT)sk;
    }
  }
  for (; k >= 0; --k) {
    const size_t o = c + sxy * k;
    sk = col[o] + Q[o] * sk;
    out[o] = (OutT)sk;
  }
}

__global__ void k_to_f32(const double* a, float* b, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// g = c x (the saved warm start, c = 1 / dt)
__global__ void k_scale_copy(const double* x, double* g, size_t n, double c) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    g[q] = c * x[q];
}

PcArgs32 pc_args32(pcf_state* st, const float* s, float* out) {
  PcArgs32 A;
  for (int t = 0; t < T_COUNT; ++t) A.a[t] = st->T32[t];
  A.s = s;
  A.out = out;
  A.nx = st->g.nx;
  A.ny = st->g.ny;
  A.nz = st->g.nz;
  return A;
}

// FP32 copies of the horizontal operator coefficients and of the line-solve
// factors of the current operator.
int prepare_mixed(pcf_state* st) {
  if (st->t32_valid) return PCF_OK;
  int rc = pc_factors(st);
  if (rc) return rc;
  const size_t N = st->g.nv();
  for (int t = 0; t < T_COUNT; ++t) {
    if (t == T_AC || t == T_AD || t == T_AU) continue;  // enter through the factors
    hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->T.a[t], st->T32[t], N);
  }
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->pc_P, st->pc32_P, N);
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->pc_E, st->pc32_E, N);
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->pc_Q, st->pc32_Q, N);
  PCF_HIP(hipGetLastError());
  st->t32_valid = true;
  return PCF_OK;
}

}  // namespace

// M^-1 sin -> sout (FP64 in and out, FP32 inside); sout must not alias sin.
int precondition_mixed(pcf_state* st, const double* sin, double* sout, const SolverScalars* sc) {
  int rc = prepare_mixed(st);
  if (rc) return rc;
  const Geo& g = st->g;
  const dim3 tgrd(blocks_for(g.nx, 64), g.ny), tblk(64);
  const bool quads = g.nx % 4 == 0;
  const dim3 sgrd = quads ? dim3(blocks_for(g.nx, 64), blocks_for(g.ny, 16), g.nz)
                          : dim3(blocks_for(g.nx, 64), blocks_for(g.ny, 4), g.nz);
  const dim3 sblk = quads ? dim3(16, 16) : dim3(64, 4);
  const float deta = (float)st->deta;
  const int iters = st->cfg.preconditioner_iterations;
  float* col = st->m_pc_tmp;  // upward-sweep scratch of the line solves
  float* cur = st->m_pc_cur;  // FP32 iterate between pseudo-time steps
  float* rhs = st->m_pc_rhs;  // FP32 right-hand side of a later pseudo-time step
  {
    ProfScope prof(st, PCF_K_PC_THOMAS);
    if (iters == 1)
      hipLaunchKernelGGL((k_pc_thomas32<true, double, double>), tgrd, tblk, 0, st->stream, st->pc32_P, st->pc32_E,
                         st->pc32_Q, sin, sout, col, g.nx, g.ny, g.nz, deta, sc);
    else
      hipLaunchKernelGGL((k_pc_thomas32<true, double, float>), tgrd, tblk, 0, st->stream, st->pc32_P, st->pc32_E,
                         st->pc32_Q, sin, cur, col, g.nx, g.ny, g.nz, deta, sc);
  }
  for (int it = 2; it <= iters; ++it) {
    {
      ProfScope prof(st, PCF_K_PC_RHS);
      if (quads)
        hipLaunchKernelGGL(k_pc_rhs32<4>, sgrd, sblk, 0, st->stream, pc_args32(st, cur, rhs), sin, deta, sc);
      else
        hipLaunchKernelGGL(k_pc_rhs32<1>, sgrd, sblk, 0, st->stream, pc_args32(st, cur, rhs), sin, deta, sc);
    }
    {
      ProfScope prof(st, PCF_K_PC_THOMAS);
      if (it == iters)
        hipLaunchKernelGGL((k_pc_thomas32<false, float, double>), tgrd, tblk, 0, st->stream, st->pc32_P, st->pc32_E,
                           st->pc32_Q, rhs, sout, col, g.nx, g.ny, g.nz, deta, sc);
      else
        hipLaunchKernelGGL((k_pc_thomas32<false, float, float>), tgrd, tblk, 0, st->stream, st->pc32_P, st->pc32_E,
                           st->pc32_Q, rhs, cur, col, g.nx, g.ny, g.nz, deta, sc);
    }
  }
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int apply_bicgstab_mixed(pcf_state* st, double tolref, double dt, int* niter_out) {
  const pcf_config& n = st->cfg;
  int rc = prepare_mixed(st);
  if (rc) return rc;
  const int slot = st->solve_slot;
  const bool use_prev = n.initial_guess == PCF_GUESS_PREVIOUS && slot >= 0 && slot < 2 && dt > 0.0;
  const double* guess = use_prev && st->m_prev_valid[slot] ? st->m_prev[slot] : nullptr;
  rc = bicgstab_core(st, tolref, true, guess, dt, niter_out);
  if (rc) return rc;
  if (use_prev) {  // keep solution / dt for the same stage of the next step
    const size_t N = st->g.nv();
    hipLaunchKernelGGL(k_scale_copy, dim3(2048), dim3(256), 0, st->stream, st->solution, st->m_prev[slot], N,
                       1.0 / dt);
    PCF_HIP(hipGetLastError());
    st->m_prev_valid[slot] = true;
  }
  return PCF_OK;
}

// Allocation of the mixed-precision buffers (pcf_create, precision MIXED).
int alloc_mixed(pcf_state* st) {
  const size_t N = st->g.nv();
  auto af = [&](float** p) -> int {
    PCF_HIP(hipMalloc((void**)p, N * sizeof(float)));
    st->allocations.push_back(*p);
    return PCF_OK;
  };
  int rc;
  for (int t = 0; t < T_COUNT; ++t) {
    if (t == T_AC || t == T_AD || t == T_AU) continue;
    if ((rc = af(&st->T32[t]))) return rc;
  }
  float** fl[] = {&st->pc32_P, &st->pc32_E, &st->pc32_Q, &st->m_pc_tmp, &st->m_pc_cur, &st->m_pc_rhs};
  for (float** p : fl)
    if ((rc = af(p))) return rc;
  double** dl[] = {&st->bzpc, &st->m_prev[0], &st->m_prev[1]};
  for (double** p : dl) {
    PCF_HIP(hipMalloc((void**)p, N * sizeof(double)));
    st->allocations.push_back(*p);
  }
  return PCF_OK;
}

}  // namespace pcf
