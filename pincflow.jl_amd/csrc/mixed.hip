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
  const float* a[T_COUNT];  // the 22 horizontal coefficients are used here
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
// operator (apply_preconditioner!.jl:62-63); sin: the FP64 or FP32 input.
// W x-adjacent cells per thread, block = (64 / W) x (4 W) threads covering
// 64 x (4 W) cells of one z-level.
template <int W, typename SinT>
__global__ void __launch_bounds__(256) k_pc_rhs32(PcArgs32 A, const SinT* __restrict__ sin, float deta,
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

// ---------------------------------------------------------------------------
// Multigrid preconditioner (FP32).  The coarse operator of level l+1 is the
// Galerkin product R A P of level l with piecewise-constant aggregation of
// fx x fy columns (P: copy a coarse value to its fine cells, R = P^T: sum
// over them); z is never coarsened, so the z-line pseudo-time step of the
// reference preconditioner remains the smoother on every level.  Because
// the fine stencil couples a cell with x or y neighbours but never with
// diagonal ones, R A P has the same 25-point shape.
// ---------------------------------------------------------------------------
constexpr int kMgNu = 2;             // pre- and post-smoothing steps
constexpr float kMgOmega = 1.5f;     // over-correction of the coarse-level correction
constexpr int kMgCoarseSteps = 20;   // smoothing steps on the coarsest level
constexpr int kMgMinSize = 8;        // a dimension is coarsened while even and >= this

// (di, dj, dk) of the 25 coefficients in the order of T_AC ... T_AFUU
__device__ constexpr int kOffI[T_COUNT] = {0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0};
__device__ constexpr int kOffJ[T_COUNT] = {0, 0, 0, -1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1};
__device__ constexpr int kOffK[T_COUNT] = {0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -2, 2, -2, 2, -2, 2, -2, 2, -2, 2};

// coefficient index of the offset (di, dj, dk); -1 if none
__host__ __device__ constexpr int coef_of(int di, int dj, int dk) {
  for (int t = 0; t < T_COUNT; ++t) {
    const int oi[T_COUNT] = {0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0};
    const int oj[T_COUNT] = {0, 0, 0, -1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1};
    const int ok[T_COUNT] = {0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -2, 2, -2, 2, -2, 2, -2, 2, -2, 2};
    if (oi[t] == di && oj[t] == dj && ok[t] == dk) return t;
  }
  return -1;
}
__host__ __device__ constexpr int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

struct MgCoefs {
  float* a[T_COUNT];
};

// R A P: one thread per coarse cell (I, J, k).
template <int FX, int FY>
__global__ void k_mg_coarsen(MgCoefs fine, MgCoefs coarse, int nx, int ny, int nz, int ncx, int ncy) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y, k = blockIdx.z;
  if (I >= ncx) return;
  constexpr int oi[T_COUNT] = {0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0};
  constexpr int oj[T_COUNT] = {0, 0, 0, -1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1};
  constexpr int ok[T_COUNT] = {0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -2, 2, -2, 2, -2, 2, -2, 2, -2, 2};
  float acc[T_COUNT];
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) acc[t] = 0.0f;
  const size_t sxy = (size_t)nx * ny;
#pragma unroll
  for (int b = 0; b < FY; ++b)
#pragma unroll
    for (int a = 0; a < FX; ++a) {
      const size_t o = (size_t)(FX * I + a) + (size_t)nx * (FY * J + b) + sxy * k;
#pragma unroll
      for (int t = 0; t < T_COUNT; ++t) {
        const int target = coef_of(floor_div(a + oi[t], FX), floor_div(b + oj[t], FY), ok[t]);
        acc[target] += fine.a[t][o];
      }
    }
  const size_t oc = (size_t)I + (size_t)ncx * J + (size_t)ncx * ncy * k;
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) coarse.a[t][oc] = acc[t];
}

// Line-solve factors (see k_pc_factors) of a coarse operator, in FP32.
__global__ void __launch_bounds__(64) k_mg_factors(const float* ac, const float* ad, const float* au, float* P, float* E,
                                                   float* Q, int nx, int ny, int nz, float deta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx) return;
  const size_t sxy = (size_t)nx * ny;
  size_t o = (size_t)i + (size_t)nx * j;
  const float denom = 1 - deta * ac[o];
  float qk = deta * au[o] / denom;
  P[o] = denom;
  E[o] = deta * ad[o];
  Q[o] = qk;
  for (int k = 1; k < nz; ++k) {
    o += sxy;
    const float e = deta * ad[o];
    const float p = 1 / (1 - deta * ac[o] - e * qk);
    qk = deta * au[o] * p;
    P[o] = p;
    E[o] = e;
    Q[o] = qk;
  }
}

// b_coarse = R (b - A s): one thread per coarse cell sums the residuals of
// its fx x fy fine cells.
template <int FX, int FY, typename BT>
__global__ void k_mg_restrict(MgCoefs A, const BT* __restrict__ b, const float* __restrict__ s, float* bc, int nx,
                              int ny, int nz, int ncx, int ncy, const SolverScalars* sc) {
  if (sc && sc->done) return;
  const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y, k = blockIdx.z;
  if (I >= ncx) return;
  constexpr int oi[T_COUNT] = {0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0};
  constexpr int oj[T_COUNT] = {0, 0, 0, -1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1};
  constexpr int ok[T_COUNT] = {0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -2, 2, -2, 2, -2, 2, -2, 2, -2, 2};
  const size_t sxy = (size_t)nx * ny;
  float sum = 0.0f;
#pragma unroll
  for (int bb = 0; bb < FY; ++bb)
#pragma unroll
    for (int a = 0; a < FX; ++a) {
      const int i = FX * I + a, j = FY * J + bb;
      const size_t o = (size_t)i + (size_t)nx * j + sxy * k;
      float v = 0.0f;
#pragma unroll
      for (int t = 0; t < T_COUNT; ++t) {
        const int kk = k + ok[t];
        if (kk < 0 || kk >= nz) continue;
        int ii = i + oi[t], jj = j + oj[t];
        ii = ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii);
        jj = jj < 0 ? jj + ny : (jj >= ny ? jj - ny : jj);
        v += A.a[t][o] * s[(size_t)ii + (size_t)nx * jj + sxy * kk];
      }
      sum += (float)b[o] - v;
    }
  bc[(size_t)I + (size_t)ncx * J + (size_t)ncx * ncy * k] = sum;
}

// s += omega P e
__global__ void k_mg_prolong(float* s, const float* e, int nx, int ny, int nz, int fx, int fy, int ncx, int ncy,
                             float omega, const SolverScalars* sc) {
  if (sc && sc->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= nx) return;
  const size_t o = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
  s[o] += omega * e[(size_t)(i / fx) + (size_t)ncx * (j / fy) + (size_t)ncx * ncy * k];
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
  for (int t = 0; t < T_COUNT; ++t)
    hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->T.a[t], st->T32[t], N);
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->pc_P, st->pc32_P, N);
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->pc_E, st->pc32_E, N);
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st->stream, st->pc_Q, st->pc32_Q, N);
  // coarse operators and their line-solve factors
  for (size_t l = 1; l < st->mg.size(); ++l) {
    const MgLevel& F = st->mg[l - 1];
    const MgLevel& C = st->mg[l];
    MgCoefs fa, ca;
    for (int t = 0; t < T_COUNT; ++t) {
      fa.a[t] = F.a[t];
      ca.a[t] = C.a[t];
    }
    const dim3 grd(blocks_for(C.nx, 64), C.ny, C.nz), blk(64);
    if (F.fx == 2 && F.fy == 2)
      hipLaunchKernelGGL((k_mg_coarsen<2, 2>), grd, blk, 0, st->stream, fa, ca, F.nx, F.ny, F.nz, C.nx, C.ny);
    else if (F.fx == 2)
      hipLaunchKernelGGL((k_mg_coarsen<2, 1>), grd, blk, 0, st->stream, fa, ca, F.nx, F.ny, F.nz, C.nx, C.ny);
    else
      hipLaunchKernelGGL((k_mg_coarsen<1, 2>), grd, blk, 0, st->stream, fa, ca, F.nx, F.ny, F.nz, C.nx, C.ny);
    hipLaunchKernelGGL(k_mg_factors, dim3(blocks_for(C.nx, 64), C.ny), dim3(64), 0, st->stream, C.a[T_AC],
                       C.a[T_AD], C.a[T_AU], C.P, C.E, C.Q, C.nx, C.ny, C.nz, C.deta);
  }
  PCF_HIP(hipGetLastError());
  st->t32_valid = true;
  return PCF_OK;
}

// Smoothing steps of one level: the first from s = 0 when from_zero; the
// last writes to out64 when given (FP64 result of the finest level).
template <typename BT>
void mg_smooth(pcf_state* st, const MgLevel& L, const BT* b, bool from_zero, int steps, double* out64,
               const SolverScalars* sc) {
  const dim3 tgrd(blocks_for(L.nx, 64), L.ny), tblk(64);
  const bool quads = L.nx % 4 == 0;
  const dim3 sgrd = quads ? dim3(blocks_for(L.nx, 64), blocks_for(L.ny, 16), L.nz)
                          : dim3(blocks_for(L.nx, 64), blocks_for(L.ny, 4), L.nz);
  const dim3 sblk = quads ? dim3(16, 16) : dim3(64, 4);
  PcArgs32 A;
  for (int t = 0; t < T_COUNT; ++t) A.a[t] = L.a[t];
  A.s = L.s;
  A.out = L.rhs;
  A.nx = L.nx;
  A.ny = L.ny;
  A.nz = L.nz;
  for (int step = 0; step < steps; ++step) {
    const bool last = step == steps - 1;
    if (step == 0 && from_zero) {
      ProfScope prof(st, PCF_K_PC_THOMAS);
      if (last && out64)
        hipLaunchKernelGGL((k_pc_thomas32<true, BT, double>), tgrd, tblk, 0, st->stream, L.P, L.E, L.Q, b, out64,
                           L.col, L.nx, L.ny, L.nz, L.deta, sc);
      else
        hipLaunchKernelGGL((k_pc_thomas32<true, BT, float>), tgrd, tblk, 0, st->stream, L.P, L.E, L.Q, b, L.s, L.col,
                           L.nx, L.ny, L.nz, L.deta, sc);
      continue;
    }
    {
      ProfScope prof(st, PCF_K_PC_RHS);
      if (quads)
        hipLaunchKernelGGL((k_pc_rhs32<4, BT>), sgrd, sblk, 0, st->stream, A, b, L.deta, sc);
      else
        hipLaunchKernelGGL((k_pc_rhs32<1, BT>), sgrd, sblk, 0, st->stream, A, b, L.deta, sc);
    }
    ProfScope prof(st, PCF_K_PC_THOMAS);
    if (last && out64)
      hipLaunchKernelGGL((k_pc_thomas32<false, float, double>), tgrd, tblk, 0, st->stream, L.P, L.E, L.Q, L.rhs,
                         out64, L.col, L.nx, L.ny, L.nz, L.deta, sc);
    else
      hipLaunchKernelGGL((k_pc_thomas32<false, float, float>), tgrd, tblk, 0, st->stream, L.P, L.E, L.Q, L.rhs, L.s,
                         L.col, L.nx, L.ny, L.nz, L.deta, sc);
  }
}

// One V-cycle on level l for the right-hand side b; the finest level writes
// its result to out64.
template <typename BT>
void mg_vcycle(pcf_state* st, size_t l, const BT* b, double* out64, const SolverScalars* sc) {
  const MgLevel& L = st->mg[l];
  if (l + 1 == st->mg.size()) {
    mg_smooth(st, L, b, true, kMgCoarseSteps, out64, sc);
    return;
  }
  const MgLevel& C = st->mg[l + 1];
  mg_smooth(st, L, b, true, kMgNu, nullptr, sc);
  MgCoefs A;
  for (int t = 0; t < T_COUNT; ++t) A.a[t] = L.a[t];
  const dim3 cgrd(blocks_for(C.nx, 64), C.ny, C.nz), blk(64);
  {
    ProfScope prof(st, PCF_K_PC_RHS);
    if (L.fx == 2 && L.fy == 2)
      hipLaunchKernelGGL((k_mg_restrict<2, 2, BT>), cgrd, blk, 0, st->stream, A, b, L.s, C.b, L.nx, L.ny, L.nz, C.nx,
                         C.ny, sc);
    else if (L.fx == 2)
      hipLaunchKernelGGL((k_mg_restrict<2, 1, BT>), cgrd, blk, 0, st->stream, A, b, L.s, C.b, L.nx, L.ny, L.nz, C.nx,
                         C.ny, sc);
    else
      hipLaunchKernelGGL((k_mg_restrict<1, 2, BT>), cgrd, blk, 0, st->stream, A, b, L.s, C.b, L.nx, L.ny, L.nz, C.nx,
                         C.ny, sc);
  }
  mg_vcycle<float>(st, l + 1, C.b, nullptr, sc);
  {
    ProfScope prof(st, PCF_K_PC_RHS);
    hipLaunchKernelGGL(k_mg_prolong, dim3(blocks_for(L.nx, 64), L.ny, L.nz), blk, 0, st->stream, L.s, C.s, L.nx,
                       L.ny, L.nz, L.fx, L.fy, C.nx, C.ny, kMgOmega, sc);
  }
  mg_smooth(st, L, b, false, kMgNu, out64, sc);
}

}  // namespace

// M^-1 sin -> sout (FP64 in and out, FP32 inside); sout must not alias sin.
int precondition_mixed(pcf_state* st, const double* sin, double* sout, const SolverScalars* sc) {
  int rc = prepare_mixed(st);
  if (rc) return rc;
  if (st->cfg.preconditioner_type == PCF_PC_MULTIGRID) {
    mg_vcycle<double>(st, 0, sin, sout, sc);
    PCF_HIP(hipGetLastError());
    return PCF_OK;
  }
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
        hipLaunchKernelGGL((k_pc_rhs32<4, double>), sgrd, sblk, 0, st->stream, pc_args32(st, cur, rhs), sin, deta,
                           sc);
      else
        hipLaunchKernelGGL((k_pc_rhs32<1, double>), sgrd, sblk, 0, st->stream, pc_args32(st, cur, rhs), sin, deta,
                           sc);
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
  const bool stage = slot >= 0 && slot < 2 && dt > 0.0;
  const bool use_prev = (n.initial_guess == PCF_GUESS_PREVIOUS || n.initial_guess == PCF_GUESS_LEARNED) && stage;
  const double* prev = use_prev && st->m_prev_valid[slot] ? st->m_prev[slot] : nullptr;
  const double* guess = n.initial_guess == PCF_GUESS_PREVIOUS ? prev : nullptr;
  double scale = dt;
  if (n.initial_guess == PCF_GUESS_LEARNED && stage) {
    if ((rc = learned_guess(st, dt, prev, &guess))) return rc;
    scale = 1.0;
  }
  rc = bicgstab_core(st, tolref, true, guess, scale, niter_out);
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
  for (int t = 0; t < T_COUNT; ++t)
    if ((rc = af(&st->T32[t]))) return rc;
  float** fl[] = {&st->pc32_P, &st->pc32_E, &st->pc32_Q, &st->m_pc_tmp, &st->m_pc_cur, &st->m_pc_rhs};
  for (float** p : fl)
    if ((rc = af(p))) return rc;
  double** dl[] = {&st->bzpc, &st->m_prev[0], &st->m_prev[1]};
  for (double** p : dl) {
    PCF_HIP(hipMalloc((void**)p, N * sizeof(double)));
    st->allocations.push_back(*p);
  }
  if (st->cfg.preconditioner_type != PCF_PC_MULTIGRID) return PCF_OK;
  // level 0 works on the buffers above; coarser levels get their own
  const Geo& g = st->g;
  MgLevel L0;
  L0.nx = g.nx;
  L0.ny = g.ny;
  L0.nz = g.nz;
  L0.deta = (float)st->deta;
  for (int t = 0; t < T_COUNT; ++t) L0.a[t] = st->T32[t];
  L0.P = st->pc32_P;
  L0.E = st->pc32_E;
  L0.Q = st->pc32_Q;
  L0.s = st->m_pc_cur;
  L0.rhs = st->m_pc_rhs;
  L0.col = st->m_pc_tmp;
  st->mg.clear();
  st->mg.push_back(L0);
  // the coarse levels' "grid spacings": Galerkin aggregation of fx x fy
  // columns multiplies the x couplings by fy and the y couplings by fx, which
  // sets the pseudo-time step of each level's smoother like the reference's
  // deta = dtau / (2 (1/dx^2 + 1/dy^2)) (apply_preconditioner!.jl:58)
  double cx = 1.0 / (g.dx * g.dx), cy = 1.0 / (g.dy * g.dy);
  for (;;) {
    MgLevel& F = st->mg.back();
    const int fx = (F.nx % 2 == 0 && F.nx >= kMgMinSize) ? 2 : 1;
    const int fy = (F.ny % 2 == 0 && F.ny >= kMgMinSize) ? 2 : 1;
    if (fx == 1 && fy == 1) break;
    F.fx = fx;
    F.fy = fy;
    MgLevel C;
    C.nx = F.nx / fx;
    C.ny = F.ny / fy;
    C.nz = F.nz;
    cx *= fy;
    cy *= fx;
    C.deta = (float)(st->cfg.dtau / (2.0 * (cx + cy)));
    const size_t Nc = (size_t)C.nx * C.ny * C.nz;
    auto ac = [&](float** p) -> int {
      PCF_HIP(hipMalloc((void**)p, Nc * sizeof(float)));
      st->allocations.push_back(*p);
      return PCF_OK;
    };
    for (int t = 0; t < T_COUNT; ++t)
      if ((rc = ac(&C.a[t]))) return rc;
    float** v[] = {&C.P, &C.E, &C.Q, &C.b, &C.s, &C.rhs, &C.col};
    for (float** p : v)
      if ((rc = ac(p))) return rc;
    st->mg.push_back(C);
  }
  return PCF_OK;
}

}  // namespace pcf
