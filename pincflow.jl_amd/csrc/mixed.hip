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
// pressure correction), or (PCF_GUESS_EXTRAPOLATED) from the linear
// extrapolation in time of the last two; the criterion stays |r| / |b|.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "pcf_device.hpp"
#include "solver_common.hpp"

namespace pcf {

namespace {

typedef float flt4 __attribute__((ext_vector_type(4)));

// CT: float, or uint16_t holding bf16 bits (the upper half of an FP32)
// Memory layout of a level's arrays: x fastest (the finest level and the
// levels smoothed with row-parallel kernels) or z fastest (the coarse levels
// smoothed one column per wavefront, whose accesses are then contiguous
// across the lanes of a wavefront).
struct Lay {
  int nx, ny, nz;
  bool zf;
  __host__ __device__ size_t at(int i, int j, int k) const {
    return zf ? (size_t)k + (size_t)nz * ((size_t)i + (size_t)nx * j)
              : (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
  }
  // 32-bit form for the levels of at most kColumnCols x kScanMaxNz cells
  __device__ int at32(int i, int j, int k) const { return zf ? k + nz * (i + nx * j) : i + nx * (j + ny * k); }
  __host__ __device__ size_t count() const { return (size_t)nx * ny * nz; }
  // the cell of the t-th element in memory order
  __device__ void cell(size_t t, int& i, int& j, int& k) const {
    if (zf) {
      k = (int)(t % nz);
      const size_t c = t / nz;
      i = (int)(c % nx);
      j = (int)(c / nx);
    } else {
      i = (int)(t % nx);
      const size_t r = t / nx;
      j = (int)(r % ny);
      k = (int)(r / ny);
    }
  }
};

template <typename CT>
struct PcArgs {
  const CT* a[T_COUNT];  // the 22 horizontal coefficients are used here
  const float* s;
  float* out;
  int nx, ny, nz;
  // coarse-level correction added to s first (k_pc_rhs32<..., PRO = true>):
  // s + omega e at the coarse cell (i >> sx, j >> sy, k) in layout cl
  const float* e;
  float omega;
  int sx, sy;
  Lay cl;
};
using PcArgs32 = PcArgs<float>;

template <typename CT>
struct Coefs {
  const CT* a[T_COUNT];
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

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// W consecutive coefficients, streamed (not kept in the caches), as FP32.
template <int W, typename CT>
__device__ __forceinline__ void ld_coef(const CT* p, float (&o)[W]) {
  if constexpr (sizeof(CT) == 4) {
    ld_stream<W>((const float*)p, o);
  } else if constexpr (W == 4) {
    const u32x2 v = __builtin_nontemporal_load((const u32x2*)p);
    o[0] = __uint_as_float(v.x << 16);
    o[1] = __uint_as_float(v.x & 0xffff0000u);
    o[2] = __uint_as_float(v.y << 16);
    o[3] = __uint_as_float(v.y & 0xffff0000u);
  } else {
    o[0] = __uint_as_float((unsigned)__builtin_nontemporal_load(p) << 16);
  }
}

// The later pseudo-time steps of the preconditioner in FP32:
// out = s + deta (H s - sin) with H the horizontal 22-point part of the
// operator (apply_preconditioner!.jl:62-63); sin: the FP64 or FP32 input.
// W x-adjacent cells per thread, block = (64 / W) x (4 W) threads covering
// 64 x (4 W) cells of one z-level; CT: the coefficients' storage.  PRO: the
// iterate is s + omega P e, the multigrid prolongation of the coarse-level
// correction fused into the first post-smoothing step.
// The body takes the block's index, so that the ensemble kernel can run it
// for a block of any member.
template <int W, typename SinT, typename CT = float, bool PRO = false>
__device__ __forceinline__ void pc_rhs32(const PcArgs<CT>& A, const SinT* __restrict__ sin, float deta, int bx,
                                         int by, int bz) {
  constexpr int BX = 64 / W;
  const int i0 = (bx * BX + threadIdx.x) * W;
  const int j = by * (4 * W) + threadIdx.y;
  const int k = bz;
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
    if constexpr (PRO) {
      auto E = [&](int ii, int jj) -> float { return A.omega * A.e[A.cl.at(ii >> A.sx, jj >> A.sy, kk)]; };
      if (W == 4 && A.sx == 1 && !A.cl.zf) {
        // cells i0, i0 + 1 and i0 + 2, i0 + 3 share the coarse cells i0 / 2
        // and i0 / 2 + 1: one 8-byte load per row
        auto E2 = [&](int jj) -> float2 { return *(const float2*)(A.e + A.cl.at(i0 >> 1, jj >> A.sy, kk)); };
        const float2 ec = E2(j), ef = E2(jp), eb = E2(jm);
#pragma unroll
        for (int c = 0; c < W; ++c) {
          const float o = A.omega;
          C[w][c] += o * (c < 2 ? ec.x : ec.y);
          F[w][c] += o * (c < 2 ? ef.x : ef.y);
          B[w][c] += o * (c < 2 ? eb.x : eb.y);
        }
      } else {
#pragma unroll
        for (int c = 0; c < W; ++c) {
          C[w][c] += E(i0 + c, j);
          F[w][c] += E(i0 + c, jp);
          B[w][c] += E(i0 + c, jm);
        }
      }
      L[w] += E(im, j);
      R[w] += E(ip, j);
    }
  }
  float a[T_COUNT][W];
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) {
    if (t == T_AC || t == T_AD || t == T_AU) continue;
    ld_coef<W, CT>(A.a[t] + o, a[t]);
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
template <int W, typename SinT, typename CT = float, bool PRO = false>
__global__ void __launch_bounds__(256) k_pc_rhs32(PcArgs<CT> A, const SinT* __restrict__ sin, float deta,
                                                  const SolverScalars* sc) {
  if (sc && sc->done) return;
  pc_rhs32<W, SinT, CT, PRO>(A, sin, deta, blockIdx.x, blockIdx.y, blockIdx.z);
}

// The preconditioner's line solve in z with FP32 factors (the recurrence of
// the FP64 k_pc_thomas); SrcT / OutT: FP64 or FP32 input and output vectors.
// One column per thread, blockDim.x = 64.  The upward sweep keeps its values
// of the lowest NLDS levels in LDS (NLDS x 64 x 4 B per wavefront) and those
// of higher levels in the FP32 scratch `col`.  FIRST: rhs = deta (0 - src).
// The compiler barriers keep each batch of loads ahead of its recurrence
// (without LDS, the possible aliasing of `col` does).  LDS pays where a level
// has enough columns for several wavefronts per CU (kThomas32LdsColumns);
// with fewer, the LDS round trips are not hidden and NLDS = 0 is faster.
constexpr int kThomas32Lds = 128;
constexpr size_t kThomas32LdsColumns = 32768;

template <bool FIRST, typename SrcT, typename OutT, int NLDS = kThomas32Lds>
__device__ __forceinline__ void pc_thomas32(const float* P, const float* E, const float* Q, const SrcT* src, OutT* out,
                                            float* col, int nx, int ny, int nz, float deta, int bx, int by) {
  __shared__ float up[NLDS > 0 ? NLDS : 1][64];
  constexpr int U = 8;
  const int tid = threadIdx.x;
  const int i = bx * 64 + tid;
  const int j = by;
  if (i >= nx) return;
  const size_t sxy = (size_t)nx * ny;
  const size_t c = (size_t)i + (size_t)nx * j;
  auto rhs_of = [&](SrcT v) -> float { return FIRST ? deta * (0.0f - (float)v) : (float)v; };
  auto keep = [&](int k, float v) {
    if (k < NLDS)
      up[k][tid] = v;
    else
      col[c + sxy * k] = v;
  };
  auto kept = [&](int k) -> float { return k < NLDS ? up[k][tid] : col[c + sxy * k]; };
  float sk = rhs_of(src[c]) / P[c];
  keep(0, sk);
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
    if constexpr (NLDS > 0) asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sk = (r[u] + e[u] * sk) * p[u];
      keep(k + u, sk);
    }
  }
  for (; k < nz; ++k) {
    const size_t o = c + sxy * k;
    sk = (rhs_of(src[o]) + E[o] * sk) * P[o];
    keep(k, sk);
  }
  out[c + sxy * (nz - 1)] = (OutT)sk;
  k = nz - 2;
  for (; k - (U - 1) >= 0; k -= U) {
    float r[U], q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = kept(k - u);
      q[u] = Q[c + sxy * (k - u)];
    }
    if constexpr (NLDS > 0) asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sk = r[u] + q[u] * sk;
      out[c + sxy * (k - u)] = (OutT)sk;
    }
  }
  for (; k >= 0; --k) {
    sk = kept(k) + Q[c + sxy * k] * sk;
    out[c + sxy * k] = (OutT)sk;
  }
}
template <bool FIRST, typename SrcT, typename OutT, int NLDS = kThomas32Lds>
__global__ void __launch_bounds__(64) k_pc_thomas32(const float* P, const float* E, const float* Q, const SrcT* src,
                                                    OutT* out, float* col, int nx, int ny, int nz, float deta,
                                                    const SolverScalars* sc) {
  if (sc && sc->done) return;
  pc_thomas32<FIRST, SrcT, OutT, NLDS>(P, E, Q, src, out, col, nx, ny, nz, deta, blockIdx.x, blockIdx.y);
}

__global__ void k_to_f32(const double* a, float* b, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// g = q1 + r (q1 - q2): the linear extrapolation in time of the last two
// solutions per unit time step (r: this step over the previous one)
__global__ void k_extrapolate(const double* q1, const double* q2, double* g, size_t n, double r) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    g[q] = q1[q] + r * (q1[q] - q2[q]);
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
constexpr float kMgOmega = 1.5f;     // over-correction of the coarse-level correction
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

// R A P: one thread per coarse cell, in the coarse level's memory order.
template <int FX, int FY>
__global__ void k_mg_coarsen(MgCoefs fine, MgCoefs coarse, Lay fl, Lay cl) {
  const size_t t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t0 >= cl.count()) return;
  int I, J, k;
  cl.cell(t0, I, J, k);
  constexpr int oi[T_COUNT] = {0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0};
  constexpr int oj[T_COUNT] = {0, 0, 0, -1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1};
  constexpr int ok[T_COUNT] = {0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -2, 2, -2, 2, -2, 2, -2, 2, -2, 2};
  float acc[T_COUNT];
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) acc[t] = 0.0f;
#pragma unroll
  for (int b = 0; b < FY; ++b)
#pragma unroll
    for (int a = 0; a < FX; ++a) {
      const size_t o = fl.at(FX * I + a, FY * J + b, k);
#pragma unroll
      for (int t = 0; t < T_COUNT; ++t) {
        const int target = coef_of(floor_div(a + oi[t], FX), floor_div(b + oj[t], FY), ok[t]);
        acc[target] += fine.a[t][o];
      }
    }
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) coarse.a[t][t0] = acc[t];
}

// Line-solve factors (see k_pc_factors) of a coarse operator, in FP32: one
// thread per column.
__global__ void __launch_bounds__(64) k_mg_factors(const float* ac, const float* ad, const float* au, float* P, float* E,
                                                   float* Q, Lay l, float deta) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= l.nx * l.ny) return;
  const int i = c % l.nx, j = c / l.nx;
  size_t o = l.at(i, j, 0);
  const float denom = 1 - deta * ac[o];
  float qk = deta * au[o] / denom;
  P[o] = denom;
  E[o] = deta * ad[o];
  Q[o] = qk;
  for (int k = 1; k < l.nz; ++k) {
    o = l.at(i, j, k);
    const float e = deta * ad[o];
    const float p = 1 / (1 - deta * ac[o] - e * qk);
    qk = deta * au[o] * p;
    P[o] = p;
    E[o] = e;
    Q[o] = qk;
  }
}

// r = b - A s on one level, one thread per cell in memory order.
template <typename BT>
__global__ void __launch_bounds__(256) k_mg_resid(MgCoefs A, const BT* __restrict__ b, const float* __restrict__ s,
                                                  float* __restrict__ r, Lay l, const SolverScalars* sc) {
  if (sc && sc->done) return;
  const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (o >= l.count()) return;
  int i, j, k;
  l.cell(o, i, j, k);
  constexpr int oi[T_COUNT] = {0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0};
  constexpr int oj[T_COUNT] = {0, 0, 0, -1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1};
  constexpr int ok[T_COUNT] = {0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -2, 2, -2, 2, -2, 2, -2, 2, -2, 2};
  float v = 0.0f;
#pragma unroll
  for (int t = 0; t < T_COUNT; ++t) {
    const int kk = k + ok[t];
    if (kk < 0 || kk >= l.nz) continue;
    int ii = i + oi[t], jj = j + oj[t];
    ii = ii < 0 ? ii + l.nx : (ii >= l.nx ? ii - l.nx : ii);
    jj = jj < 0 ? jj + l.ny : (jj >= l.ny ? jj - l.ny : jj);
    v += A.a[t][o] * s[l.at(ii, jj, kk)];
  }
  r[o] = (float)b[o] - v;
}

// b_coarse = R r: one thread per coarse cell (coarse memory order) sums its
// fx x fy fine cells.
template <int FX, int FY>
__global__ void k_mg_sum(const float* __restrict__ r, float* __restrict__ bc, Lay fl, Lay cl,
                         const SolverScalars* sc) {
  if (sc && sc->done) return;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= cl.count()) return;
  int I, J, k;
  cl.cell(t, I, J, k);
  float sum = 0.0f;
#pragma unroll
  for (int bb = 0; bb < FY; ++bb)
#pragma unroll
    for (int a = 0; a < FX; ++a) sum += r[fl.at(FX * I + a, FY * J + bb, k)];
  bc[t] = sum;
}

// b_coarse = R (b - A s) for 2 x 2 aggregation and nx % 4 == 0: block
// 16 x 16 threads, 4 x-adjacent fine cells per thread (the loads of
// k_pc_rhs32); the x pairs are summed in the thread, the y pairs across the
// lanes l and l + 16 of a wavefront (rows j and j + 1).
template <typename BT, typename CT>
__global__ void __launch_bounds__(256) k_resid_restrict4(Coefs<CT> A, const BT* __restrict__ b,
                                                         const float* __restrict__ s, float* __restrict__ bc, int nx,
                                                         int ny, int nz, Lay cl, const SolverScalars* sc) {
  if (sc && sc->done) return;
  constexpr int W = 4;
  const int i0 = (blockIdx.x * 16 + threadIdx.x) * W;
  const int j = blockIdx.y * 16 + threadIdx.y;
  const int k = blockIdx.z;
  const bool act = i0 < nx && j < ny;
  float c0 = 0.0f, c1 = 0.0f;
  if (act) {
    const size_t sxy = (size_t)nx * ny;
    const size_t o = (size_t)i0 + (size_t)nx * j + sxy * k;
    const int im = i0 == 0 ? nx - 1 : i0 - 1, ip = i0 + W == nx ? 0 : i0 + W;
    const int jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
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
    for (int t = 0; t < T_COUNT; ++t) ld_coef<W, CT>(A.a[t] + o, a[t]);
    float r[W];
#pragma unroll
    for (int c = 0; c < W; ++c) {
      auto XL = [&](int w) -> float { return c == 0 ? L[w] : C[w][c - 1]; };
      auto XR = [&](int w) -> float { return c == W - 1 ? R[w] : C[w][c + 1]; };
      auto X0 = [&](int w) -> float { return C[w][c]; };
      auto YF = [&](int w) -> float { return F[w][c]; };
      auto YB = [&](int w) -> float { return B[w][c]; };
      float v = a[T_AL][c] * XL(2) + a[T_AR][c] * XR(2) + a[T_AF][c] * YF(2) + a[T_AB][c] * YB(2) +
                a[T_AU][c] * X0(3) + a[T_AD][c] * X0(1) + a[T_AC][c] * X0(2);
      v += a[T_ARU][c] * XR(3) + a[T_ARD][c] * XR(1) + a[T_ALU][c] * XL(3) + a[T_ALD][c] * XL(1) +
           a[T_AFU][c] * YF(3) + a[T_AFD][c] * YF(1) + a[T_ABU][c] * YB(3) + a[T_ABD][c] * YB(1) +
           a[T_AUU][c] * X0(4) + a[T_ADD][c] * X0(0) + a[T_ARUU][c] * XR(4) + a[T_ARDD][c] * XR(0) +
           a[T_ALUU][c] * XL(4) + a[T_ALDD][c] * XL(0) + a[T_AFUU][c] * YF(4) + a[T_AFDD][c] * YF(0) +
           a[T_ABUU][c] * YB(4) + a[T_ABDD][c] * YB(0);
      r[c] = (float)b[o + c] - v;
    }
    c0 = r[0] + r[1];
    c1 = r[2] + r[3];
  }
  // every lane takes part in the exchange; lane + 16 is row j + 1 of the same x
  const float u0 = __shfl_down(c0, 16), u1 = __shfl_down(c1, 16);
  if (act && (j & 1) == 0) {
    if (cl.zf) {
      bc[cl.at(i0 / 2, j / 2, k)] = c0 + u0;
      bc[cl.at(i0 / 2 + 1, j / 2, k)] = c1 + u1;
    } else {
      float2 v;
      v.x = c0 + u0;
      v.y = c1 + u1;
      *(float2*)(bc + cl.at(i0 / 2, j / 2, k)) = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Line solves with one wavefront per column.  The z recurrences of the line
// solve are affine maps, s_k = fa_k + fb_k s_(k-1) upwards and
// y_k = col_k + q_k y_(k+1) downwards; lane l holds NPL consecutive levels,
// composes its maps and an inclusive scan over the 64 lanes (shuffles)
// composes the rest, so a column costs one round trip to memory instead of
// nz dependent steps.  Used where there are too few columns to hide the
// latency of the thread-per-column sweep.
// ---------------------------------------------------------------------------
template <int NPL>
__device__ __forceinline__ void scan_line(const float (&fa)[NPL], const float (&fb)[NPL], const float (&q)[NPL],
                                          int nz, float (&y)[NPL]) {
  const int lane = threadIdx.x & 63;
  float A = 0.0f, B = 1.0f;
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    A = fa[u] + fb[u] * A;
    B = fb[u] * B;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const float Ap = __shfl_up(A, off), Bp = __shfl_up(B, off);
    if (lane >= off) {
      A = A + B * Ap;
      B = B * Bp;
    }
  }
  float sv = __shfl_up(A, 1);
  if (lane == 0) sv = 0.0f;  // the map of level 0 does not depend on it
  float col[NPL];
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    sv = fa[u] + fb[u] * sv;
    col[u] = NPL * lane + u < nz ? sv : 0.0f;
  }
  float Cc = 0.0f, D = 1.0f;
#pragma unroll
  for (int u = NPL - 1; u >= 0; --u) {
    Cc = col[u] + q[u] * Cc;
    D = q[u] * D;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const float Cn = __shfl_down(Cc, off), Dn = __shfl_down(D, off);
    if (lane + off < 64) {
      Cc = Cc + D * Cn;
      D = D * Dn;
    }
  }
  float t = __shfl_down(Cc, 1);
  if (lane == 63) t = 0.0f;  // above the top level: q = 0 there
#pragma unroll
  for (int u = NPL - 1; u >= 0; --u) {
    t = col[u] + q[u] * t;
    y[u] = t;
  }
}

// The forward map of level k of a column from its line-solve factors and
// right-hand side r (k = 0: P holds the pivot itself); identity maps above nz.
template <typename IdxT>
__device__ __forceinline__ void line_maps(const float* P, const float* E, const float* Q, IdxT o, int k, int nz,
                                          float r, float& fa, float& fb, float& q) {
  if (k >= nz) {
    fa = 0.0f;
    fb = 1.0f;
    q = 1.0f;
    return;
  }
  const float p = P[o];
  if (k == 0) {
    fa = r / p;
    fb = 0.0f;
  } else {
    fa = r * p;
    fb = E[o] * p;
  }
  q = k == nz - 1 ? 0.0f : Q[o];
}

// Line solve of ncols columns, block 256 = 4 columns.  FIRST: rhs = deta (0 - src).
template <int NPL, bool FIRST, typename SrcT, typename OutT>
__global__ void __launch_bounds__(256) k_line_scan(const float* __restrict__ P, const float* __restrict__ E,
                                                   const float* __restrict__ Q, const SrcT* __restrict__ src,
                                                   OutT* __restrict__ out, int ncols, int nz, float deta,
                                                   const SolverScalars* sc) {
  if (sc && sc->done) return;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncols) return;  // uniform over the wavefront
  const int lane = threadIdx.x & 63;
  float fa[NPL], fb[NPL], q[NPL], y[NPL];
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    const int k = NPL * lane + u;
    const size_t o = (size_t)c + (size_t)ncols * k;
    float r = 0.0f;
    if (k < nz) r = FIRST ? deta * (0.0f - (float)src[o]) : (float)src[o];
    line_maps(P, E, Q, o, k, nz, r, fa[u], fb[u], q[u]);
  }
  scan_line<NPL>(fa, fb, q, nz, y);
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    const int k = NPL * lane + u;
    if (k < nz) out[(size_t)c + (size_t)ncols * k] = (OutT)y[u];
  }
}

// One pseudo-time step of column (i, j): the right-hand side
// s + deta (H s - b) at the lane's levels (FIRST: deta (0 - b), from s = 0),
// then the line solve.  rd(i, j, k) reads the current iterate; the level's
// arrays are in layout l.
template <int NPL, bool FIRST, typename BT, class Rd>
__device__ __forceinline__ void column_step(const MgCoefs& A, const float* P, const float* E, const float* Q,
                                            const BT* b, Rd rd, const Lay& l, int i, int j, float deta,
                                            float (&y)[NPL]) {
  const int lane = threadIdx.x & 63;
  const int nx = l.nx, ny = l.ny, nz = l.nz;
  const int im = i == 0 ? nx - 1 : i - 1, ip = i + 1 == nx ? 0 : i + 1;
  const int jm = j == 0 ? ny - 1 : j - 1, jp = j + 1 == ny ? 0 : j + 1;
  float fa[NPL], fb[NPL], q[NPL];
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    const int k = NPL * lane + u;
    const int o = l.at32(i, j, k < nz ? k : nz - 1);
    float r = 0.0f;
    if (k < nz) {
      if (FIRST) {
        r = deta * (0.0f - (float)b[o]);
      } else {
        auto S = [&](int ii, int jj, int kk) -> float { return kk < 0 || kk >= nz ? 0.0f : rd(ii, jj, kk); };
        auto T = [&](int t) -> float { return A.a[t][o]; };
        float v = T(T_AL) * S(im, j, k) + T(T_AR) * S(ip, j, k) + T(T_AF) * S(i, jp, k) + T(T_AB) * S(i, jm, k);
        v += T(T_ARU) * S(ip, j, k + 1) + T(T_ARD) * S(ip, j, k - 1) + T(T_ALU) * S(im, j, k + 1) +
             T(T_ALD) * S(im, j, k - 1) + T(T_AFU) * S(i, jp, k + 1) + T(T_AFD) * S(i, jp, k - 1) +
             T(T_ABU) * S(i, jm, k + 1) + T(T_ABD) * S(i, jm, k - 1) + T(T_AUU) * S(i, j, k + 2) +
             T(T_ADD) * S(i, j, k - 2) + T(T_ARUU) * S(ip, j, k + 2) + T(T_ARDD) * S(ip, j, k - 2) +
             T(T_ALUU) * S(im, j, k + 2) + T(T_ALDD) * S(im, j, k - 2) + T(T_AFUU) * S(i, jp, k + 2) +
             T(T_AFDD) * S(i, jp, k - 2) + T(T_ABUU) * S(i, jm, k + 2) + T(T_ABDD) * S(i, jm, k - 2);
        r = S(i, j, k) + deta * (v - (float)b[o]);
      }
    }
    line_maps(P, E, Q, o, k, nz, r, fa[u], fb[u], q[u]);
  }
  scan_line<NPL>(fa, fb, q, nz, y);
}

// One smoothing step of a level with few columns, one wavefront per column
// (block 256 = 4 columns); s_old and out must differ.  PRO: the iterate is
// s_old + omega P e (e: the coarse-level correction in layout cl,
// fx = 1 << sx, fy = 1 << sy).
template <int NPL, bool FIRST, bool PRO, typename BT, typename OutT>
__global__ void __launch_bounds__(256) k_col_smooth(MgCoefs A, const float* __restrict__ P,
                                                    const float* __restrict__ E, const float* __restrict__ Q,
                                                    const BT* __restrict__ b, const float* __restrict__ s_old,
                                                    const float* __restrict__ e, float omega, int sx, int sy, Lay l,
                                                    Lay cl, OutT* __restrict__ out, float deta,
                                                    const SolverScalars* sc) {
  if (sc && sc->done) return;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= l.nx * l.ny) return;  // uniform over the wavefront
  const int i = c % l.nx, j = c / l.nx;
  auto rd = [&](int ii, int jj, int kk) -> float {
    float v = s_old[l.at32(ii, jj, kk)];
    if (PRO) v += omega * e[cl.at32(ii >> sx, jj >> sy, kk)];
    return v;
  };
  float y[NPL];
  column_step<NPL, FIRST, BT>(A, P, E, Q, b, rd, l, i, j, deta, y);
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    const int k = NPL * lane + u;
    if (k < l.nz) out[l.at32(i, j, k)] = (OutT)y[u];
  }
}

// All smoothing steps of the coarsest level in one workgroup of up to 16
// wavefronts, wavefront w smoothing the columns w, w + 16, ... (nx ny <= 16),
// the iterate ping-ponging between two z-fastest LDS buffers (2 nx ny nz
// floats of dynamic shared memory), so a column's new values are stored as
// soon as they are computed.
template <int NPL, typename BT, typename OutT>
__global__ void __launch_bounds__(1024) k_coarse_solve(MgCoefs A, const float* __restrict__ P,
                                                       const float* __restrict__ E, const float* __restrict__ Q,
                                                       const BT* __restrict__ b, OutT* __restrict__ out, Lay l,
                                                       float deta, int steps, const SolverScalars* sc) {
  if (sc && sc->done) return;
  extern __shared__ float lds[];
  const int nx = l.nx, nz = l.nz;
  const int ncols = l.nx * l.ny, n = ncols * nz;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int step = 0; step < steps; ++step) {
    const float* cur = lds + (step & 1) * n;
    float* nxt = lds + ((step + 1) & 1) * n;
    auto rd = [&](int ii, int jj, int kk) -> float { return cur[kk + nz * (ii + nx * jj)]; };
#pragma unroll 1
    for (int c = w; c < ncols; c += nw) {  // uniform over the wavefront
      float y[NPL];
      if (step == 0)
        column_step<NPL, true, BT>(A, P, E, Q, b, rd, l, c % nx, c / nx, deta, y);
      else
        column_step<NPL, false, BT>(A, P, E, Q, b, rd, l, c % nx, c / nx, deta, y);
#pragma unroll
      for (int u = 0; u < NPL; ++u) {
        const int k = NPL * lane + u;
        if (k < nz) {
          if (step + 1 < steps)
            nxt[k + nz * c] = y[u];
          else
            out[l.at32(c % nx, c / nx, k)] = (OutT)y[u];
        }
      }
    }
    __syncthreads();  // the step's writes before the next step's reads
  }
}

Lay lay_of(const MgLevel& L) { return Lay{L.nx, L.ny, L.nz, L.zf}; }

bool bf16_storage(const pcf_state* st) { return st->cfg.preconditioner_storage == PCF_STORAGE_BF16; }

// out = s + deta (H s - sin) on an x-fastest level with FP32 coefficients
// a32; the finest level reads its coefficients in the configured storage.
template <typename SinT>
void pc_rhs(pcf_state* st, float* const* a32, const uint16_t* const* a16, const float* s, float* out,
            const SinT* sin, int nx, int ny, int nz, float deta, const SolverScalars* sc, const float* e = nullptr,
            const MgLevel* F = nullptr, const MgLevel* C = nullptr) {
  ProfScope prof(st, PCF_K_PC_RHS);
  const bool quads = nx % 4 == 0;
  const dim3 grd = quads ? dim3(blocks_for(nx, 64), blocks_for(ny, 16), nz)
                         : dim3(blocks_for(nx, 64), blocks_for(ny, 4), nz);
  const dim3 blk = quads ? dim3(16, 16) : dim3(64, 4);
  auto go = [&](auto A) {
    using CT = std::remove_const_t<std::remove_pointer_t<std::remove_reference_t<decltype(A.a[0])>>>;
    A.s = s;
    A.out = out;
    A.nx = nx;
    A.ny = ny;
    A.nz = nz;
    A.e = e;
    A.omega = kMgOmega;
    A.sx = F && F->fx == 2;
    A.sy = F && F->fy == 2;
    A.cl = C ? lay_of(*C) : Lay{1, 1, 1, false};
    if (e) {
      if (quads)
        hipLaunchKernelGGL((k_pc_rhs32<4, SinT, CT, true>), grd, blk, 0, st->stream, A, sin, deta, sc);
      else
        hipLaunchKernelGGL((k_pc_rhs32<1, SinT, CT, true>), grd, blk, 0, st->stream, A, sin, deta, sc);
    } else {
      if (quads)
        hipLaunchKernelGGL((k_pc_rhs32<4, SinT, CT>), grd, blk, 0, st->stream, A, sin, deta, sc);
      else
        hipLaunchKernelGGL((k_pc_rhs32<1, SinT, CT>), grd, blk, 0, st->stream, A, sin, deta, sc);
    }
  };
  if (a16 && bf16_storage(st)) {
    PcArgs<uint16_t> A;
    for (int t = 0; t < T_COUNT; ++t) A.a[t] = a16[t];
    go(A);
  } else {
    PcArgs<float> A;
    for (int t = 0; t < T_COUNT; ++t) A.a[t] = a32[t];
    go(A);
  }
}

__global__ void k_to_bf16(const float* a, uint16_t* b, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const unsigned u = __float_as_uint(a[q]);
    b[q] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
  }
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
  if (bf16_storage(st))
    for (int t = 0; t < T_COUNT; ++t)
      hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, st->stream, st->T32[t], st->T16[t], N);
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
    const Lay fl = lay_of(F), cl = lay_of(C);
    const dim3 grd(blocks_for(cl.count(), 256)), blk(256);
    if (F.fx == 2 && F.fy == 2)
      hipLaunchKernelGGL((k_mg_coarsen<2, 2>), grd, blk, 0, st->stream, fa, ca, fl, cl);
    else if (F.fx == 2)
      hipLaunchKernelGGL((k_mg_coarsen<2, 1>), grd, blk, 0, st->stream, fa, ca, fl, cl);
    else
      hipLaunchKernelGGL((k_mg_coarsen<1, 2>), grd, blk, 0, st->stream, fa, ca, fl, cl);
    hipLaunchKernelGGL(k_mg_factors, dim3(blocks_for((long long)C.nx * C.ny, 64)), dim3(64), 0, st->stream,
                       C.a[T_AC], C.a[T_AD], C.a[T_AU], C.P, C.E, C.Q, cl, C.deta);
    if (C.a16[0])
      for (int t = 0; t < T_COUNT; ++t)
        hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, st->stream, C.a[t], C.a16[t], cl.count());
  }
  PCF_HIP(hipGetLastError());
  st->t32_valid = true;
  return PCF_OK;
}

// Level paths.  Levels with at most kColumnCols columns smooth with one
// fused wavefront-per-column kernel per step; larger ones with the
// row-parallel horizontal kernel and a line solve, the latter one wavefront
// per column up to kScanCols columns and one thread per column beyond (where
// the columns alone fill the chip).  The coarsest level runs all its steps
// in one workgroup when it has at most 16 columns.
constexpr size_t kColumnCols = 1024;
constexpr size_t kScanCols = 4096;
constexpr int kScanMaxNz = 512;
constexpr size_t kCoarseLdsFloats = 16384;

size_t columns(const MgLevel& L) { return (size_t)L.nx * L.ny; }
bool column_path(const MgLevel& L) { return columns(L) <= kColumnCols && L.nz <= kScanMaxNz; }

template <class F>
void with_npl(int nz, F&& f) {
  if (nz <= 64)
    f(std::integral_constant<int, 1>{});
  else if (nz <= 128)
    f(std::integral_constant<int, 2>{});
  else if (nz <= 256)
    f(std::integral_constant<int, 4>{});
  else
    f(std::integral_constant<int, 8>{});
}

MgCoefs coefs_of(const MgLevel& L) {
  MgCoefs A;
  for (int t = 0; t < T_COUNT; ++t) A.a[t] = L.a[t];
  return A;
}

// Line solve of a row-path level (FIRST: rhs = deta (0 - src)).
template <bool FIRST, typename SrcT, typename OutT>
void line_solve(pcf_state* st, const MgLevel& L, const SrcT* src, OutT* out, const SolverScalars* sc) {
  ProfScope prof(st, PCF_K_PC_THOMAS);
  const size_t nc = columns(L);
  if (nc <= kScanCols && L.nz <= kScanMaxNz) {
    with_npl(L.nz, [&](auto npl) {
      hipLaunchKernelGGL((k_line_scan<decltype(npl)::value, FIRST, SrcT, OutT>), dim3(blocks_for(nc, 4)), dim3(256),
                         0, st->stream, L.P, L.E, L.Q, src, out, (int)nc, L.nz, L.deta, sc);
    });
  } else {
    if (nc >= kThomas32LdsColumns)
      hipLaunchKernelGGL((k_pc_thomas32<FIRST, SrcT, OutT>), dim3(blocks_for(L.nx, 64), L.ny), dim3(64), 0,
                         st->stream, L.P, L.E, L.Q, src, out, L.col, L.nx, L.ny, L.nz, L.deta, sc);
    else
      hipLaunchKernelGGL((k_pc_thomas32<FIRST, SrcT, OutT, 0>), dim3(blocks_for(L.nx, 64), L.ny), dim3(64), 0,
                         st->stream, L.P, L.E, L.Q, src, out, L.col, L.nx, L.ny, L.nz, L.deta, sc);
  }
}

// steps smoothing steps on level L.  s: the buffer holding the current
// iterate, null to start from zero; e (with the coarser level C): the coarse
// correction, s += omega P e before the first step.  The last step writes to
// out64 when given (the FP64 result of the finest level); otherwise returns
// the buffer holding the result.
template <typename BT>
float* mg_smooth(pcf_state* st, const MgLevel& L, const BT* b, float* s, const float* e, const MgLevel* C, int steps,
                 double* out64, const SolverScalars* sc) {
  const bool col = column_path(L);
  const size_t nc = columns(L);
  int step = 0;
  if (!s) {  // first step from zero
    const bool last = steps == 1;
    if (col) {
      ProfScope prof(st, PCF_K_PC_THOMAS);
      with_npl(L.nz, [&](auto npl) {
        constexpr int NPL = decltype(npl)::value;
        const Lay l = lay_of(L);
        if (last && out64)
          hipLaunchKernelGGL((k_col_smooth<NPL, true, false, BT, double>), dim3(blocks_for(nc, 4)), dim3(256), 0,
                             st->stream, coefs_of(L), L.P, L.E, L.Q, b, nullptr, nullptr, 0.0f, 0, 0, l, l, out64,
                             L.deta, sc);
        else
          hipLaunchKernelGGL((k_col_smooth<NPL, true, false, BT, float>), dim3(blocks_for(nc, 4)), dim3(256), 0,
                             st->stream, coefs_of(L), L.P, L.E, L.Q, b, nullptr, nullptr, 0.0f, 0, 0, l, l, L.s,
                             L.deta, sc);
      });
    } else {
      if (last && out64)
        line_solve<true>(st, L, b, out64, sc);
      else
        line_solve<true>(st, L, b, L.s, sc);
    }
    s = L.s;
    step = 1;
  }
  // the row path adds the coarse correction in its first horizontal kernel
  const int first_post = step;
  for (; step < steps; ++step) {
    const bool last = step == steps - 1;
    if (col) {
      const bool pro = e && step == 0;
      float* dst = s == L.s ? L.rhs : L.s;
      ProfScope prof(st, PCF_K_PC_THOMAS);
      with_npl(L.nz, [&](auto npl) {
        constexpr int NPL = decltype(npl)::value;
        const int sx = L.fx == 2, sy = L.fy == 2;
        const Lay l = lay_of(L), cl = C ? lay_of(*C) : l;
        const dim3 g(blocks_for(nc, 4)), blk(256);
        if (last && out64) {
          if (pro)
            hipLaunchKernelGGL((k_col_smooth<NPL, false, true, BT, double>), g, blk, 0, st->stream, coefs_of(L), L.P,
                               L.E, L.Q, b, s, e, kMgOmega, sx, sy, l, cl, out64, L.deta, sc);
          else
            hipLaunchKernelGGL((k_col_smooth<NPL, false, false, BT, double>), g, blk, 0, st->stream, coefs_of(L), L.P,
                               L.E, L.Q, b, s, e, kMgOmega, sx, sy, l, cl, out64, L.deta, sc);
        } else {
          if (pro)
            hipLaunchKernelGGL((k_col_smooth<NPL, false, true, BT, float>), g, blk, 0, st->stream, coefs_of(L), L.P,
                               L.E, L.Q, b, s, e, kMgOmega, sx, sy, l, cl, dst, L.deta, sc);
          else
            hipLaunchKernelGGL((k_col_smooth<NPL, false, false, BT, float>), g, blk, 0, st->stream, coefs_of(L), L.P,
                               L.E, L.Q, b, s, e, kMgOmega, sx, sy, l, cl, dst, L.deta, sc);
        }
      });
      s = dst;
      continue;
    }
    pc_rhs<BT>(st, L.a, L.a16[0] ? L.a16 : nullptr, s, L.rhs, b, L.nx, L.ny, L.nz, L.deta, sc,
               step == first_post ? e : nullptr, &L, C);
    if (last && out64)
      line_solve<false>(st, L, (const float*)L.rhs, out64, sc);
    else
      line_solve<false>(st, L, (const float*)L.rhs, L.s, sc);
    s = L.s;
  }
  return s;
}

// b_coarse = R (b - A s)
template <typename BT>
void mg_restrict(pcf_state* st, const MgLevel& L, const MgLevel& C, const BT* b, const float* s,
                 const SolverScalars* sc) {
  ProfScope prof(st, PCF_K_PC_RHS);
  const MgCoefs A = coefs_of(L);
  if (!column_path(L) && L.fx == 2 && L.fy == 2 && L.nx % 4 == 0) {
    const dim3 grd(blocks_for(L.nx, 64), blocks_for(L.ny, 16), L.nz), blk(16, 16);
    if (L.a16[0] && bf16_storage(st)) {
      Coefs<uint16_t> A16;
      for (int t = 0; t < T_COUNT; ++t) A16.a[t] = L.a16[t];
      hipLaunchKernelGGL((k_resid_restrict4<BT, uint16_t>), grd, blk, 0, st->stream, A16, b, s, C.b, L.nx, L.ny,
                         L.nz, lay_of(C), sc);
    } else {
      Coefs<float> A32;
      for (int t = 0; t < T_COUNT; ++t) A32.a[t] = L.a[t];
      hipLaunchKernelGGL((k_resid_restrict4<BT, float>), grd, blk, 0, st->stream, A32, b, s, C.b, L.nx, L.ny, L.nz,
                         lay_of(C), sc);
    }
    return;
  }
  float* r = L.col;  // the line-solve scratch is free here
  const Lay fl = lay_of(L), cl = lay_of(C);
  hipLaunchKernelGGL((k_mg_resid<BT>), dim3(blocks_for(fl.count(), 256)), dim3(256), 0, st->stream, A, b, s, r, fl,
                     sc);
  const dim3 cgrd(blocks_for(cl.count(), 256)), blk(256);
  if (L.fx == 2 && L.fy == 2)
    hipLaunchKernelGGL((k_mg_sum<2, 2>), cgrd, blk, 0, st->stream, r, C.b, fl, cl, sc);
  else if (L.fx == 2)
    hipLaunchKernelGGL((k_mg_sum<2, 1>), cgrd, blk, 0, st->stream, r, C.b, fl, cl, sc);
  else
    hipLaunchKernelGGL((k_mg_sum<1, 2>), cgrd, blk, 0, st->stream, r, C.b, fl, cl, sc);
}

// The coarsest level: mg_coarse_steps smoothing steps from zero.
template <typename BT>
float* mg_coarsest(pcf_state* st, const MgLevel& L, const BT* b, double* out64, const SolverScalars* sc) {
  const size_t nc = columns(L);
  const size_t n = nc * (size_t)L.nz;
  if (nc <= 16 && L.nz <= 256 && 2 * n <= kCoarseLdsFloats) {
    ProfScope prof(st, PCF_K_PC_THOMAS);
    const unsigned lds = (unsigned)(2 * n * sizeof(float));
    const dim3 blk(64 * (unsigned)nc);
    with_npl(L.nz, [&](auto npl) {
      constexpr int NPL = decltype(npl)::value;
      if constexpr (NPL <= 4) {
        if (out64)
          hipLaunchKernelGGL((k_coarse_solve<NPL, BT, double>), dim3(1), blk, lds, st->stream, coefs_of(L), L.P, L.E,
                             L.Q, b, out64, lay_of(L), L.deta, st->cfg.mg_coarse_steps, sc);
        else
          hipLaunchKernelGGL((k_coarse_solve<NPL, BT, float>), dim3(1), blk, lds, st->stream, coefs_of(L), L.P, L.E,
                             L.Q, b, L.s, lay_of(L), L.deta, st->cfg.mg_coarse_steps, sc);
      }
    });
    return L.s;
  }
  return mg_smooth<BT>(st, L, b, nullptr, nullptr, nullptr, st->cfg.mg_coarse_steps, out64, sc);
}

// One V-cycle on level l for the right-hand side b; the finest level writes
// its result to out64; returns the buffer holding the level's result.
template <typename BT>
float* mg_vcycle(pcf_state* st, size_t l, const BT* b, double* out64, const SolverScalars* sc) {
  const MgLevel& L = st->mg[l];
  if (l + 1 == st->mg.size()) return mg_coarsest<BT>(st, L, b, out64, sc);
  const MgLevel& C = st->mg[l + 1];
  float* s = mg_smooth<BT>(st, L, b, nullptr, nullptr, nullptr, st->cfg.mg_smoothing_steps, nullptr, sc);
  mg_restrict<BT>(st, L, C, b, s, sc);
  const float* e = mg_vcycle<float>(st, l + 1, C.b, nullptr, sc);
  return mg_smooth<BT>(st, L, b, s, e, &C, st->cfg.mg_smoothing_steps, out64, sc);
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
      pc_rhs<double>(st, st->T32, st->T16[0] ? st->T16 : nullptr, cur, rhs, sin, g.nx, g.ny, g.nz, deta, sc);
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
  const bool extrap = n.initial_guess == PCF_GUESS_EXTRAPOLATED;
  const bool use_prev =
      (n.initial_guess == PCF_GUESS_PREVIOUS || n.initial_guess == PCF_GUESS_LEARNED || extrap) && stage;
  const double* prev = use_prev && st->m_prev_valid[slot] ? st->m_prev[slot] : nullptr;
  const double* guess = n.initial_guess == PCF_GUESS_PREVIOUS || extrap ? prev : nullptr;
  const size_t N = st->g.nv();
  if (extrap && prev && st->m_prev2_valid[slot] && st->m_prev_dt[slot] > 0.0) {
    hipLaunchKernelGGL(k_extrapolate, dim3(2048), dim3(256), 0, st->stream, prev, st->m_prev2[slot], st->m_guess, N,
                       dt / st->m_prev_dt[slot]);
    PCF_HIP(hipGetLastError());
    guess = st->m_guess;
  }
  double scale = dt;
  if (n.initial_guess == PCF_GUESS_LEARNED && stage) {
    if ((rc = learned_guess(st, dt, prev, &guess))) return rc;
    scale = 1.0;
  }
  rc = bicgstab_core(st, tolref, true, guess, scale, niter_out);
  if (rc == PCF_E_NOCONV && st->stage_net && st->stage_iterations > 0) {
    // the flexible solve with the learned stage failed: the plain
    // preconditioned BiCGSTAB from zero is the fallback
    const int failed = *niter_out;
    pcf_net* keep = st->stage_net;
    st->stage_net = nullptr;
    rc = bicgstab_core(st, tolref, true, nullptr, 0.0, niter_out);
    st->stage_net = keep;
    st->stage_fallbacks += 1;
    *niter_out += failed;
  }
  if (rc) return rc;
  if (use_prev) {  // keep solution / dt for the same stage of the next step
    if (extrap) {  // and the one before it
      std::swap(st->m_prev[slot], st->m_prev2[slot]);
      st->m_prev2_valid[slot] = st->m_prev_valid[slot];
    }
    hipLaunchKernelGGL(k_scale_copy, dim3(2048), dim3(256), 0, st->stream, st->solution, st->m_prev[slot], N,
                       1.0 / dt);
    PCF_HIP(hipGetLastError());
    st->m_prev_valid[slot] = true;
    st->m_prev_dt[slot] = dt;
  }
  return PCF_OK;
}

// ---------------------------------------------------------------------------
// Ensemble batch: the FP32 line preconditioner over the solves of a batched
// iteration (see the ensemble kernels of poisson.hip).  Each member runs the
// blocks of its single solve's kernels (precondition_mixed; x_size a
// multiple of 4).
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ int ens_member32(const int* start, int count, int b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// FIRST: from the FP64 vector v[in]; else from the member's FP32 rhs.
// OutT double: to v[out]; float: to the member's FP32 iterate.
template <bool FIRST, typename OutT>
__global__ void __launch_bounds__(64) kEm_thomas(const EnsTable* __restrict__ E, const EnsList* __restrict__ L, int in,
                                                 int out) {
  const int a = ens_member32(L->thomas_start, L->count, blockIdx.x);
  const EnsMember& M = E->m[L->member[a]];
  if (M.sc->done) return;
  const int b = blockIdx.x - L->thomas_start[a];
  const int gx = (M.nx + 63) / 64;
  OutT* dst;
  if constexpr (sizeof(OutT) == 8) dst = M.v[out];
  else dst = M.cur32;
  const float deta = (float)M.deta;
  if constexpr (FIRST)
    pc_thomas32<true, double, OutT>(M.P32, M.E32, M.Q32, (const double*)M.v[in], dst, M.col32, M.nx, M.ny, M.nz,
                                    deta, b % gx, b / gx);
  else
    pc_thomas32<false, float, OutT>(M.P32, M.E32, M.Q32, (const float*)M.rhs32, dst, M.col32, M.nx, M.ny, M.nz,
                                    deta, b % gx, b / gx);
}

// rhs32 = cur32 + deta (H cur32 - v[in]) with the coefficients in storage CT
template <typename CT>
__global__ void __launch_bounds__(256) kEm_rhs(const EnsTable* __restrict__ E, const EnsList* __restrict__ L, int in) {
  const int a = ens_member32(L->rhs32_start, L->count, blockIdx.x);
  const EnsMember& M = E->m[L->member[a]];
  if (M.sc->done) return;
  const int b = blockIdx.x - L->rhs32_start[a];
  const int gx = (M.nx + 63) / 64, gy = (M.ny + 15) / 16;
  PcArgs<CT> A;
  for (int t = 0; t < T_COUNT; ++t) {
    if constexpr (sizeof(CT) == 2) A.a[t] = M.T16[t];
    else A.a[t] = M.T32[t];
  }
  A.s = M.cur32;
  A.out = M.rhs32;
  A.nx = M.nx;
  A.ny = M.ny;
  A.nz = M.nz;
  A.e = nullptr;
  A.omega = 0.0f;
  A.sx = A.sy = 0;
  A.cl = Lay{1, 1, 1, false};
  pc_rhs32<4, double, CT>(A, (const double*)M.v[in], (float)M.deta, b % gx, (b / gx) % gy, b / (gx * gy));
}

}  // namespace

int ensemble_rhs32_blocks(const pcf_state* st) {
  if (st->cfg.precision != PCF_PRECISION_MIXED) return 0;
  const Geo& g = st->g;
  return (int)(blocks_for(g.nx, 64) * blocks_for(g.ny, 16) * g.nz);
}

// M^-1 v[in] -> v[out] for every solve of the list (precondition_mixed's
// line branch)
void ensemble_precondition_mixed(hipStream_t s, const EnsTable* dT, const EnsList* dL, const EnsList& L, int iters,
                                 bool bf16, int in, int out) {
  const int tb = L.thomas_start[L.count], rb = L.rhs32_start[L.count];
  if (iters == 1) {
    hipLaunchKernelGGL((kEm_thomas<true, double>), dim3(tb), dim3(64), 0, s, dT, dL, in, out);
    return;
  }
  hipLaunchKernelGGL((kEm_thomas<true, float>), dim3(tb), dim3(64), 0, s, dT, dL, in, out);
  for (int it = 2; it <= iters; ++it) {
    if (bf16) hipLaunchKernelGGL(kEm_rhs<uint16_t>, dim3(rb), dim3(16, 16), 0, s, dT, dL, in);
    else hipLaunchKernelGGL(kEm_rhs<float>, dim3(rb), dim3(16, 16), 0, s, dT, dL, in);
    if (it == iters) hipLaunchKernelGGL((kEm_thomas<false, double>), dim3(tb), dim3(64), 0, s, dT, dL, in, out);
    else hipLaunchKernelGGL((kEm_thomas<false, float>), dim3(tb), dim3(64), 0, s, dT, dL, in, out);
  }
}

// Allocation of the mixed-precision buffers (pcf_create, precision MIXED).
int alloc_mixed(pcf_state* st) {
  const size_t N = st->g.nv();
  auto af = [&](float** p) -> int { return device_alloc(st, p, N); };
  int rc;
  for (int t = 0; t < T_COUNT; ++t)
    if ((rc = af(&st->T32[t]))) return rc;
  if (st->cfg.preconditioner_storage == PCF_STORAGE_BF16)
    for (int t = 0; t < T_COUNT; ++t)
      if ((rc = device_alloc(st, &st->T16[t], N))) return rc;
  float** fl[] = {&st->pc32_P, &st->pc32_E, &st->pc32_Q, &st->m_pc_tmp, &st->m_pc_cur, &st->m_pc_rhs};
  for (float** p : fl)
    if ((rc = af(p))) return rc;
  double** dl[] = {&st->bzpc, &st->m_prev[0], &st->m_prev[1]};
  for (double** p : dl)
    if ((rc = device_alloc(st, p, N))) return rc;
  if (st->cfg.initial_guess == PCF_GUESS_EXTRAPOLATED) {
    double** el[] = {&st->m_prev2[0], &st->m_prev2[1], &st->m_guess};
    for (double** p : el)
      if ((rc = device_alloc(st, p, N))) return rc;
  }
  if (st->cfg.preconditioner_type != PCF_PC_MULTIGRID) return PCF_OK;
  // level 0 works on the buffers above; coarser levels get their own
  const Geo& g = st->g;
  MgLevel L0;
  L0.nx = g.nx;
  L0.ny = g.ny;
  L0.nz = g.nz;
  L0.deta = (float)st->deta;
  for (int t = 0; t < T_COUNT; ++t) {
    L0.a[t] = st->T32[t];
    L0.a16[t] = st->T16[t];
  }
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
    auto ac = [&](float** p) -> int { return device_alloc(st, p, Nc); };
    for (int t = 0; t < T_COUNT; ++t)
      if ((rc = ac(&C.a[t]))) return rc;
    float** v[] = {&C.P, &C.E, &C.Q, &C.b, &C.s, &C.rhs, &C.col};
    for (float** p : v)
      if ((rc = ac(p))) return rc;
    C.zf = column_path(C);  // coarse levels smoothed column by column: z fastest
    if (!C.zf && st->cfg.preconditioner_storage == PCF_STORAGE_BF16)
      for (int t = 0; t < T_COUNT; ++t)
        if ((rc = device_alloc(st, &C.a16[t], Nc))) return rc;
    st->mg.push_back(C);
  }
  return PCF_OK;
}

}  // namespace pcf
