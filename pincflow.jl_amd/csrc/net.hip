// The initial-guess network of the learned solve (no reference counterpart).
//
// Six 3 x 3 x 3 convolutions, 3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1 channels,
// GELU after the first five, bias on every layer; fields are periodic in x
// and y and zero beyond the walls in z, like the pressure operator.  Every
// layer's forward runs as an implicit GEMM on the bf16 matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation) over an input brick staged
// in LDS (k_brick); weights and activations are bf16, the network's input is
// rounded to bf16 on staging.  The backward pass reuses the 32 -> 32 kernel
// for the input gradient (transposed, tap-flipped weights) and has its own
// MFMA kernel for the weight gradient (a GEMM over the voxels, split into
// chunks whose partial sums are added in a fixed order); the weight
// gradients of the first and last layers and the last layer's input
// gradient run on the vector units from the FP32 master weights.
//
// Layout: activations are channels-last [voxel][channel] with voxel =
// i + nx (j + ny k); parameters are one FP32 vector, layer by layer W[27][cout][cin]
// then b[cout], tap t = (dx + 1) + 3 (dy + 1) + 9 (dz + 1) weighing the input
// at (i + dx, j + dy, k + dz).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pcf_internal.hpp"
#include "solver_common.hpp"

namespace pcf {

namespace {

typedef __bf16 bf16;
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kC = 32;        // hidden channels
constexpr int kLayers = 6;
constexpr int kChunk = 4096;  // voxels per partial sum of a weight gradient
constexpr int kCin[kLayers] = {3, 32, 32, 32, 32, 32};
constexpr int kCout[kLayers] = {32, 32, 32, 32, 32, 1};

__host__ __device__ constexpr int layer_size(int l) { return 27 * kCout[l] * kCin[l] + kCout[l]; }
__host__ __device__ constexpr int layer_offset(int l) {
  int o = 0;
  for (int q = 0; q < l; ++q) o += layer_size(q);
  return o;
}
constexpr int kParams = layer_offset(kLayers);

struct Dims {
  int nx, ny, nz;
  size_t n;
};

// Voxel index of the neighbour at tap t of voxel (i, j, k); -1 beyond the walls.
__device__ __forceinline__ long long neighbour(const Dims& d, int i, int j, int k, int t) {
  const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
  const int kk = k + dz;
  if (kk < 0 || kk >= d.nz) return -1;
  int ii = i + dx, jj = j + dy;
  ii = ii < 0 ? ii + d.nx : (ii >= d.nx ? ii - d.nx : ii);
  jj = jj < 0 ? jj + d.ny : (jj >= d.ny ? jj - d.ny : jj);
  return (long long)ii + (long long)d.nx * (jj + (long long)d.ny * kk);
}

// (i, j, k) of voxel p in 32-bit arithmetic: pcf_net_create admits at most
// 2^31 - 1 voxels
__device__ __forceinline__ void coords(const Dims& d, size_t p, int& i, int& j, int& k) {
  const unsigned q = (unsigned)p, nx = (unsigned)d.nx, ny = (unsigned)d.ny;
  const unsigned r = q / nx;
  i = (int)(q - r * nx);
  j = (int)(r % ny);
  k = (int)(r / ny);
}

// GELU x Phi(x) = x (1 + erf(x / sqrt 2)) / 2 with 1 + erf(z) from the
// rational approximation of Abramowitz and Stegun 7.1.26 (|error of erf| <=
// 1.5e-7): for z >= 0, 1 + erf(z) = 2 - q, for z < 0 it is q itself, with
// q = t (a1 + t (a2 + ...)) exp(-z^2), t = 1 / (1 + p |z|), so the negative
// side has no cancellation.  Branch-free, about a dozen instructions; the
// activations are rounded to bf16 (relative 3.9e-3) right after, far above
// its 6.2e-7 absolute error (measured against erf on [-12, 12]).
__device__ __forceinline__ float gelu(float x) {
  const float z = x * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * fabsf(z));
  float p = 1.061405429f;
  p = fmaf(p, t, -1.453152027f);
  p = fmaf(p, t, 1.421413741f);
  p = fmaf(p, t, -0.284496736f);
  p = fmaf(p, t, 0.254829592f);
  const float q = p * t * __builtin_amdgcn_exp2f(-(z * z) * 1.4426950408889634f);
  return 0.5f * x * (z >= 0.0f ? 2.0f - q : q);
}
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.39894228040143268f;
}

// ---- the convolutions on the matrix cores (v_mfma_f32_32x32x16_bf16, FP32
// accumulation).  Operand layout: lane l holds A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32], e = 0..7, and D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32],
// q = 0..15.  A row is the output channel, a column one of 32 voxels
// consecutive in x (nx is a multiple of 32), K runs over input channels (and,
// for the first layer, taps).
//
// A workgroup of 4 wavefronts computes a block of 32 (x) x 4 (y) x 4 (z)
// voxels.  Its input brick, the block with a one-voxel halo (34 x 6 x 6
// voxels, periodic in x and y, zero beyond the walls in z), is staged in LDS
// once as bf16; each wavefront computes 2 (y) x 2 (z) rows of 32 voxels, so
// one B operand read from LDS (the input row at one (y, z), shifted by dx)
// feeds the MFMAs of every output row of the wavefront whose 3 x 3 x 3
// neighbourhood holds that row (up to four).  The A operands of one (dx,
// channel half) -- nine taps -- stay in registers over the 16 input rows.
// Blocks are numbered so that each XCD (blocks are dealt to the 8 XCDs in
// turn) gets a contiguous range of neighbouring blocks, whose halos its L2
// then shares.
//
// KIND_FIRST: 3 -> 32, K = 27 taps x 3 channels padded to 96; the brick holds
//   the FP32 input rounded to bf16, 4 channels (8 B) per voxel.
// KIND_MID:   32 -> 32 (FWD: W x + bias -> z (FP32, may be null), gelu -> a
//   (bf16); DATA: W x -> g (FP32), the input gradient with transposed,
//   tap-flipped W).
// KIND_LAST:  32 -> 1, the 32 x 32 MFMA with only row 0 of A nonzero; out (FP32).
enum { KIND_FIRST = 0, KIND_MID = 1, KIND_LAST = 2 };
enum { MID_FWD = 0, MID_DATA = 1 };
constexpr int kBX = 34, kBY = 6, kBZ = 6, kBrickVox = kBX * kBY * kBZ;  // 1224 voxels
constexpr int kK1 = 96;  // padded K of the first layer

struct BrickGrid {
  int tiles_x, tiles_y, tiles_z;
};

// 16-byte chunk c (0..3) of brick voxel v, XOR-swizzled over neighbouring
// voxels so that 8 lanes reading one chunk of consecutive voxels hit distinct banks
__device__ __forceinline__ int brick_chunk(int v, int c) { return 4 * v + (c ^ ((v >> 1) & 3)); }

__device__ __forceinline__ int wrap(int q, int n) {
  q %= n;
  return q < 0 ? q + n : q;
}

__device__ __forceinline__ bf8 as_bf8(uint4 u) {
  bf8 r;
  __builtin_memcpy(&r, &u, sizeof(r));
  return r;
}

__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
  bf16 h[4] = {(bf16)a, (bf16)b, (bf16)c, (bf16)d};
  uint2 r;
  __builtin_memcpy(&r, h, sizeof(r));
  return r;
}

template <int KIND, int MODE>
__global__ void __launch_bounds__(256, 2) k_brick(Dims d, BrickGrid bg, const void* __restrict__ in_,
                                                  const bf16* __restrict__ W, const float* __restrict__ bias,
                                                  float* zout, bf16* aout, float* gout) {
  __shared__ uint4 brick[KIND == KIND_FIRST ? kBrickVox / 2 : kBrickVox * 4];
  const int nb = (int)gridDim.x;
  int b = (int)blockIdx.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int tx = b % bg.tiles_x, tyz = b / bg.tiles_x, ty = tyz % bg.tiles_y, tz = tyz / bg.tiles_y;
  const int x0 = 32 * tx, y0 = 4 * ty, z0 = 4 * tz;

  // ---- stage the input brick
  if (KIND == KIND_FIRST) {
    const float* in = static_cast<const float*>(in_);
    uint2* br = reinterpret_cast<uint2*>(brick);
    for (int v = threadIdx.x; v < kBrickVox; v += 256) {
      const int bx = v % kBX, byz = v / kBX, by = byz % kBY, bz = byz / kBY;
      const int gz = z0 + bz - 1;
      float x0v = 0.0f, x1v = 0.0f, x2v = 0.0f;
      if (gz >= 0 && gz < d.nz) {
        const size_t q = (size_t)wrap(x0 + bx - 1, d.nx) + (size_t)d.nx * (wrap(y0 + by - 1, d.ny) + (size_t)d.ny * gz);
        x0v = in[3 * q];
        x1v = in[3 * q + 1];
        x2v = in[3 * q + 2];
      }
      br[v] = pack4(x0v, x1v, x2v, 0.0f);
    }
  } else {
    const bf16* in = static_cast<const bf16*>(in_);
    constexpr int kChunks = kBrickVox * 4, kPer = (kChunks + 255) / 256;
    uint4 buf[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 256 * u;
      buf[u] = uint4{0u, 0u, 0u, 0u};
      if (e < kChunks) {
        const int v = e >> 2, c = e & 3;
        const int bx = v % kBX, byz = v / kBX, by = byz % kBY, bz = byz / kBY;
        const int gz = z0 + bz - 1;
        if (gz >= 0 && gz < d.nz) {
          const size_t q =
              (size_t)wrap(x0 + bx - 1, d.nx) + (size_t)d.nx * (wrap(y0 + by - 1, d.ny) + (size_t)d.ny * gz);
          buf[u] = *reinterpret_cast<const uint4*>(in + q * kC + 8 * c);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 256 * u;
      if (e < kChunks) brick[brick_chunk(e >> 2, e & 3)] = buf[u];
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, half = lane >> 5;
  const int wy = 2 * (wv & 1), wz = 2 * (wv >> 1);  // the wavefront's first output row in the block
  f16v acc[2][2];
#pragma unroll
  for (int oy = 0; oy < 2; ++oy)
#pragma unroll
    for (int oz = 0; oz < 2; ++oz) acc[oy][oz] = f16v{};

  if (KIND == KIND_FIRST) {
    // K index k = 3 t + ci; lane (col, half) of k-block kb holds k = 16 kb + 8 half + e
    const uint16_t* br = reinterpret_cast<const uint16_t*>(brick);
    bf8 wa[6];
#pragma unroll
    for (int kb = 0; kb < 6; ++kb) wa[kb] = *reinterpret_cast<const bf8*>(W + (size_t)col * kK1 + 16 * kb + 8 * half);
#pragma unroll
    for (int oy = 0; oy < 2; ++oy)
#pragma unroll
      for (int oz = 0; oz < 2; ++oz) {
        // brick voxel of the output voxel (col, wy + oy, wz + oz) at tap (0, 0, 0)
        const int vc = (col + 1) + kBX * ((wy + oy + 1) + kBY * (wz + oz + 1));
#pragma unroll
        for (int kb = 0; kb < 6; ++kb) {
          bf8 bv;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k0 = 16 * kb + e, k1 = k0 + 8;  // half 0 / half 1
            const int t0 = k0 / 3, c0 = k0 % 3, t1 = k1 / 3, c1 = k1 % 3;
            const int o0 = (t0 % 3 - 1) + kBX * ((t0 / 3) % 3 - 1) + kBX * kBY * (t0 / 9 - 1);
            const int o1 = (t1 % 3 - 1) + kBX * ((t1 / 3) % 3 - 1) + kBX * kBY * (t1 / 9 - 1);
            const bool z0k = k0 >= 81, z1k = k1 >= 81;
            const int off = half ? 4 * (vc + o1) + c1 : 4 * (vc + o0) + c0;
            const bool zero = half ? z1k : z0k;
            const uint16_t raw = br[zero ? 0 : off];
            bv[e] = zero ? (bf16)0.0f : __builtin_bit_cast(bf16, raw);
          }
          acc[oy][oz] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[kb], bv, acc[oy][oz], 0, 0, 0);
        }
      }
  } else {
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) {
        const int c = 2 * kc + half;  // the lane's 8 channels 8c..8c+7
        bf8 wa[3][3];                 // [dy + 1][dz + 1]
#pragma unroll
        for (int ty2 = 0; ty2 < 3; ++ty2)
#pragma unroll
          for (int tz2 = 0; tz2 < 3; ++tz2) {
            const int t = dx + 3 * ty2 + 9 * tz2;
            if (KIND == KIND_LAST) {
              if (col == 0) {
                wa[ty2][tz2] = *reinterpret_cast<const bf8*>(W + (size_t)t * kC + 8 * c);
              } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) wa[ty2][tz2][e] = (bf16)0.0f;
              }
            } else {
              wa[ty2][tz2] = *reinterpret_cast<const bf8*>(W + ((size_t)t * kC + col) * kC + 8 * c);
            }
          }
#pragma unroll
        for (int iy = 0; iy < 4; ++iy)
#pragma unroll
          for (int iz = 0; iz < 4; ++iz) {
            const int v = (col + dx) + kBX * ((wy + iy) + kBY * (wz + iz));
            const bf8 bv = as_bf8(brick[brick_chunk(v, c)]);
#pragma unroll
            for (int oy = 0; oy < 2; ++oy)
#pragma unroll
              for (int oz = 0; oz < 2; ++oz) {
                const int ty2 = iy - oy, tz2 = iz - oz;
                if (ty2 >= 0 && ty2 <= 2 && tz2 >= 0 && tz2 <= 2)
                  acc[oy][oz] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[ty2][tz2], bv, acc[oy][oz], 0, 0, 0);
              }
          }
      }
    }
  }

  // ---- epilogue
#pragma unroll
  for (int oy = 0; oy < 2; ++oy)
#pragma unroll
    for (int oz = 0; oz < 2; ++oz) {
      const int y = y0 + wy + oy, z = z0 + wz + oz;
      if (y >= d.ny || z >= d.nz) continue;
      const size_t p = (size_t)(x0 + col) + (size_t)d.nx * (y + (size_t)d.ny * z);
      if (KIND == KIND_LAST) {
        if (half == 0) zout[p] = acc[oy][oz][0] + bias[0];
        continue;
      }
#pragma unroll
      for (int gq = 0; gq < 4; ++gq) {
        const int co = 8 * gq + 4 * half;
        float v0 = acc[oy][oz][4 * gq], v1 = acc[oy][oz][4 * gq + 1], v2 = acc[oy][oz][4 * gq + 2],
              v3 = acc[oy][oz][4 * gq + 3];
        if (KIND == KIND_MID && MODE == MID_DATA) {
          *reinterpret_cast<float4*>(gout + p * kC + co) = float4{v0, v1, v2, v3};
          continue;
        }
        v0 += bias[co];
        v1 += bias[co + 1];
        v2 += bias[co + 2];
        v3 += bias[co + 3];
        if (zout) *reinterpret_cast<float4*>(zout + p * kC + co) = float4{v0, v1, v2, v3};
        *reinterpret_cast<uint2*>(aout + p * kC + co) = pack4(gelu(v0), gelu(v1), gelu(v2), gelu(v3));
      }
    }
}

// ---- loss: dout = 2 (out - target) / n; partial sums of (out - target)^2 per block
__global__ void k_loss(Dims d, const float* out, const float* target, float* dout, double* partials) {
  __shared__ double red[4];
  double s = 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < d.n; p += (size_t)gridDim.x * blockDim.x) {
    const float e = out[p] - target[p];
    dout[p] = 2.0f * e / (float)d.n;
    s += (double)e * e;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ---- layer 6 backward: input gradient g[p][c] = sum_t W[t][c] dout(p - tap_t)
__global__ void k_last_data(Dims d, const float* dout, const float* W, float* g) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= d.n) return;
  int i, j, k;
  coords(d, p, i, j, k);
  float acc[kC];
#pragma unroll
  for (int c = 0; c < kC; ++c) acc[c] = 0.0f;
  for (int t = 0; t < 27; ++t) {
    const long long q = neighbour(d, i, j, k, 26 - t);  // p - tap_t
    if (q < 0) continue;
    const float dv = dout[q];
    const float* w = W + t * kC;
#pragma unroll
    for (int c = 0; c < kC; ++c) acc[c] += w[c] * dv;
  }
#pragma unroll
  for (int c = 0; c < kC; ++c) g[p * kC + c] = acc[c];
}

// ---- layer 6 weight gradient: per chunk of voxels, W[t][c] += dout(p) in(p + tap_t)[c], b += dout(p).
// Thread (t, 4 channels) of 216, thread 216 the bias; the chunk's dout in LDS.
constexpr int kLastWThreads = 27 * kC / 4 + 1;
__global__ void __launch_bounds__(256) k_last_weights(Dims d, const float* dout, const bf16* in, float* partials) {
  __shared__ float ld[256];
  const int slot = threadIdx.x;
  const size_t p0 = (size_t)blockIdx.x * kChunk, p1 = p0 + kChunk < d.n ? p0 + kChunk : d.n;
  const int t = slot / (kC / 4), c0 = 4 * (slot % (kC / 4));
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (size_t pb = p0; pb < p1; pb += 256) {
    const size_t pl = pb + threadIdx.x;
    ld[threadIdx.x] = pl < p1 ? dout[pl] : 0.0f;
    __syncthreads();
    const int m = (int)(p1 - pb < 256 ? p1 - pb : 256);
    if (slot < kLastWThreads - 1) {
      for (int v = 0; v < m; ++v) {
        const size_t p = pb + v;
        int i, j, k;
        coords(d, p, i, j, k);
        const long long q = neighbour(d, i, j, k, t);
        if (q < 0) continue;
        const float dv = ld[v];
        typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
        const bf4 x = *(const bf4*)(in + (size_t)q * kC + c0);
#pragma unroll
        for (int e = 0; e < 4; ++e) s[e] += dv * (float)x[e];
      }
    } else if (slot == kLastWThreads - 1) {
      for (int v = 0; v < m; ++v) s[0] += ld[v];
    }
    __syncthreads();
  }
  float* out = partials + (size_t)blockIdx.x * (27 * kC + 1);
  if (slot < kLastWThreads - 1) {
#pragma unroll
    for (int e = 0; e < 4; ++e) out[t * kC + c0 + e] = s[e];
  } else if (slot == kLastWThreads - 1) {
    out[27 * kC] = s[0];
  }
}

// ---- GELU backward: dz = g * gelu'(z) (bf16 for the matrix cores), and per
// block partial sums of dz per channel (the bias gradient, FP32)
__global__ void k_gelu_back(Dims d, const float* g, const float* z, bf16* dz, float* bias_partials) {
  __shared__ float red[256];
  const size_t p0 = (size_t)blockIdx.x * kChunk, p1 = p0 + kChunk < d.n ? p0 + kChunk : d.n;
  const int c = threadIdx.x % kC, lane_group = threadIdx.x / kC;  // 8 groups of 32
  float s = 0.0f;
  for (size_t p = p0 + lane_group; p < p1; p += 8) {
    const float v = g[p * kC + c] * gelu_grad(z[p * kC + c]);
    dz[p * kC + c] = (bf16)v;
    s += v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < kC) {
    float t = 0.0f;
    for (int q = 0; q < 8; ++q) t += red[q * kC + threadIdx.x];
    bias_partials[(size_t)blockIdx.x * kC + threadIdx.x] = t;
  }
}

// ---- weight gradient of a 32 -> 32 layer on the matrix cores: for tap t and
// a chunk of voxels, dW[t] (32 co x 32 ci) += sum_p dz(p)[co] in(p + tap_t)[ci],
// with 16 voxels per MFMA step.  A workgroup of four wavefronts takes a chunk
// and four taps (blockIdx.y = tap group, wavefront w: tap 4 y + w, the last
// group has three): batches of 64 voxels are staged in LDS, dz once for the
// four taps and each tap's neighbours apart, and the next batch's global
// loads are issued before the current batch's MFMAs.  Each tap's MFMAs run
// over the chunk's 16-voxel steps in order.
constexpr int kWBatch = 64;  // voxels per staged batch

__global__ void __launch_bounds__(256) k_mid_weights(Dims d, const bf16* __restrict__ dz, const bf16* __restrict__ in,
                                                     float* partials) {
  __shared__ bf16 ld[kWBatch][kC];
  __shared__ bf16 la[4][kWBatch][kC];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = 4 * (int)blockIdx.y + w;  // 27: no tap (the wavefront only stages)
  const int col = lane & 31, half = lane >> 5;
  const size_t p0 = (size_t)blockIdx.x * kChunk, p1 = p0 + kChunk < d.n ? p0 + kChunk : d.n;
  // staging: thread T loads dz of voxel T / 4 (8 channels 8 (T % 4)); lane L
  // of wavefront w loads its tap's neighbour of voxels L / 4 + 16 u, u < 4
  const int vd = (int)threadIdx.x >> 2, cd = 8 * (threadIdx.x & 3);
  const int va = lane >> 2, ca = 8 * (lane & 3);
  auto load = [&](size_t pb, bf8& dv, bf8 (&av)[4]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) dv[e] = (bf16)0.0f;
    const size_t p = pb + vd;
    if (p < p1) dv = *(const bf8*)(dz + p * kC + cd);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int e = 0; e < 8; ++e) av[u][e] = (bf16)0.0f;
      const size_t q = pb + va + 16 * u;
      if (t < 27 && q < p1) {
        int i, j, k;
        coords(d, q, i, j, k);
        const long long nq = neighbour(d, i, j, k, t);
        if (nq >= 0) av[u] = *(const bf8*)(in + (size_t)nq * kC + ca);
      }
    }
  };
  f16v acc = {};
  bf8 dv, av[4];
  load(p0, dv, av);
  for (size_t pb = p0; pb < p1; pb += kWBatch) {
#pragma unroll
    for (int e = 0; e < 8; ++e) ld[vd][cd + e] = dv[e];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < 8; ++e) la[w][va + 16 * u][ca + e] = av[u][e];
    __syncthreads();
    if (pb + kWBatch < p1) load(pb + kWBatch, dv, av);  // in flight during the MFMAs
    if (t < 27) {
#pragma unroll
      for (int s = 0; s < kWBatch / 16; ++s) {
        if (pb + 16 * s >= p1) break;
        bf8 A, B;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          A[e] = ld[16 * s + 8 * half + e][col];     // A[co][k = voxel]
          B[e] = la[w][16 * s + 8 * half + e][col];  // B[k = voxel][ci]
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc, 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if (t >= 27) return;
  float* out = partials + ((size_t)blockIdx.x * 27 + t) * kC * kC;
#pragma unroll
  for (int qi = 0; qi < 16; ++qi) {
    const int co = 8 * (qi / 4) + 4 * half + (qi % 4);
    out[co * kC + col] = acc[qi];
  }
}

// ---- layer 1 weight gradient: W[t][co][ci] += dz(p)[co] x(p + tap_t)[ci].
// A batch of 64 voxels of the chunk is staged in LDS (dz and the 27 x 3
// neighbour inputs); thread (t, 4 output channels) of 216 accumulates 12
// weights over the batch, voxel by voxel in order.
constexpr int kFirstWBatch = 64;
__global__ void __launch_bounds__(256) k_first_weights(Dims d, const bf16* dz, const float* x, float* partials) {
  __shared__ float ldz[kFirstWBatch][kC];
  __shared__ float lx[kFirstWBatch][27 * 3];
  const size_t p0 = (size_t)blockIdx.x * kChunk, p1 = p0 + kChunk < d.n ? p0 + kChunk : d.n;
  const int slot = threadIdx.x;
  const int t = slot / (kC / 4), co0 = 4 * (slot % (kC / 4));
  float s[4][3] = {};
  for (size_t pb = p0; pb < p1; pb += kFirstWBatch) {
    const int m = (int)(p1 - pb < kFirstWBatch ? p1 - pb : kFirstWBatch);
    for (int e = threadIdx.x; e < kFirstWBatch * kC; e += blockDim.x) {
      const int v = e / kC, c = e % kC;
      ldz[v][c] = v < m ? (float)dz[(pb + v) * kC + c] : 0.0f;
    }
    for (int e = threadIdx.x; e < kFirstWBatch * 27; e += blockDim.x) {
      const int v = e / 27, tt = e % 27;
      float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
      if (v < m) {
        int i, j, k;
        coords(d, pb + v, i, j, k);
        const long long q = neighbour(d, i, j, k, tt);
        if (q >= 0) {
          x0 = x[3 * q];
          x1 = x[3 * q + 1];
          x2 = x[3 * q + 2];
        }
      }
      lx[v][3 * tt] = x0;
      lx[v][3 * tt + 1] = x1;
      lx[v][3 * tt + 2] = x2;
    }
    __syncthreads();
    if (slot < 27 * kC / 4) {
      for (int v = 0; v < m; ++v) {
        const float x0 = lx[v][3 * t], x1 = lx[v][3 * t + 1], x2 = lx[v][3 * t + 2];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float dv = ldz[v][co0 + c];
          s[c][0] += dv * x0;
          s[c][1] += dv * x1;
          s[c][2] += dv * x2;
        }
      }
    }
    __syncthreads();
  }
  if (slot < 27 * kC / 4) {
    float* out = partials + (size_t)blockIdx.x * (27 * kC * 3) + (size_t)t * kC * 3;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) out[(co0 + c) * 3 + ci] = s[c][ci];
  }
}

// sum over chunks in a fixed order, in two levels: groups of kSumGroup
// consecutive chunks (blockIdx.y) into group sums, then the group sums in
// group order (k_sum_groups): out[s] = sum_g sum_{c in g} partials[c][s]
constexpr int kSumGroup = 64;
__global__ void k_sum_chunks(const float* partials, int nchunks, int size, float* groups) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= size) return;
  const int c0 = blockIdx.y * kSumGroup, c1 = c0 + kSumGroup < nchunks ? c0 + kSumGroup : nchunks;
  float acc = 0.0f;
  for (int c = c0; c < c1; ++c) acc += partials[(size_t)c * size + s];
  groups[(size_t)blockIdx.y * size + s] = acc;
}
__global__ void k_sum_groups(const float* groups, int ngroups, int size, float* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= size) return;
  float acc = 0.0f;
  for (int c = 0; c < ngroups; ++c) acc += groups[(size_t)c * size + s];
  out[s] = acc;
}

// bf16 copies of the 32 -> 32 weights: W16[t][co][ci] and the transposed,
// tap-flipped WT16[t][ci][co] = W[26 - t][co][ci] of the input gradient
__global__ void k_pack(const float* params, bf16* w16, bf16* wt16) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;  // within one layer
  const int l = 1 + blockIdx.y;
  if (s >= 27 * kC * kC) return;
  const float* W = params + layer_offset(l);
  const int t = s / (kC * kC), co = (s / kC) % kC, ci = s % kC;
  const size_t base = (size_t)(l - 1) * 27 * kC * kC;
  w16[base + s] = (bf16)W[s];
  wt16[base + ((size_t)(26 - t) * kC + ci) * kC + co] = (bf16)W[s];
}

// bf16 copies of the first layer's weights as the A operand wants them,
// wf16[co][3 t + ci] padded with zeros to kK1 columns, and of the last
// layer's, wl16[t][ci]
__global__ void k_pack_ends(const float* params, bf16* wf16, bf16* wl16) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < kC * kK1) {
    const int co = s / kK1, k = s % kK1;
    const float* W = params + layer_offset(0);
    wf16[s] = k < 81 ? (bf16)W[((k / 3) * kC + co) * 3 + k % 3] : (bf16)0.0f;
  }
  if (s < 27 * kC) wl16[s] = (bf16)params[layer_offset(5) + s];
}

// input channels of the learned guess: b / |b|, c_prev prev / |b| (prev may
// be null: 0) and jac - 1, channels-last FP32
__global__ void k_guess_input(pcf::Geo g, const double* b, const double* prev, const double* jac, float* in,
                              double inv_bnorm, double prev_scale) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= g.nv()) return;
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((size_t)g.nx * g.ny));
  in[3 * p] = (float)(b[p] * inv_bnorm);
  in[3 * p + 1] = prev ? (float)(prev[p] * prev_scale * inv_bnorm) : 0.0f;
  in[3 * p + 2] = (float)(jac[g.h(g.i0 + i, g.j0 + j, g.k0 + k)] - 1.0);
}

// x = c prev + |b| out: the network corrects the rescaled previous solution
// (prev may be null: the first step of a run)
__global__ void k_guess_output(const float* out, double* x, size_t n, double bnorm, const double* prev, double c) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    x[q] = (prev ? c * prev[q] : 0.0) + bnorm * (double)out[q];
}

// ---- the learned preconditioning stage and the residual training
constexpr int kStageGrid = 2048;  // blocks of the sum-of-squares partials

__global__ void k_sumsq(const double* v, size_t n, double* partials) {
  double val[1] = {0.0};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    val[0] += v[q] * v[q];
  block_partials<1>(val, partials, blockIdx.x);
}

// *out = sqrt(sum of the partials / n), the RMS norm
__global__ void k_rms_final(const double* partials, int nparts, double* out, double n) {
  const double s = finalize_sum(partials, 0, nparts);
  if (threadIdx.x == 0) *out = sqrt(s / n);
}

// the stage's input channels v / |v|, y / |v| and J - 1 with |v| on the device
__global__ void k_stage_input(pcf::Geo g, const double* v, const double* y, const double* jac, float* in,
                              const double* nrm, const SolverScalars* sc) {
  if (sc && sc->done) return;
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= g.nv()) return;
  const int i = (int)(p % g.nx), j = (int)((p / g.nx) % g.ny), k = (int)(p / ((size_t)g.nx * g.ny));
  const double inv = *nrm > 0.0 ? 1.0 / *nrm : 0.0;
  in[3 * p] = (float)(v[p] * inv);
  in[3 * p + 1] = (float)(y[p] * inv);
  in[3 * p + 2] = (float)(jac[g.h(g.i0 + i, g.j0 + j, g.k0 + k)] - 1.0);
}

// y += |v| out
__global__ void k_stage_output(const float* out, double* y, size_t n, const double* nrm, const SolverScalars* sc) {
  if (sc && sc->done) return;
  const double s = *nrm;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    y[q] += s * (double)out[q];
}

// x = base + c out (base may be null: 0)
__global__ void k_compose(const float* out, const double* base, double* x, size_t n, double c) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    x[q] = (base ? base[q] : 0.0) + c * (double)out[q];
}

// r = a - b
__global__ void k_sub(const double* a, const double* b, double* r, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    r[q] = a[q] - b[q];
}

// the output gradient of the residual loss, c A^T r, in FP32
__global__ void k_dout(const double* atr, float* dout, size_t n, double c) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    dout[q] = (float)(c * atr[q]);
}

// Adam (Kingma & Ba 2015) on the FP32 parameters
__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace
}  // namespace pcf

using pcf::bf16;
struct pcf_net;

struct pcf_net {
  pcf::Dims d;
  bool training = false;
  hipStream_t stream = nullptr;
  int device = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  bf16 *w16 = nullptr, *wt16 = nullptr;
  bf16 *w16_first = nullptr, *w16_last = nullptr;  // the first and last layers' weights in bf16
  float* z[5] = {};        // pre-activations of layers 1..5 (training)
  bf16* a[5] = {};         // activations of layers 1..5 (inference: a[0], a[1] alternate)
  float* out = nullptr;    // network output
  float *dout = nullptr, *g = nullptr;
  bf16* dzb = nullptr;
  float* partials = nullptr;
  float* groups = nullptr;  // group sums of the chunk partials
  double* loss_partials = nullptr;
  double* loss_host = nullptr;
  int step = 0;
  std::vector<void*> allocations;
};

namespace pcf {
namespace {

int net_alloc(pcf_net* net, void** p, size_t bytes) {
  PCF_HIP(hipMalloc(p, bytes));
  net->allocations.push_back(*p);
  return PCF_OK;
}

int nchunks(const pcf_net* net) { return (int)((net->d.n + kChunk - 1) / kChunk); }

// blocks of the k_brick launches
BrickGrid brick_grid(const Dims& d) {
  return BrickGrid{d.nx / 32, (d.ny + 3) / 4, (d.nz + 3) / 4};
}
unsigned brick_blocks(const BrickGrid& bg) { return (unsigned)(bg.tiles_x * bg.tiles_y * bg.tiles_z); }

// out[0..size) = sum over the nc chunk partials (fixed order, two levels)
void sum_chunks(pcf_net* net, int nc, int size, float* out) {
  const unsigned ng = (unsigned)((nc + kSumGroup - 1) / kSumGroup);
  const unsigned blocks = (unsigned)((size + 255) / 256);
  hipLaunchKernelGGL(k_sum_chunks, dim3(blocks, ng), dim3(256), 0, net->stream, net->partials, nc, size, net->groups);
  hipLaunchKernelGGL(k_sum_groups, dim3(blocks), dim3(256), 0, net->stream, net->groups, (int)ng, size, out);
}

// forward pass over the FP32 input (n x 3); output in net->out
int net_forward(pcf_net* net, const float* input, hipStream_t s) {
  const Dims& d = net->d;
  const BrickGrid bg = brick_grid(d);
  const unsigned nb = brick_blocks(bg);
  const float* P = net->params;
  bf16* cur = net->a[0];
  hipLaunchKernelGGL((k_brick<KIND_FIRST, MID_FWD>), dim3(nb), dim3(256), 0, s, d, bg, (const void*)input,
                     net->w16_first, P + layer_offset(0) + 27 * kC * 3, net->training ? net->z[0] : nullptr, cur,
                     nullptr);
  for (int l = 1; l <= 4; ++l) {
    bf16* nxt = net->training ? net->a[l] : net->a[l & 1];
    hipLaunchKernelGGL((k_brick<KIND_MID, MID_FWD>), dim3(nb), dim3(256), 0, s, d, bg, (const void*)cur,
                       net->w16 + (size_t)(l - 1) * 27 * kC * kC, P + layer_offset(l) + 27 * kC * kC,
                       net->training ? net->z[l] : nullptr, nxt, nullptr);
    cur = nxt;
  }
  hipLaunchKernelGGL((k_brick<KIND_LAST, MID_FWD>), dim3(nb), dim3(256), 0, s, d, bg, (const void*)cur,
                     net->w16_last, P + layer_offset(5) + 27 * kC, net->out, nullptr, nullptr);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// backward pass from the output gradient in net->dout; gradients in net->grads
int net_backward_dout(pcf_net* net, const float* input) {
  const Dims& d = net->d;
  hipStream_t s = net->stream;
  const unsigned vb = 256, vg = (unsigned)((d.n + vb - 1) / vb);
  const BrickGrid bg = brick_grid(d);
  const int nc = nchunks(net);
  const float* P = net->params;
  float* G = net->grads;
  // layer 6
  hipLaunchKernelGGL(k_last_weights, dim3(nc), dim3(256), 0, s, d, net->dout, net->a[4], net->partials);
  sum_chunks(net, nc, 27 * kC + 1, G + layer_offset(5));
  hipLaunchKernelGGL(k_last_data, dim3(vg), dim3(vb), 0, s, d, net->dout, P + layer_offset(5), net->g);
  // layers 5..2
  for (int l = 4; l >= 1; --l) {
    hipLaunchKernelGGL(k_gelu_back, dim3(nc), dim3(256), 0, s, d, net->g, net->z[l], net->dzb, net->partials);
    sum_chunks(net, nc, kC, G + layer_offset(l) + 27 * kC * kC);
    hipLaunchKernelGGL(k_mid_weights, dim3(nc, 7), dim3(256), 0, s, d, net->dzb, net->a[l - 1], net->partials);
    sum_chunks(net, nc, 27 * kC * kC, G + layer_offset(l));
    hipLaunchKernelGGL((k_brick<KIND_MID, MID_DATA>), dim3(brick_blocks(bg)), dim3(256), 0, s, d, bg,
                       (const void*)net->dzb, net->wt16 + (size_t)(l - 1) * 27 * kC * kC, nullptr, nullptr, nullptr,
                       net->g);
  }
  // layer 1
  hipLaunchKernelGGL(k_gelu_back, dim3(nc), dim3(256), 0, s, d, net->g, net->z[0], net->dzb, net->partials);
  sum_chunks(net, nc, kC, G + layer_offset(0) + 27 * kC * 3);
  hipLaunchKernelGGL(k_first_weights, dim3(nc), dim3(256), 0, s, d, net->dzb, input, net->partials);
  sum_chunks(net, nc, 27 * kC * 3, G + layer_offset(0));
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

// backward pass for the mean squared error against target; loss on the host
int net_backward(pcf_net* net, const float* input, const float* target, double* loss) {
  const Dims& d = net->d;
  hipStream_t s = net->stream;
  const unsigned lg = 1024;
  hipLaunchKernelGGL(k_loss, dim3(lg), dim3(256), 0, s, d, net->out, target, net->dout, net->loss_partials);
  int rc = net_backward_dout(net, input);
  if (rc) return rc;
  std::vector<double> lp(lg);
  PCF_HIP(hipMemcpyAsync(lp.data(), net->loss_partials, lg * sizeof(double), hipMemcpyDeviceToHost, s));
  PCF_HIP(hipStreamSynchronize(s));
  double sum = 0.0;
  for (double v : lp) sum += v;
  if (loss) *loss = sum / (double)d.n;
  return PCF_OK;
}

// input channels for b (and prev scaled by prev_scale) and the RMS norm of b
int guess_input(pcf_state* st, const double* b, const double* prev, double prev_scale, float* input,
                double* bnorm) {
  double bb = 0.0;
  int rc = dot(st, b, b, &bb);
  if (rc) return rc;
  const double nrm = std::sqrt(bb / (double)st->g.nv());
  *bnorm = nrm;
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  hipLaunchKernelGGL(k_guess_input, dim3((unsigned)((st->g.nv() + 255) / 256)), dim3(256), 0, st->stream, st->g, b,
                     prev, st->jac, input, inv, prev_scale);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

int net_pack(pcf_net* net) {
  hipLaunchKernelGGL(k_pack, dim3((27 * kC * kC + 255) / 256, 4), dim3(256), 0, net->stream, net->params, net->w16,
                     net->wt16);
  hipLaunchKernelGGL(k_pack_ends, dim3((kC * kK1 + 255) / 256), dim3(256), 0, net->stream, net->params,
                     net->w16_first, net->w16_last);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

}  // namespace

// The learned initial guess of the mixed solve (apply_bicgstab_mixed): x0 in
// st->guess_x, the stage's previous solution rescaled to this dt plus the
// network's correction from the right-hand side, that rescaled solution and
// the Jacobian.
int learned_guess(pcf_state* st, double dt, const double* prev, const double** x0) {
  pcf_net* net = st->guess_net;
  if (!net) {
    set_error("initial_guess = learned needs a network (pcf_set_guess_network)");
    return PCF_E_CONFIG;
  }
  double bnorm = 0.0;
  int rc = guess_input(st, st->rhs, prev, dt, st->guess_in, &bnorm);
  if (rc) return rc;
  ProfScope prof(st, PCF_K_NET);
  if ((rc = net_forward(net, st->guess_in, st->stream))) return rc;
  hipLaunchKernelGGL(k_guess_output, dim3(2048), dim3(256), 0, st->stream, net->out, st->guess_x, st->g.nv(), bnorm,
                     prev, dt);
  PCF_HIP(hipGetLastError());
  *x0 = st->guess_x;
  return PCF_OK;
}

int learned_stage(pcf_state* st, const double* v, double* y, const SolverScalars* sc) {
  pcf_net* net = st->stage_net;
  const Geo& g = st->g;
  const size_t N = g.nv();
  ProfScope prof(st, PCF_K_NET);
  // the fourth dot slot of the partials: the solver's own dot products use
  // slots 0-2 between a kernel that writes them and the one that reads them
  double* part = st->partials + (size_t)3 * kPartialStride;
  hipLaunchKernelGGL(k_sumsq, dim3(kStageGrid), dim3(kReduceBlock), 0, st->stream, v, N, part);
  hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(1024), 0, st->stream, part, kStageGrid, st->stage_norm, (double)N);
  hipLaunchKernelGGL(k_stage_input, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st->stream, g, v, y, st->jac,
                     st->stage_in, st->stage_norm, sc);
  int rc = net_forward(net, st->stage_in, st->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_stage_output, dim3(2048), dim3(256), 0, st->stream, net->out, y, N, st->stage_norm, sc);
  PCF_HIP(hipGetLastError());
  return PCF_OK;
}

}  // namespace pcf

extern "C" {

int pcf_guess_input(pcf_state* st, const double* d_b, const double* d_prev, float* d_input, double* b_norm) {
  if (!st || !d_b || !d_input) return PCF_E_ARG;
  double nrm = 0.0;
  int rc = pcf::guess_input(st, d_b, d_prev, 1.0, d_input, &nrm);
  if (rc) return rc;
  PCF_HIP(hipStreamSynchronize(st->stream));
  if (b_norm) *b_norm = nrm;
  return PCF_OK;
}

int pcf_guess_fallbacks(const pcf_state* st, int32_t* count) {
  if (!st || !count) return PCF_E_ARG;
  *count = st->guess_fallbacks;
  return PCF_OK;
}

int pcf_set_guess_network(pcf_state* st, pcf_net* net) {
  using namespace pcf;
  if (!st) return PCF_E_ARG;
  if (net && ((size_t)net->d.nx != (size_t)st->g.nx || net->d.ny != st->g.ny || net->d.nz != st->g.nz)) {
    set_error("pcf_set_guess_network: the network's grid differs from the state's");
    return PCF_E_ARG;
  }
  if (net && !st->guess_in) {
    const size_t N = st->g.nv();
    int rc = device_alloc(st, &st->guess_in, 3 * N);
    if (rc == PCF_OK) rc = device_alloc(st, &st->guess_x, N);
    if (rc) return rc;
  }
  st->guess_net = net;
  return PCF_OK;
}


int pcf_set_stage_network(pcf_state* st, pcf_net* net, int32_t iterations) {
  using namespace pcf;
  if (!st || iterations < 0) return PCF_E_ARG;
  if (net && ((size_t)net->d.nx != (size_t)st->g.nx || net->d.ny != st->g.ny || net->d.nz != st->g.nz)) {
    set_error("pcf_set_stage_network: the network's grid differs from the state's");
    return PCF_E_ARG;
  }
  if (net && (st->cfg.precision != PCF_PRECISION_MIXED || !st->cfg.preconditioner)) {
    set_error("pcf_set_stage_network: the learned stage needs precision = mixed and a preconditioner "
              "(a nonlinear stage needs the flexible form of the solution update)");
    return PCF_E_CONFIG;
  }
  if (net && !st->stage_in) {
    int rc = device_alloc(st, &st->stage_in, 3 * st->g.nv());
    if (rc == PCF_OK) rc = device_alloc(st, &st->stage_norm, 1);
    if (rc) return rc;
  }
  st->stage_net = net;
  st->stage_iterations = net ? iterations : 0;
  return PCF_OK;
}

int pcf_stage_fallbacks(const pcf_state* st, int32_t* count) {
  if (!st || !count) return PCF_E_ARG;
  *count = st->stage_fallbacks;
  return PCF_OK;
}

int pcf_net_param_count(void) { return pcf::kParams; }

int pcf_net_create(int32_t nx, int32_t ny, int32_t nz, int32_t training, pcf_net** out) {
  using namespace pcf;
  if (!out || nx < 32 || nx % 32 != 0 || ny < 1 || nz < 1) {
    set_error("pcf_net_create: nx must be a positive multiple of 32");
    return PCF_E_ARG;
  }
  if ((long long)nx * ny * nz > 0x7fffffffLL) {
    set_error("pcf_net_create: at most 2^31 - 1 voxels");
    return PCF_E_ARG;
  }
  pcf_net* net = new pcf_net();
  net->d = Dims{nx, ny, nz, (size_t)nx * ny * nz};
  net->training = training != 0;
  auto fail = [&](int rc) {
    for (void* p : net->allocations) (void)hipFree(p);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return rc;
  };
  if (hipGetDevice(&net->device) != hipSuccess) return fail(PCF_E_HIP);
  if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess) return fail(PCF_E_HIP);
  const size_t N = net->d.n;
  int rc;
  if ((rc = net_alloc(net, (void**)&net->params, kParams * sizeof(float)))) return fail(rc);
  if ((rc = net_alloc(net, (void**)&net->w16, 4 * 27 * kC * kC * sizeof(bf16)))) return fail(rc);
  if ((rc = net_alloc(net, (void**)&net->wt16, 4 * 27 * kC * kC * sizeof(bf16)))) return fail(rc);
  if ((rc = net_alloc(net, (void**)&net->w16_first, kC * kK1 * sizeof(bf16)))) return fail(rc);
  if ((rc = net_alloc(net, (void**)&net->w16_last, 27 * kC * sizeof(bf16)))) return fail(rc);
  if ((rc = net_alloc(net, (void**)&net->out, N * sizeof(float)))) return fail(rc);
  const int na = net->training ? 5 : 2;
  for (int l = 0; l < na; ++l)
    if ((rc = net_alloc(net, (void**)&net->a[l], N * kC * sizeof(bf16)))) return fail(rc);
  if (net->training) {
    for (int l = 0; l < 5; ++l)
      if ((rc = net_alloc(net, (void**)&net->z[l], N * kC * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->grads, kParams * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->adam_m, kParams * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->adam_v, kParams * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->dout, N * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->g, N * kC * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->dzb, N * kC * sizeof(bf16)))) return fail(rc);
    const size_t part = (size_t)nchunks(net) * 27 * kC * kC;
    if ((rc = net_alloc(net, (void**)&net->partials, part * sizeof(float)))) return fail(rc);
    const size_t ngroups = (size_t)(nchunks(net) + kSumGroup - 1) / kSumGroup;
    if ((rc = net_alloc(net, (void**)&net->groups, ngroups * 27 * kC * kC * sizeof(float)))) return fail(rc);
    if ((rc = net_alloc(net, (void**)&net->loss_partials, 1024 * sizeof(double)))) return fail(rc);
    if (hipMemsetAsync(net->adam_m, 0, kParams * sizeof(float), net->stream) != hipSuccess) return fail(PCF_E_HIP);
    if (hipMemsetAsync(net->adam_v, 0, kParams * sizeof(float), net->stream) != hipSuccess) return fail(PCF_E_HIP);
  }
  if (hipMemsetAsync(net->params, 0, kParams * sizeof(float), net->stream) != hipSuccess) return fail(PCF_E_HIP);
  if (hipStreamSynchronize(net->stream) != hipSuccess) return fail(PCF_E_HIP);
  *out = net;
  return PCF_OK;
}

int pcf_net_destroy(pcf_net* net) {
  if (!net) return PCF_E_ARG;
  (void)hipStreamSynchronize(net->stream);
  for (void* p : net->allocations) (void)hipFree(p);
  (void)hipStreamDestroy(net->stream);
  delete net;
  return PCF_OK;
}

int pcf_net_set_params(pcf_net* net, const float* host) {
  using namespace pcf;
  if (!net || !host) return PCF_E_ARG;
  PCF_HIP(hipMemcpyAsync(net->params, host, kParams * sizeof(float), hipMemcpyHostToDevice, net->stream));
  int rc = net_pack(net);
  if (rc) return rc;
  PCF_HIP(hipStreamSynchronize(net->stream));
  return PCF_OK;
}

int pcf_net_get_params(pcf_net* net, float* host) {
  using namespace pcf;
  if (!net || !host) return PCF_E_ARG;
  PCF_HIP(hipMemcpyAsync(host, net->params, kParams * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  PCF_HIP(hipStreamSynchronize(net->stream));
  return PCF_OK;
}

int pcf_net_forward(pcf_net* net, const float* d_input, float* d_output) {
  using namespace pcf;
  if (!net || !d_input || !d_output) return PCF_E_ARG;
  int rc = net_forward(net, d_input, net->stream);
  if (rc) return rc;
  PCF_HIP(hipMemcpyAsync(d_output, net->out, net->d.n * sizeof(float), hipMemcpyDeviceToDevice, net->stream));
  PCF_HIP(hipStreamSynchronize(net->stream));
  return PCF_OK;
}

int pcf_net_gradients(pcf_net* net, const float* d_input, const float* d_target, float* host_grads, double* loss) {
  using namespace pcf;
  if (!net || !d_input || !d_target || !net->training) return PCF_E_ARG;
  int rc = net_forward(net, d_input, net->stream);
  if (rc) return rc;
  if ((rc = net_backward(net, d_input, d_target, loss))) return rc;
  if (host_grads) {
    PCF_HIP(hipMemcpyAsync(host_grads, net->grads, kParams * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    PCF_HIP(hipStreamSynchronize(net->stream));
  }
  return PCF_OK;
}

// one Adam step on the gradients in net->grads
static int adam_step(pcf_net* net, float lr) {
  using namespace pcf;
  net->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)net->step), c2 = 1.0f - std::pow(b2, (float)net->step);
  hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, net->stream, net->params, net->grads,
                     net->adam_m, net->adam_v, kParams, lr, b1, b2, 1e-8f, c1, c2);
  int rc = net_pack(net);
  if (rc) return rc;
  PCF_HIP(hipStreamSynchronize(net->stream));
  return PCF_OK;
}

int pcf_net_train_step(pcf_net* net, const float* d_input, const float* d_target, float lr, double* loss) {
  using namespace pcf;
  int rc = pcf_net_gradients(net, d_input, d_target, nullptr, loss);
  if (rc) return rc;
  return adam_step(net, lr);
}

// L = |A x - v|^2 / |v|^2 for x = base + |v| out, out the network's output
// (x in st->train_buf[0]); the residual stays in train_buf[1]
static int residual_loss(pcf_state* st, pcf_net* net, const double* v, const double* base, double vnorm, double vv,
                         double* loss) {
  using namespace pcf;
  const size_t N = st->g.nv();
  double *x = st->train_buf[0], *r = st->train_buf[1];
  int rc = net_forward(net, st->guess_in, st->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compose, dim3(2048), dim3(256), 0, st->stream, net->out, base, x, N, vnorm);
  PCF_HIP(hipGetLastError());
  if ((rc = apply_operator(st, x, r, true))) return rc;
  hipLaunchKernelGGL(k_sub, dim3(2048), dim3(256), 0, st->stream, r, v, r, N);
  PCF_HIP(hipGetLastError());
  double rr = 0.0;
  if ((rc = dot(st, r, r, &rr))) return rc;
  *loss = rr / vv;
  return PCF_OK;
}

int pcf_net_train_residual(pcf_state* st, pcf_net* net, const double* d_v, const double* d_base, int32_t objective,
                           int32_t inner, float lr, double* loss_base, double* loss_net) {
  using namespace pcf;
  if (!st || !net || !d_v || inner < 0 || (inner > 0 && !net->training)) return PCF_E_ARG;
  if (objective != PCF_OBJECTIVE_RESIDUAL && objective != PCF_OBJECTIVE_ENERGY) {
    set_error("pcf_net_train_residual: unknown objective");
    return PCF_E_OPTION;
  }
  if ((size_t)net->d.nx != (size_t)st->g.nx || net->d.ny != st->g.ny || net->d.nz != st->g.nz) {
    set_error("pcf_net_train_residual: the network's grid differs from the state's");
    return PCF_E_ARG;
  }
  const size_t N = st->g.nv();
  int rc;
  if (!st->guess_in && (rc = device_alloc(st, &st->guess_in, 3 * N))) return rc;
  for (double*& b : st->train_buf)
    if (!b && (rc = device_alloc(st, &b, N))) return rc;
  double vnorm = 0.0;
  if ((rc = guess_input(st, d_v, d_base, 1.0, st->guess_in, &vnorm))) return rc;
  const double vv = vnorm * vnorm * (double)N;
  if (!(vv > 0.0)) {
    set_error("pcf_net_train_residual: v is zero");
    return PCF_E_ARG;
  }
  double* r = st->train_buf[1];
  double* atr = st->train_buf[2];
  if (loss_base) {  // the residual of base alone
    double rr = vv;
    if (d_base) {
      if ((rc = apply_operator(st, d_base, r, true))) return rc;
      hipLaunchKernelGGL(k_sub, dim3(2048), dim3(256), 0, st->stream, r, d_v, r, N);
      PCF_HIP(hipGetLastError());
      if ((rc = dot(st, r, r, &rr))) return rc;
    }
    *loss_base = rr / vv;
  }
  double loss = 0.0;
  for (int it = 0; it < inner; ++it) {
    if ((rc = residual_loss(st, net, d_v, d_base, vnorm, vv, &loss))) return rc;
    if (objective == PCF_OBJECTIVE_RESIDUAL) {
      if ((rc = apply_operator_transpose(st, r, atr))) return rc;
      hipLaunchKernelGGL(k_dout, dim3(2048), dim3(256), 0, st->stream, atr, net->dout, N, 2.0 * vnorm / vv);
    } else {  // the energy norm of the error under -A: its gradient in x is -(A x - v)
      hipLaunchKernelGGL(k_dout, dim3(2048), dim3(256), 0, st->stream, r, net->dout, N, -2.0 * vnorm / vv);
    }
    PCF_HIP(hipGetLastError());
    PCF_HIP(hipStreamSynchronize(st->stream));  // the backward pass runs on the network's stream
    if ((rc = net_backward_dout(net, st->guess_in))) return rc;
    if ((rc = adam_step(net, lr))) return rc;
  }
  if (loss_net) {
    if ((rc = residual_loss(st, net, d_v, d_base, vnorm, vv, &loss))) return rc;
    *loss_net = loss;
  }
  PCF_HIP(hipStreamSynchronize(st->stream));
  return PCF_OK;
}

int pcf_net_train_step_dout(pcf_net* net, const float* d_input, const float* d_dout, float lr) {
  using namespace pcf;
  if (!net || !d_input || !d_dout || !net->training) return PCF_E_ARG;
  int rc = net_forward(net, d_input, net->stream);
  if (rc) return rc;
  PCF_HIP(hipMemcpyAsync(net->dout, d_dout, net->d.n * sizeof(float), hipMemcpyDeviceToDevice, net->stream));
  if ((rc = net_backward_dout(net, d_input))) return rc;
  return adam_step(net, lr);
}

}  // extern "C"
