// cmi_net.hip -- the learned warm start: a 4-layer 3-D convolutional network
// (3x3x3 kernels, 32 channels, ReLU) that maps (log n_H, log of the
// first-pass photoionization rate) to the logit of the converged neutral
// fraction.  Trained and run on the GPU by the kernels below:
//
//   conv3_kernel   implicit-GEMM 3-D convolution on BF16 MFMA
//                  (v_mfma_f32_16x16x32_bf16, FP32 accumulation): one wave
//                  computes 16 voxels (a run along the last axis) x 16 or 32
//                  output channels; K = 27 taps x 32 input channels, one MFMA
//                  per tap and output block.  Serves the forward pass
//                  (weights [tap][cout][cin], bias, ReLU) and the input
//                  gradient (the same convolution over the output gradient
//                  with the flipped, transposed weights, masked by the ReLU of
//                  the layer below).
//   wgrad_kernel   weight gradient dW[tap][cout][cin] = sum_v dy[v][cout] *
//                  x[v + tap][cin] on the same MFMA, K = 32 voxels per
//                  instruction; FP32 accumulation, one atomic add per weight
//                  per wave.
//   plus the loss (mean squared error on logits), bias gradients, Adam on
//   FP32 master weights, weight packing to BF16, crop sampling and input
//   normalisation.
//
// Activations are BF16, channels-last with 32 channels ([B][D][H][W][32]);
// the first layer's 2 input channels and the last layer's single output are
// zero-padded to the MFMA shapes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/cmi.h"
#include "cmi_ctx.hpp"
#include "cmi_device.hpp"

namespace cmi {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kC = 32;        // channels of every activation tensor
constexpr int kTaps = 27;
constexpr int kLayers = 4;
// real input / output channels of each layer
__host__ __device__ constexpr int layer_cin(int l) { return l == 0 ? 2 : 32; }
__host__ __device__ constexpr int layer_cout(int l) { return l == kLayers - 1 ? 1 : 32; }

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

// ---------------------------------------------------------------------------
// 3-D convolution (padding 1) as implicit GEMM on MFMA 16x16x32 BF16.
// Lane l of a wave: A rows (voxels) l%16, A/B k-group (input channels)
// 8*(l/16)..+7, B column (output channel) l%16; accumulator rows (voxels)
// 4*(l/16)..+3, column l%16.
// ---------------------------------------------------------------------------
template <int NB>
__global__ void __launch_bounds__(256) conv3_kernel(
    const __bf16 *__restrict__ x, const __bf16 *__restrict__ w,
    const float *__restrict__ bias, int nrun, int D, int H, int W, int dil,
    __bf16 *__restrict__ y, float *__restrict__ yf,
    const __bf16 *__restrict__ mask, int relu) {
  const int lane = threadIdx.x & 63;
  const int run = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (run >= nrun) return;  // whole wave leaves together
  // run -> (b, d, h, w0): 16 voxels along the last axis
  const int wr = W >> 4;
  int t = run;
  const int w0 = (t % wr) << 4;
  t /= wr;
  const int h = t % H;
  t /= H;
  const int d = t % D;
  const int b = t / D;
  const int vl = lane & 15;   // voxel (A row) / output channel (B column)
  const int kg = lane >> 4;   // k group: channels 8*kg .. 8*kg+7
  f32x4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int NOUT = 16 * NB;
#pragma unroll 1
  for (int tap = 0; tap < kTaps; ++tap) {
    const int kd = tap / 9, kh = (tap / 3) % 3, kw = tap % 3;
    const int zd = d + (kd - 1) * dil, zh = h + (kh - 1) * dil,
              zw = w0 + vl + (kw - 1) * dil;
    bf16x8 a;
    if ((unsigned)zd < (unsigned)D && (unsigned)zh < (unsigned)H &&
        (unsigned)zw < (unsigned)W) {
      const size_t vox = (((size_t)b * D + zd) * H + zh) * W + zw;
      a = *reinterpret_cast<const bf16x8 *>(x + vox * kC + 8 * kg);
    } else {
      a = bf16x8{};
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const bf16x8 bw = *reinterpret_cast<const bf16x8 *>(
          w + ((size_t)tap * NOUT + nb * 16 + vl) * kC + 8 * kg);
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[nb], 0, 0, 0);
    }
  }
  const size_t vbase = (((size_t)b * D + d) * H + h) * W + w0;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = nb * 16 + vl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t v = vbase + 4 * kg + i;
      float r = acc[nb][i] + (bias ? bias[co] : 0.f);
      if (mask) r = bf2f(mask[v * kC + co]) > 0.f ? r : 0.f;
      if (relu) r = fmaxf(r, 0.f);
      if (y) y[v * kC + co] = f2bf(r);
      if (yf && co == 0) yf[v] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled forward convolution, for the shapes it divides: the same outputs as
// conv3_kernel bit for bit (every output voxel's 27 taps go through the same
// MFMAs with the same operands in the same order).  A convolution with
// dilation d is d^3 ordinary 3x3x3 convolutions on the sub-lattices of voxels
// with equal coordinates mod d.  A workgroup takes a block of one sub-lattice,
// kTZ x kTY runs of 16 voxels along the last axis, and stages the block with
// its one-voxel halo in LDS, so that each input voxel is read from memory
// once per block instead of once per tap and output run; each wave computes
// kRunsPerWave runs and loads each tap's weights once for all of them.
// ---------------------------------------------------------------------------
constexpr int kTZ = 4, kTY = 8;
constexpr int kTiledWaves = 4;
constexpr int kRunsPerWave = kTZ * kTY / kTiledWaves;
constexpr int kHZ = kTZ + 2, kHY = kTY + 2, kHX = 16 + 2;  // block + halo (voxels)

__host__ __device__ inline bool tiled_shape_ok(int D, int H, int W, int dil) {
  return dil >= 1 && W % (16 * dil) == 0 && H % (kTY * dil) == 0 && D % (kTZ * dil) == 0;
}

template <int NB>
__global__ void __launch_bounds__(64 * kTiledWaves) conv3_tiled_kernel(
    const __bf16 *__restrict__ x, const __bf16 *__restrict__ w,
    const float *__restrict__ bias, int D, int H, int W, int dil,
    __bf16 *__restrict__ y, float *__restrict__ yf, int relu) {
  __shared__ __align__(16) __bf16 blk[kHZ * kHY * kHX * kC];
  const int Ds = D / dil, Hs = H / dil, Ws = W / dil;  // sub-lattice extents
  const int nbz = Ds / kTZ, nby = Hs / kTY, nbx = Ws / 16;
  int t = blockIdx.x;
  const int bx = t % nbx;
  t /= nbx;
  const int by = t % nby;
  t /= nby;
  const int bz = t % nbz;
  t /= nbz;
  const int rx = t % dil;
  t /= dil;
  const int ry = t % dil;
  t /= dil;
  const int rz = t % dil;
  const int b = t / dil;
  // the block with its halo, in sub-lattice coordinates from (z0, y0, x0);
  // out of the volume is the zero padding of conv3_kernel
  const int z0 = bz * kTZ - 1, y0 = by * kTY - 1, x0 = bx * 16 - 1;
  // every thread issues all of its loads before its first LDS store, so the
  // block's memory latency is paid once, not once per load
  constexpr int kQuarters = kHZ * kHY * kHX * 4;
  constexpr int kThreadsT = 64 * kTiledWaves;
  constexpr int kLoads = (kQuarters + kThreadsT - 1) / kThreadsT;
  bf16x8 vals[kLoads];
#pragma unroll
  for (int k = 0; k < kLoads; ++k) {
    const int i = (int)threadIdx.x + k * kThreadsT;
    vals[k] = bf16x8{};
    if (i < kQuarters) {
      const int q = i & 3;  // 16-byte quarter of the voxel's 32 channels
      const int v = i >> 2;
      const int hx = v % kHX, hy = (v / kHX) % kHY, hz = v / (kHX * kHY);
      const int sz = z0 + hz, sy = y0 + hy, sx = x0 + hx;
      if ((unsigned)sz < (unsigned)Ds && (unsigned)sy < (unsigned)Hs &&
          (unsigned)sx < (unsigned)Ws) {
        const size_t vox = (((size_t)b * D + (rz + dil * sz)) * H + (ry + dil * sy)) * W +
                           (rx + dil * sx);
        vals[k] = *reinterpret_cast<const bf16x8 *>(x + vox * kC + 8 * q);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kLoads; ++k) {
    const int i = (int)threadIdx.x + k * kThreadsT;
    if (i < kQuarters) *reinterpret_cast<bf16x8 *>(blk + (size_t)i * 8) = vals[k];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int vl = lane & 15;  // voxel (A row) / output channel (B column)
  const int kg = lane >> 4;  // k group: channels 8*kg .. 8*kg+7
  const int NOUT = 16 * NB;
  f32x4 acc[kRunsPerWave][NB];
#pragma unroll
  for (int r = 0; r < kRunsPerWave; ++r)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[r][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  // each tap's weights are loaded one tap ahead
  bf16x8 bw[NB], bn[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
    bn[nb] = *reinterpret_cast<const bf16x8 *>(w + ((size_t)nb * 16 + vl) * kC + 8 * kg);
#pragma unroll 1
  for (int tap = 0; tap < kTaps; ++tap) {
    const int kd = tap / 9, kh = (tap / 3) % 3, kw = tap % 3;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bw[nb] = bn[nb];
    if (tap + 1 < kTaps) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        bn[nb] = *reinterpret_cast<const bf16x8 *>(
            w + ((size_t)(tap + 1) * NOUT + nb * 16 + vl) * kC + 8 * kg);
    }
#pragma unroll
    for (int r = 0; r < kRunsPerWave; ++r) {
      const int j = wave * kRunsPerWave + r;
      const int jz = j / kTY, jy = j % kTY;
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(
          blk + ((size_t)((jz + kd) * kHY + (jy + kh)) * kHX + (vl + kw)) * kC + 8 * kg);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[nb], acc[r][nb], 0, 0, 0);
    }
  }
  if (NB == 2 && y && !yf) {
    // 32-channel output: through the (now free) LDS block as [run][16
    // voxels][32 channels], then written out in 16-byte pieces, a voxel's
    // 64 bytes by 4 neighbouring lanes
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRunsPerWave; ++r) {
      const int j = wave * kRunsPerWave + r;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int co = nb * 16 + vl;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float o = acc[r][nb][i] + (bias ? bias[co] : 0.f);
          if (relu) o = fmaxf(o, 0.f);
          blk[((size_t)j * 16 + 4 * kg + i) * kC + co] = f2bf(o);
        }
      }
    }
    __syncthreads();
    constexpr int kOut = kTZ * kTY * 16 * 4;  // 16-byte pieces of the block's outputs
#pragma unroll
    for (int k = 0; k < kOut / kThreadsT; ++k) {
      const int i = (int)threadIdx.x + k * kThreadsT;
      const int q = i & 3, vx = (i >> 2) & 15, j = i >> 6;
      const int gz = rz + dil * (bz * kTZ + (j / kTY)), gy = ry + dil * (by * kTY + (j % kTY));
      const size_t v = (((size_t)b * D + gz) * H + gy) * W + rx + dil * (bx * 16 + vx);
      *reinterpret_cast<bf16x8 *>(y + v * kC + 8 * q) =
          *reinterpret_cast<const bf16x8 *>(blk + (size_t)i * 8);
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < kRunsPerWave; ++r) {
    const int j = wave * kRunsPerWave + r;
    const int gz = rz + dil * (bz * kTZ + (j / kTY)), gy = ry + dil * (by * kTY + (j % kTY));
    const size_t row = (((size_t)b * D + gz) * H + gy) * W;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = nb * 16 + vl;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const size_t v = row + rx + dil * (bx * 16 + 4 * kg + i);
        float o = acc[r][nb][i] + (bias ? bias[co] : 0.f);
        if (relu) o = fmaxf(o, 0.f);
        if (y) y[v * kC + co] = f2bf(o);
        if (yf && co == 0) yf[v] = o;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[tap][cout][cin] += sum over voxels of
// dy[v][cout] * x[v + tap - 1][cin].  A wave owns (tap, cout block) and both
// cin blocks and sweeps `rows_per_wave` runs of 32 voxels.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) wgrad_kernel(
    const __bf16 *__restrict__ dy, const __bf16 *__restrict__ x, int nb_out,
    int nrow, int rows_per_wave, int D, int H, int W, int dil,
    float *__restrict__ dW, int NOUT) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nchunks = (nrow + rows_per_wave - 1) / rows_per_wave;
  const int job = wave;
  if (job >= kTaps * nb_out * nchunks) return;
  const int chunk = job % nchunks;
  const int mb = (job / nchunks) % nb_out;
  const int tap = job / (nchunks * nb_out);
  const int kd = tap / 9, kh = (tap / 3) % 3, kw = tap % 3;
  const int ml = lane & 15;   // A row: cout; B column: cin
  const int kg = lane >> 4;   // voxels 8*kg .. 8*kg+7 of the 32-voxel run
  f32x4 acc0 = f32x4{0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
  const int wr = W >> 5;
  const int r0 = chunk * rows_per_wave, r1 = min(nrow, r0 + rows_per_wave);
  for (int row = r0; row < r1; ++row) {
    int t = row;
    const int w0 = (t % wr) << 5;
    t /= wr;
    const int h = t % H;
    t /= H;
    const int d = t % D;
    const int b = t / D;
    const size_t vbase = (((size_t)b * D + d) * H + h) * W + w0 + 8 * kg;
    bf16x8 a, b0, b1;
    const int zd = d + (kd - 1) * dil, zh = h + (kh - 1) * dil;
    const bool rowin = (unsigned)zd < (unsigned)D && (unsigned)zh < (unsigned)H;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      a[j] = dy[(vbase + j) * kC + mb * 16 + ml];
      const int zw = w0 + 8 * kg + j + (kw - 1) * dil;
      if (rowin && (unsigned)zw < (unsigned)W) {
        const size_t vx = (((size_t)b * D + zd) * H + zh) * W + zw;
        b0[j] = x[vx * kC + ml];
        b1[j] = x[vx * kC + 16 + ml];
      } else {
        b0[j] = (__bf16)0.f;
        b1[j] = (__bf16)0.f;
      }
    }
    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
  }
  // accumulator: rows (cout) 4*kg..+3, column (cin) ml
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int co = mb * 16 + 4 * kg + i;
    atomicAdd(&dW[((size_t)tap * NOUT + co) * kC + ml], acc0[i]);
    atomicAdd(&dW[((size_t)tap * NOUT + co) * kC + 16 + ml], acc1[i]);
  }
}

// db[c] += sum over voxels of dy[v][c], c < nc
__global__ void bgrad_kernel(const __bf16 *__restrict__ dy, size_t nvox, int nc,
                             float *__restrict__ db) {
  __shared__ float sh[kC];
  if (threadIdx.x < kC) sh[threadIdx.x] = 0.f;
  __syncthreads();
  const int c = threadIdx.x & 31;
  float s = 0.f;
  if (c < nc) {
    for (size_t v = (size_t)blockIdx.x * (blockDim.x >> 5) + (threadIdx.x >> 5); v < nvox;
         v += (size_t)gridDim.x * (blockDim.x >> 5))
      s += bf2f(dy[v * kC + c]);
    atomicAdd(&sh[c], s);
  }
  __syncthreads();
  if (threadIdx.x < nc) atomicAdd(&db[threadIdx.x], sh[threadIdx.x]);
}

// Weighted mean squared error on logits: with e = z - t and w = w_pos for
// e > 0 (a prediction more neutral than the target), 1 otherwise,
// loss += sum w e^2 / n; dy[v][0] = 2 w e / n, other channels 0.
__global__ void mse_kernel(const float *__restrict__ z, const float *__restrict__ tgt,
                           size_t n, float w_pos, __bf16 *__restrict__ dy,
                           float *__restrict__ loss) {
  __shared__ float sh[4];
  float s = 0.f;
  const float inv = 1.f / (float)n;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (size_t)gridDim.x * blockDim.x) {
    const float e = z[v] - tgt[v];
    const float w = e > 0.f ? w_pos : 1.f;
    s += w * e * e;
    __bf16 *o = dy + v * kC;
    o[0] = f2bf(2.f * w * e * inv);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    atomicAdd(loss, t * inv);
  }
}

// Adam (Kingma & Ba 2015) on FP32 master parameters.
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, size_t n,
                            float lr, float b1, float b2, float eps, float c1,
                            float c2, float wd) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// FP32 master layout of layer l (padded to the MFMA shapes): W[27][NOUT][32]
// (tap, output channel, input channel; entries beyond the real channels stay
// 0: their gradients are 0) then b[NOUT].  Packed BF16 GEMM weights: fwd =
// W as is, and the input-gradient weights bwd[t'][cin][cout] = W[26-t'][cout]
// [cin] ([27][32][32], zero for cout >= NOUT).
__global__ void pack_kernel(const float *__restrict__ master, int NOUT,
                            __bf16 *__restrict__ fwd, __bf16 *__restrict__ bwd) {
  const int n = kTaps * 32 * 32;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int tap = i / 1024, o = (i / 32) % 32, k = i % 32;
    if (o < NOUT) fwd[((size_t)tap * NOUT + o) * kC + k] = f2bf(master[((size_t)tap * NOUT + o) * kC + k]);
    const float vb = k < NOUT ? master[((size_t)(kTaps - 1 - tap) * NOUT + k) * kC + o] : 0.f;
    bwd[((size_t)tap * 32 + o) * kC + k] = f2bf(vb);
  }
}

// Input features of one sample (log10 n_H, log10 jH, FP32) -> BF16 channels
// 0, 1 of the activation tensor (normalised), zeros elsewhere; an optional
// crop [o..o+S) of a larger sample.
__global__ void features_kernel(const float *__restrict__ feat, int D, int H, int W,
                                int od, int oh, int ow, int S0, int S1, int S2,
                                float m0, float s0, float m1, float s1,
                                __bf16 *__restrict__ out) {
  const size_t n = (size_t)S0 * S1 * S2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * kC;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t v = i / kC;
    const int c = (int)(i % kC);
    float r = 0.f;
    if (c < 2) {
      const int z = (int)(v % S2), y = (int)((v / S2) % S1), xx = (int)(v / ((size_t)S1 * S2));
      const size_t src = (((size_t)(od + xx) * H) + (oh + y)) * W + (ow + z);
      r = c == 0 ? (feat[2 * src] - m0) * s0 : (feat[2 * src + 1] - m1) * s1;
    }
    out[i] = f2bf(r);
  }
}

__global__ void crop_target_kernel(const float *__restrict__ t, int D, int H, int W,
                                   int od, int oh, int ow, int S0, int S1, int S2,
                                   float inv_scale, float *__restrict__ out) {
  const size_t n = (size_t)S0 * S1 * S2;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(v % S2), y = (int)((v / S2) % S1), xx = (int)(v / ((size_t)S1 * S2));
    out[v] = t[(((size_t)(od + xx) * H) + (oh + y)) * W + (ow + z)] * inv_scale;
  }
}

// Lattice cell (gx, gy, gz) of a context -> its device address (subgrids in
// the reference's order (sx * nsub_y + sy) * nsub_z + sz, tile-major inside).
__device__ __forceinline__ size_t lattice_addr(const GridDev &g, int gx, int gy, int gz) {
  const int sx = gx / g.sc[0], sy = gy / g.sc[1], sz = gz / g.sc[2];
  const int sub = (sx * g.nsub[1] + sy) * g.nsub[2] + sz;
  return cell_addr(g, sub, gx - sx * g.sc[0], gy - sy * g.sc[1], gz - sz * g.sc[2]);
}

constexpr double kJHFloor = 1e-12;  // s^-1, added before the logarithm

// logit of x_n clamped to [1e-7, 1 - 1/150]: above that every cell is
// opaque (a 256^3-workload cell of neutral gas has optical depth ~76), and
// the clamp keeps the Monte Carlo noise of near-neutral cells (1 - x ~ 1e-6)
// out of the features and targets
constexpr double kXnHigh = 1. - 1. / 150.;
__device__ __forceinline__ double clamped_logit(double x) {
  x = fmin(fmax(x, 1e-7), kXnHigh);
  return log(x / (1. - x));
}

// Features of a context's cells in lattice order: log10 n_H and, in input
// mode 0, log10(jconv * J + floor) (the photoionization rate of the last
// transport), in mode 1 the logit of the current x_n; normalised BF16 into
// act (channels 0, 1 of 32) and/or raw FP32 into feat ([v][2]).
__global__ void ctx_features_kernel(GridDev g, const double *__restrict__ n,
                                    const double *__restrict__ J,
                                    const double *__restrict__ xn, int mode, int D, int H,
                                    int W, double jconv, float m0, float s0, float m1,
                                    float s1, __bf16 *__restrict__ act,
                                    float *__restrict__ feat) {
  const size_t nv = (size_t)D * H * W;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (size_t)gridDim.x * blockDim.x) {
    const int gz = (int)(v % W), gy = (int)((v / W) % H), gx = (int)(v / ((size_t)W * H));
    const size_t a = lattice_addr(g, gx, gy, gz);
    const float f0 = (float)log10(n[a]);
    const float f1 = mode == 1 ? (float)clamped_logit(xn[a])
                               : (float)log10(jconv * J[a] + kJHFloor);
    if (feat) {
      feat[2 * v] = f0;
      feat[2 * v + 1] = f1;
    }
    if (act) {
      bf16x8 p0 = bf16x8{}, pz = bf16x8{};
      p0[0] = f2bf((f0 - m0) * s0);
      p0[1] = f2bf((f1 - m1) * s1);
      bf16x8 *o = reinterpret_cast<bf16x8 *>(act + v * kC);
      o[0] = p0;
      o[1] = pz;
      o[2] = pz;
      o[3] = pz;
    }
  }
}

// Network outputs (lattice order) -> x_n of the context, J = 0, FP32
// opacity.  Mode 0: the output is the logit of x_n; mode 1: it is added to
// the logit of the current x_n.  A cell whose logit exceeds `cut`
// (predicted neutral) starts at `fallback` instead.
__global__ void ctx_apply_kernel(GridDev g, const float *__restrict__ z, int mode, int D,
                                 int H, int W, float out_scale, float cut, double fallback,
                                 const double *__restrict__ n,
                                 double *__restrict__ xn, double *__restrict__ J,
                                 float *__restrict__ kappa, double kscale) {
  const size_t nv = (size_t)D * H * W;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (size_t)gridDim.x * blockDim.x) {
    const int gz = (int)(v % W), gy = (int)((v / W) % H), gx = (int)(v / ((size_t)W * H));
    const size_t a = lattice_addr(g, gx, gy, gz);
    const double logit =
        (double)z[v] * (double)out_scale + (mode == 1 ? clamped_logit(xn[a]) : 0.);
    double x = 1. / (1. + exp(-logit));
    x = fmin(fmax(x, 1e-7), 1.);
    if (logit > (double)cut) x = fallback;
    xn[a] = x;
    J[a] = 0.;
    if (kappa) kappa[a] = (float)(kscale * n[a] * x);
  }
}

// Coarse start of the warm start.  restrict: coarse n_H = the mean of each
// f^3 block of fine cells (the mass of the block), coarse x_n = 1e-6, J = 0,
// FP32 opacity of that state.
__global__ void restrict_density_kernel(GridDev gf, const double *__restrict__ nf,
                                        GridDev gc, int f, int Dc, int Hc, int Wc,
                                        double *__restrict__ nc, double *__restrict__ xnc,
                                        double *__restrict__ Jc, float *__restrict__ kappa,
                                        double kscale) {
  const size_t nv = (size_t)Dc * Hc * Wc;
  const double inv = 1. / ((double)f * f * f);
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (size_t)gridDim.x * blockDim.x) {
    const int cz = (int)(v % Wc), cy = (int)((v / Wc) % Hc), cx = (int)(v / ((size_t)Wc * Hc));
    double s = 0.;
    for (int i = 0; i < f; ++i)
      for (int j = 0; j < f; ++j)
        for (int k = 0; k < f; ++k)
          s += nf[lattice_addr(gf, cx * f + i, cy * f + j, cz * f + k)];
    const size_t a = lattice_addr(gc, cx, cy, cz);
    const double n = s * inv;
    nc[a] = n;
    xnc[a] = 1e-6;
    Jc[a] = 0.;
    if (kappa) kappa[a] = (float)(kscale * n * 1e-6);
  }
}

// prolong: the fine x_n is the coarse field's logit (clamped like the
// features) interpolated trilinearly at the fine cell's centre (coarse cell
// centres at (c + 1/2) f fine cells, the edge cells extended), J = 0, FP32
// opacity refreshed.
__global__ void prolong_xn_kernel(GridDev gc, const double *__restrict__ xnc, int f,
                                  int Dc, int Hc, int Wc, GridDev gf,
                                  const double *__restrict__ nf, double *__restrict__ xnf,
                                  double *__restrict__ Jf, float *__restrict__ kappa,
                                  double kscale) {
  const int Df = Dc * f, Hf = Hc * f, Wf = Wc * f;
  const size_t nv = (size_t)Df * Hf * Wf;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(v % Wf), y = (int)((v / Wf) % Hf), x = (int)(v / ((size_t)Wf * Hf));
    int c0[3], c1[3];
    double w[3];
    const int fi[3] = {x, y, z}, nc[3] = {Dc, Hc, Wc};
    for (int a = 0; a < 3; ++a) {
      const double u = ((double)fi[a] + 0.5) / (double)f - 0.5;
      const double fl = floor(u);
      w[a] = u - fl;
      c0[a] = min(max((int)fl, 0), nc[a] - 1);
      c1[a] = min(max((int)fl + 1, 0), nc[a] - 1);
    }
    double l = 0.;
    for (int k = 0; k < 8; ++k) {
      const int ix = (k & 4) ? c1[0] : c0[0], iy = (k & 2) ? c1[1] : c0[1],
                iz = (k & 1) ? c1[2] : c0[2];
      const double wk = ((k & 4) ? w[0] : 1. - w[0]) * ((k & 2) ? w[1] : 1. - w[1]) *
                        ((k & 1) ? w[2] : 1. - w[2]);
      l += wk * clamped_logit(xnc[lattice_addr(gc, ix, iy, iz)]);
    }
    const size_t a = lattice_addr(gf, x, y, z);
    const double xn = 1. / (1. + exp(-l));
    xnf[a] = xn;
    Jf[a] = 0.;
    if (kappa) kappa[a] = (float)(kscale * nf[a] * xn);
  }
}

__global__ void zero_f32_kernel(float *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

__global__ void zero_bf16_kernel(__bf16 *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = f2bf(0.f);
}

}  // namespace cmi

using namespace cmi;

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct cmi_net {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t nparam = 0;
  size_t off_w[kLayers], off_b[kLayers];  // offsets into the master vector
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32
  __bf16 *wf[kLayers] = {}, *wb[kLayers] = {};                    // packed
  int step = 0;
  // normalisation of the features and of the target logits
  float mean[2] = {8.f, -10.f}, scale[2] = {1.f, 0.25f};
  float out_scale = 8.f;  // logit = out_scale * network output
  // start rule of the warm start: cells predicted more neutral than logit
  // `cut` start at `fallback` (default: no cut)
  float cut = 3.0e38f;
  double fallback = 1e-6;
  int dil[kLayers] = {1, 1, 1, 1};  // dilation of each layer's 3x3x3 taps
  float w_pos = 1.f;  // loss weight of predictions more neutral than the target
  int mode = 0;       // input mode of the warm start (0: jH, 1: refine x_n)
  int tiled = 1;      // forward convolutions through conv3_tiled_kernel where it fits
  // dataset (device): features [S][N][2] FP32, targets [S][N] FP32
  float *ds_feat = nullptr, *ds_tgt = nullptr;
  int ds_n = 0, ds_dim[3] = {0, 0, 0};
  size_t ds_cap = 0;
  // work buffers (sized for the largest batch seen)
  size_t work_vox = 0;
  __bf16 *act[kLayers] = {};  // act[0] = input, act[l] = output of layer l-1
  __bf16 *dact[2] = {};       // gradient ping-pong
  float *z = nullptr, *tgt = nullptr, *loss = nullptr;
  float *feat_tmp = nullptr;
  size_t feat_cap = 0;
};

namespace {

int nfail(cmi_net *n, int code, const std::string &m) {
  if (n) n->err = m;
  return code;
}

#define NET_TRY(n, call)                                                   \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return nfail(n, CMI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int nout_of(int l) { return layer_cout(l) > 16 ? 32 : 16; }

int ensure_work(cmi_net *n, size_t nvox) {
  if (nvox <= n->work_vox) return CMI_OK;
  for (int l = 0; l < kLayers; ++l) hipFree(n->act[l]);
  for (int k = 0; k < 2; ++k) hipFree(n->dact[k]);
  hipFree(n->z);
  hipFree(n->tgt);
  n->work_vox = 0;
  for (int l = 0; l < kLayers; ++l) NET_TRY(n, hipMalloc(&n->act[l], nvox * kC * sizeof(__bf16)));
  for (int k = 0; k < 2; ++k) NET_TRY(n, hipMalloc(&n->dact[k], nvox * kC * sizeof(__bf16)));
  NET_TRY(n, hipMalloc(&n->z, nvox * sizeof(float)));
  NET_TRY(n, hipMalloc(&n->tgt, nvox * sizeof(float)));
  n->work_vox = nvox;
  return CMI_OK;
}

int pack(cmi_net *n) {
  for (int l = 0; l < kLayers; ++l)
    hipLaunchKernelGGL(pack_kernel, dim3(108), dim3(256), 0, n->stream, n->p + n->off_w[l],
                       nout_of(l), n->wf[l], n->wb[l]);
  NET_TRY(n, hipGetLastError());
  return CMI_OK;
}

unsigned grid_for(size_t items, unsigned per_block, unsigned cap = 65535) {
  size_t g = (items + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return (unsigned)(g < cap ? g : cap);
}

// forward over act[0] (nvox voxels in B samples of D x H x W) -> z
int forward(cmi_net *n, int B, int D, int H, int W, float *zout) {
  const int nrun = B * D * H * (W / 16);
  const unsigned blocks = grid_for((size_t)nrun, 4, 1u << 30);  // one wave per run
  for (int l = 0; l < kLayers; ++l) {
    const float *bias = n->p + n->off_b[l];
    const int dil = n->dil[l];
    if (n->tiled && tiled_shape_ok(D, H, W, dil)) {
      // the same outputs, blocks of a dilation sub-lattice staged in LDS
      const unsigned tb = (unsigned)((size_t)B * D * H * W / (16 * kTZ * kTY));
      if (l < kLayers - 1)
        hipLaunchKernelGGL(conv3_tiled_kernel<2>, dim3(tb), dim3(64 * kTiledWaves), 0,
                           n->stream, n->act[l], n->wf[l], bias, D, H, W, dil, n->act[l + 1],
                           (float *)nullptr, 1);
      else
        hipLaunchKernelGGL(conv3_tiled_kernel<1>, dim3(tb), dim3(64 * kTiledWaves), 0,
                           n->stream, n->act[l], n->wf[l], bias, D, H, W, dil,
                           (__bf16 *)nullptr, zout, 0);
      continue;
    }
    if (l < kLayers - 1) {
      hipLaunchKernelGGL(conv3_kernel<2>, dim3(blocks), dim3(256), 0, n->stream, n->act[l],
                         n->wf[l], bias, nrun, D, H, W, n->dil[l], n->act[l + 1], (float *)nullptr,
                         (const __bf16 *)nullptr, 1);
    } else {
      hipLaunchKernelGGL(conv3_kernel<1>, dim3(blocks), dim3(256), 0, n->stream, n->act[l],
                         n->wf[l], bias, nrun, D, H, W, n->dil[l], (__bf16 *)nullptr, zout,
                         (const __bf16 *)nullptr, 0);
    }
  }
  NET_TRY(n, hipGetLastError());
  return CMI_OK;
}

// backward from the MSE loss on z vs tgt; gradients accumulate into n->g
int backward(cmi_net *n, int B, int D, int H, int W) {
  const size_t nvox = (size_t)B * D * H * W;
  const int nrun = B * D * H * (W / 16);
  const unsigned blocks = grid_for((size_t)nrun, 4, 1u << 30);  // one wave per run
  const int nrow = B * D * H * (W / 32);
  const int rows_per_wave = 16;
  const int nchunks = (nrow + rows_per_wave - 1) / rows_per_wave;
  __bf16 *dy = n->dact[0], *dx = n->dact[1];
  hipLaunchKernelGGL(zero_bf16_kernel, dim3(grid_for(nvox * kC, 256, 4096)), dim3(256), 0,
                     n->stream, dy, nvox * kC);
  hipLaunchKernelGGL(mse_kernel, dim3(grid_for(nvox, 256, 1024)), dim3(256), 0, n->stream,
                     n->z, n->tgt, nvox, n->w_pos, dy, n->loss);
  for (int l = kLayers - 1; l >= 0; --l) {
    const int NOUT = nout_of(l);
    const int nb_out = NOUT / 16;
    const size_t waves = (size_t)kTaps * nb_out * nchunks;
    hipLaunchKernelGGL(wgrad_kernel, dim3(grid_for(waves, 4, 1u << 30)), dim3(256), 0, n->stream, dy,
                       n->act[l], nb_out, nrow, rows_per_wave, D, H, W, n->dil[l],
                       n->g + n->off_w[l], NOUT);
    hipLaunchKernelGGL(bgrad_kernel, dim3(grid_for(nvox, 8, 2048)), dim3(256), 0, n->stream,
                       dy, nvox, layer_cout(l), n->g + n->off_b[l]);
    if (l > 0) {
      // dx = conv(dy, flipped transposed W), masked by the ReLU of layer l-1
      hipLaunchKernelGGL(conv3_kernel<2>, dim3(blocks), dim3(256), 0, n->stream, dy, n->wb[l],
                         (const float *)nullptr, nrun, D, H, W, n->dil[l], dx, (float *)nullptr,
                         (const __bf16 *)n->act[l], 0);
      std::swap(dy, dx);
    }
  }
  NET_TRY(n, hipGetLastError());
  return CMI_OK;
}

}  // namespace

extern "C" {

int cmi_net_create(uint64_t seed, cmi_net **out) {
  if (!out) return CMI_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CMI_ERR_NODEVICE;
  cmi_net *n = new cmi_net();
  (void)hipGetDevice(&n->device);
  size_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    n->off_w[l] = off;
    off += (size_t)kTaps * nout_of(l) * kC;
    n->off_b[l] = off;
    off += nout_of(l);
  }
  n->nparam = off;
  hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&n->p, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->g, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->m, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->v, off * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&n->loss, 65536 * sizeof(float));
  for (int l = 0; l < kLayers && e == hipSuccess; ++l) {
    e = hipMalloc(&n->wf[l], (size_t)kTaps * nout_of(l) * kC * sizeof(__bf16));
    if (e == hipSuccess) e = hipMalloc(&n->wb[l], (size_t)kTaps * 32 * kC * sizeof(__bf16));
  }
  if (e != hipSuccess) {
    cmi_net_destroy(n);
    return CMI_ERR_ALLOC;
  }
  // He-uniform initial weights over the real fan-in, zero biases, from a
  // Philox stream
  std::vector<float> h(off, 0.f);
  uint64_t ctr = 0;
  for (int l = 0; l < kLayers; ++l) {
    const float bound = sqrtf(6.f / (kTaps * layer_cin(l))) * (l == kLayers - 1 ? 0.5f : 1.f);
    const int NOUT = nout_of(l);
    for (int tap = 0; tap < kTaps; ++tap)
      for (int o = 0; o < layer_cout(l); ++o)
        for (int k = 0; k < layer_cin(l); ++k) {
          const u32x4 r = philox4x32_10(u32x4{(uint32_t)ctr, (uint32_t)(ctr >> 32), 0xC0u, 0u},
                                        (uint32_t)seed, (uint32_t)(seed >> 32));
          ++ctr;
          const double u = uniform_open(r.x, r.y);
          h[n->off_w[l] + ((size_t)tap * NOUT + o) * kC + k] = (float)((2. * u - 1.) * bound);
        }
  }
  e = hipMemcpy(n->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(n->m, 0, off * sizeof(float));
  if (e == hipSuccess) e = hipMemset(n->v, 0, off * sizeof(float));
  if (e != hipSuccess || pack(n) != CMI_OK || hipStreamSynchronize(n->stream) != hipSuccess) {
    cmi_net_destroy(n);
    return CMI_ERR_HIP;
  }
  *out = n;
  return CMI_OK;
}

void cmi_net_destroy(cmi_net *n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  hipFree(n->p);
  hipFree(n->g);
  hipFree(n->m);
  hipFree(n->v);
  hipFree(n->loss);
  for (int l = 0; l < kLayers; ++l) {
    hipFree(n->wf[l]);
    hipFree(n->wb[l]);
    hipFree(n->act[l]);
  }
  for (int k = 0; k < 2; ++k) hipFree(n->dact[k]);
  hipFree(n->z);
  hipFree(n->tgt);
  hipFree(n->ds_feat);
  hipFree(n->ds_tgt);
  hipFree(n->feat_tmp);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

const char *cmi_net_last_error(const cmi_net *n) { return n ? n->err.c_str() : "null net"; }

size_t cmi_net_param_count(const cmi_net *n) { return n ? n->nparam : 0; }

int cmi_net_get_params(cmi_net *n, float *out) {
  if (!n || !out) return CMI_ERR_ARG;
  (void)hipSetDevice(n->device);
  NET_TRY(n, hipStreamSynchronize(n->stream));
  NET_TRY(n, hipMemcpy(out, n->p, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return CMI_OK;
}

int cmi_net_set_params(cmi_net *n, const float *in) {
  if (!n || !in) return CMI_ERR_ARG;
  (void)hipSetDevice(n->device);
  NET_TRY(n, hipStreamSynchronize(n->stream));  // null-stream copies below
  NET_TRY(n, hipMemcpy(n->p, in, n->nparam * sizeof(float), hipMemcpyHostToDevice));
  NET_TRY(n, hipMemset(n->m, 0, n->nparam * sizeof(float)));
  NET_TRY(n, hipMemset(n->v, 0, n->nparam * sizeof(float)));
  n->step = 0;
  const int rc = pack(n);
  if (rc != CMI_OK) return rc;
  NET_TRY(n, hipStreamSynchronize(n->stream));
  return CMI_OK;
}

int cmi_net_set_normalisation(cmi_net *n, const float mean[2], const float scale[2],
                              float logit_scale) {
  if (!n || !mean || !scale || !(logit_scale > 0.f)) return CMI_ERR_ARG;
  n->out_scale = logit_scale;
  for (int c = 0; c < 2; ++c) {
    n->mean[c] = mean[c];
    n->scale[c] = scale[c];
  }
  return CMI_OK;
}

int cmi_net_add_samples(cmi_net *n, const float *feat, const float *target, int nsamples,
                        const int dim[3]) {
  if (!n || !feat || !target || nsamples <= 0 || !dim) return CMI_ERR_ARG;
  (void)hipSetDevice(n->device);
  NET_TRY(n, hipStreamSynchronize(n->stream));  // null-stream copies below
  if (n->ds_n > 0 && (dim[0] != n->ds_dim[0] || dim[1] != n->ds_dim[1] || dim[2] != n->ds_dim[2]))
    return nfail(n, CMI_ERR_ARG, "all samples must have the same shape");
  const size_t nv = (size_t)dim[0] * dim[1] * dim[2];
  const size_t need = (size_t)(n->ds_n + nsamples);
  if (need > n->ds_cap) {
    size_t cap = std::max(need, 2 * n->ds_cap);
    float *f2 = nullptr, *t2 = nullptr;
    NET_TRY(n, hipMalloc(&f2, cap * nv * 2 * sizeof(float)));
    NET_TRY(n, hipMalloc(&t2, cap * nv * sizeof(float)));
    if (n->ds_n) {
      NET_TRY(n, hipMemcpy(f2, n->ds_feat, (size_t)n->ds_n * nv * 2 * sizeof(float),
                           hipMemcpyDeviceToDevice));
      NET_TRY(n, hipMemcpy(t2, n->ds_tgt, (size_t)n->ds_n * nv * sizeof(float),
                           hipMemcpyDeviceToDevice));
    }
    hipFree(n->ds_feat);
    hipFree(n->ds_tgt);
    n->ds_feat = f2;
    n->ds_tgt = t2;
    n->ds_cap = cap;
  }
  NET_TRY(n, hipMemcpy(n->ds_feat + (size_t)n->ds_n * nv * 2, feat,
                       (size_t)nsamples * nv * 2 * sizeof(float), hipMemcpyHostToDevice));
  NET_TRY(n, hipMemcpy(n->ds_tgt + (size_t)n->ds_n * nv, target,
                       (size_t)nsamples * nv * sizeof(float), hipMemcpyHostToDevice));
  for (int a = 0; a < 3; ++a) n->ds_dim[a] = dim[a];
  n->ds_n += nsamples;
  return CMI_OK;
}

int cmi_net_sample_count(const cmi_net *n) { return n ? n->ds_n : 0; }

int cmi_net_set_dilation(cmi_net *n, const int dilation[4]) {
  if (!n || !dilation) return CMI_ERR_ARG;
  for (int l = 0; l < kLayers; ++l)
    if (dilation[l] < 1 || dilation[l] > 64) return nfail(n, CMI_ERR_ARG, "dilation must be 1..64");
  for (int l = 0; l < kLayers; ++l) n->dil[l] = dilation[l];
  return CMI_OK;
}

int cmi_net_set_input_mode(cmi_net *n, int mode) {
  if (!n || (mode != 0 && mode != 1)) return CMI_ERR_ARG;
  n->mode = mode;
  return CMI_OK;
}

int cmi_net_get_input_mode(const cmi_net *n) { return n ? n->mode : -1; }

int cmi_net_set_tiled(cmi_net *n, int on) {
  if (!n) return CMI_ERR_ARG;
  n->tiled = on ? 1 : 0;
  return CMI_OK;
}

int cmi_net_set_loss_weight(cmi_net *n, float w_too_neutral) {
  if (!n || !(w_too_neutral > 0.f)) return CMI_ERR_ARG;
  n->w_pos = w_too_neutral;
  return CMI_OK;
}

int cmi_net_set_start_rule(cmi_net *n, float logit_cut, double fallback_xn) {
  if (!n || !(fallback_xn > 0. && fallback_xn <= 1.) || logit_cut != logit_cut)
    return CMI_ERR_ARG;
  n->cut = logit_cut;
  n->fallback = fallback_xn;
  return CMI_OK;
}

int cmi_net_get_start_rule(const cmi_net *n, float *logit_cut, double *fallback_xn) {
  if (!n || !logit_cut || !fallback_xn) return CMI_ERR_ARG;
  *logit_cut = n->cut;
  *fallback_xn = n->fallback;
  return CMI_OK;
}

int cmi_net_get_dilation(const cmi_net *n, int dilation[4]) {
  if (!n || !dilation) return CMI_ERR_ARG;
  for (int l = 0; l < kLayers; ++l) dilation[l] = n->dil[l];
  return CMI_OK;
}

int cmi_net_train(cmi_net *n, int steps, int batch, const int crop[3], float lr,
                  uint64_t seed, float *losses) {
  if (!n || steps <= 0 || steps > 65536 || batch <= 0 || !crop || !(lr > 0.f))
    return nfail(n, CMI_ERR_ARG, "bad training arguments");
  if (n->ds_n == 0) return nfail(n, CMI_ERR_STATE, "no samples");
  for (int a = 0; a < 3; ++a)
    if (crop[a] <= 0 || crop[a] > n->ds_dim[a]) return nfail(n, CMI_ERR_ARG, "crop larger than samples");
  if (crop[2] % 32 != 0) return nfail(n, CMI_ERR_ARG, "crop: last extent must be a multiple of 32");
  (void)hipSetDevice(n->device);
  const size_t cv = (size_t)crop[0] * crop[1] * crop[2];
  int rc = ensure_work(n, cv * batch);
  if (rc != CMI_OK) return rc;
  const size_t nv = (size_t)n->ds_dim[0] * n->ds_dim[1] * n->ds_dim[2];
  hipLaunchKernelGGL(zero_f32_kernel, dim3(64), dim3(256), 0, n->stream, n->loss, (size_t)steps);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < batch; ++b) {
      const u32x4 r = philox4x32_10(u32x4{(uint32_t)s, (uint32_t)b, (uint32_t)n->step, 0x7A1Eu},
                                    (uint32_t)seed, (uint32_t)(seed >> 32));
      const int si = (int)(r.x % (uint32_t)n->ds_n);
      const int od = (int)(r.y % (uint32_t)(n->ds_dim[0] - crop[0] + 1));
      const int oh = (int)(r.z % (uint32_t)(n->ds_dim[1] - crop[1] + 1));
      const int ow = (int)(r.w % (uint32_t)(n->ds_dim[2] - crop[2] + 1));
      hipLaunchKernelGGL(features_kernel, dim3(grid_for(cv * kC, 256, 4096)), dim3(256), 0,
                         n->stream, n->ds_feat + (size_t)si * nv * 2, n->ds_dim[0],
                         n->ds_dim[1], n->ds_dim[2], od, oh, ow, crop[0], crop[1], crop[2],
                         n->mean[0], n->scale[0], n->mean[1], n->scale[1],
                         n->act[0] + (size_t)b * cv * kC);
      hipLaunchKernelGGL(crop_target_kernel, dim3(grid_for(cv, 256, 4096)), dim3(256), 0,
                         n->stream, n->ds_tgt + (size_t)si * nv, n->ds_dim[0], n->ds_dim[1],
                         n->ds_dim[2], od, oh, ow, crop[0], crop[1], crop[2],
                         1.f / n->out_scale, n->tgt + (size_t)b * cv);
    }
    hipLaunchKernelGGL(zero_f32_kernel, dim3(64), dim3(256), 0, n->stream, n->g, n->nparam);
    rc = forward(n, batch, crop[0], crop[1], crop[2], n->z);
    if (rc != CMI_OK) return rc;
    float *loss_slot = n->loss + s;
    std::swap(n->loss, loss_slot);  // backward accumulates into n->loss
    rc = backward(n, batch, crop[0], crop[1], crop[2]);
    std::swap(n->loss, loss_slot);
    if (rc != CMI_OK) return rc;
    ++n->step;
    const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
    hipLaunchKernelGGL(adam_kernel, dim3(128), dim3(256), 0, n->stream, n->p, n->g, n->m, n->v,
                       n->nparam, lr, b1, b2, 1e-8f, c1, c2, 0.f);
    rc = pack(n);
    if (rc != CMI_OK) return rc;
  }
  NET_TRY(n, hipGetLastError());
  if (losses)
    NET_TRY(n, hipMemcpyAsync(losses, n->loss, steps * sizeof(float), hipMemcpyDeviceToHost,
                              n->stream));
  NET_TRY(n, hipStreamSynchronize(n->stream));
  return CMI_OK;
}

// One forward + backward on the given batch (features [B][D][H][W][2]
// log10 n_H, log10 jH; targets [B][D][H][W] logits): the loss and the FP32
// gradient of every master parameter, no update.
int cmi_net_loss_and_grad(cmi_net *n, const float *feat, const float *target, int B,
                          const int dim[3], float *loss, float *grad) {
  if (!n || !feat || !target || B <= 0 || !dim) return CMI_ERR_ARG;
  if (dim[2] % 32 != 0) return nfail(n, CMI_ERR_ARG, "last extent must be a multiple of 32");
  (void)hipSetDevice(n->device);
  NET_TRY(n, hipStreamSynchronize(n->stream));  // null-stream copies below
  const size_t nv = (size_t)dim[0] * dim[1] * dim[2];
  int rc = ensure_work(n, nv * B);
  if (rc != CMI_OK) return rc;
  std::vector<float> ts(target, target + nv * B);
  for (float &t : ts) t /= n->out_scale;
  float *df = nullptr;
  NET_TRY(n, hipMalloc(&df, nv * B * 2 * sizeof(float)));
  hipError_t e = hipMemcpy(df, feat, nv * B * 2 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(n->tgt, ts.data(), nv * B * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(df);
    return nfail(n, CMI_ERR_HIP, "loss_and_grad: copy failed");
  }
  for (int b = 0; b < B; ++b)
    hipLaunchKernelGGL(features_kernel, dim3(grid_for(nv * kC, 256, 4096)), dim3(256), 0,
                       n->stream, df + (size_t)b * nv * 2, dim[0], dim[1], dim[2], 0, 0, 0,
                       dim[0], dim[1], dim[2], n->mean[0], n->scale[0], n->mean[1],
                       n->scale[1], n->act[0] + (size_t)b * nv * kC);
  hipLaunchKernelGGL(zero_f32_kernel, dim3(64), dim3(256), 0, n->stream, n->g, n->nparam);
  hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, n->stream, n->loss, (size_t)1);
  rc = forward(n, B, dim[0], dim[1], dim[2], n->z);
  if (rc == CMI_OK) rc = backward(n, B, dim[0], dim[1], dim[2]);
  if (rc == CMI_OK) {
    e = hipMemcpyAsync(grad, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) rc = nfail(n, CMI_ERR_HIP, "loss_and_grad failed");
  }
  hipFree(df);
  return rc;
}

// Logits of x_n for one grid (features [D][H][W][2]).
int cmi_net_predict(cmi_net *n, const float *feat, const int dim[3], float *logits) {
  if (!n || !feat || !dim || !logits) return CMI_ERR_ARG;
  if (dim[2] % 32 != 0) return nfail(n, CMI_ERR_ARG, "last extent must be a multiple of 32");
  (void)hipSetDevice(n->device);
  NET_TRY(n, hipStreamSynchronize(n->stream));  // null-stream copies below
  const size_t nv = (size_t)dim[0] * dim[1] * dim[2];
  int rc = ensure_work(n, nv);
  if (rc != CMI_OK) return rc;
  if (n->feat_cap < nv) {
    hipFree(n->feat_tmp);
    n->feat_tmp = nullptr;
    n->feat_cap = 0;
    NET_TRY(n, hipMalloc(&n->feat_tmp, nv * 2 * sizeof(float)));
    n->feat_cap = nv;
  }
  NET_TRY(n, hipMemcpy(n->feat_tmp, feat, nv * 2 * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(features_kernel, dim3(grid_for(nv * kC, 256, 8192)), dim3(256), 0,
                     n->stream, n->feat_tmp, dim[0], dim[1], dim[2], 0, 0, 0, dim[0], dim[1],
                     dim[2], n->mean[0], n->scale[0], n->mean[1], n->scale[1], n->act[0]);
  rc = forward(n, 1, dim[0], dim[1], dim[2], n->z);
  if (rc != CMI_OK) return rc;
  NET_TRY(n, hipMemcpyAsync(logits, n->z, nv * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  NET_TRY(n, hipStreamSynchronize(n->stream));
  for (size_t v = 0; v < nv; ++v) logits[v] *= n->out_scale;
  return CMI_OK;
}

namespace {

// jH = jconv * J of a context's stored intensity (J is kept in units of
// sigma * Delta_0 in FP32 contexts)
double context_jconv(const cmi_ctx *c, double luminosity, uint64_t npackets) {
  double j = luminosity / (double)npackets * c->g.inv_cell_size[0] * c->g.inv_cell_size[1] *
             c->g.inv_cell_size[2];
  if (c->precision != CMI_PREC_FP64) j *= c->sigma * c->g.cell_size[0];
  return j;
}

int check_context(cmi_net *n, const cmi_ctx *c, double luminosity, uint64_t npackets) {
  if (!c || npackets == 0 || !(luminosity > 0.))
    return nfail(n, CMI_ERR_ARG, "bad warm-start arguments");
  if (c->device != n->device) return nfail(n, CMI_ERR_ARG, "context on another device");
  return CMI_OK;
}

}  // namespace

int cmi_net_warm_start(cmi_net *n, cmi_ctx *c, double luminosity, uint64_t npackets) {
  if (!n) return CMI_ERR_ARG;
  int rc = check_context(n, c, luminosity, npackets);
  if (rc != CMI_OK) return rc;
  const int D = c->ncell[0], H = c->ncell[1], W = c->ncell[2];
  if (W % 16 != 0) return nfail(n, CMI_ERR_ARG, "ncell[2] must be a multiple of 16");
  (void)hipSetDevice(n->device);
  const size_t nv = (size_t)D * H * W;
  rc = ensure_work(n, nv);
  if (rc != CMI_OK) return rc;
  NET_TRY(n, hipStreamSynchronize(c->stream));  // the first pass has finished
  hipLaunchKernelGGL(ctx_features_kernel, dim3(grid_for(nv, 256, 8192)), dim3(256), 0,
                     n->stream, c->g, (const double *)c->f.n, (const double *)c->f.J,
                     (const double *)c->f.xn, n->mode, D, H, W,
                     context_jconv(c, luminosity, npackets), n->mean[0], n->scale[0],
                     n->mean[1], n->scale[1], n->act[0], (float *)nullptr);
  rc = forward(n, 1, D, H, W, n->z);
  if (rc != CMI_OK) return rc;
  const bool fp32 = c->precision == CMI_PREC_FP32;
  hipLaunchKernelGGL(ctx_apply_kernel, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, n->stream,
                     c->g, (const float *)n->z, n->mode, D, H, W, n->out_scale, n->cut,
                     n->fallback,
                     (const double *)c->f.n,
                     c->f.xn, c->f.J, fp32 ? c->f.kappa : (float *)nullptr,
                     c->sigma * c->g.cell_size[0]);
  NET_TRY(n, hipGetLastError());
  NET_TRY(n, hipStreamSynchronize(n->stream));
  return CMI_OK;
}

// A coarse context of the same box with ncell f times smaller per axis
// (f >= 2), on the same device; returns f or 0.
static int coarse_factor(const cmi_ctx *fine, const cmi_ctx *coarse) {
  if (!fine || !coarse || fine->device != coarse->device) return 0;
  const int f = fine->ncell[0] / coarse->ncell[0];
  if (f < 2) return 0;
  for (int a = 0; a < 3; ++a) {
    if (coarse->ncell[a] * f != fine->ncell[a]) return 0;
    if (fine->box[a] != coarse->box[a] || fine->box[3 + a] != coarse->box[3 + a]) return 0;
  }
  return f;
}

int cmi_restrict_density(cmi_ctx *fine, cmi_ctx *coarse) {
  const int f = coarse_factor(fine, coarse);
  if (!f)
    return coarse ? fail(coarse, CMI_ERR_ARG,
                         "restrict: the coarse grid must be the fine one's box with "
                         "ncell an integer factor >= 2 smaller per axis, on its device")
                  : CMI_ERR_ARG;
  (void)hipSetDevice(coarse->device);
  // the fine context's stream has written its fields
  CMI_TRY_HIP(fine, hipStreamSynchronize(fine->stream));
  const size_t nv = (size_t)coarse->ncell[0] * coarse->ncell[1] * coarse->ncell[2];
  const bool fp32 = coarse->precision == CMI_PREC_FP32;
  hipLaunchKernelGGL(restrict_density_kernel, dim3(grid_for(nv, 256, 8192)), dim3(256), 0,
                     coarse->stream, fine->g, (const double *)fine->f.n, coarse->g, f,
                     coarse->ncell[0], coarse->ncell[1], coarse->ncell[2], coarse->f.n,
                     coarse->f.xn, coarse->f.J, fp32 ? coarse->f.kappa : (float *)nullptr,
                     coarse->sigma * coarse->g.cell_size[0]);
  CMI_TRY_HIP(coarse, hipGetLastError());
  return CMI_OK;
}

int cmi_prolong_neutral_fraction(cmi_ctx *coarse, cmi_ctx *fine) {
  const int f = coarse_factor(fine, coarse);
  if (!f)
    return fine ? fail(fine, CMI_ERR_ARG,
                       "prolong: the coarse grid must be the fine one's box with "
                       "ncell an integer factor >= 2 smaller per axis, on its device")
                : CMI_ERR_ARG;
  (void)hipSetDevice(fine->device);
  CMI_TRY_HIP(coarse, hipStreamSynchronize(coarse->stream));
  const size_t nv = (size_t)fine->ncell[0] * fine->ncell[1] * fine->ncell[2];
  const bool fp32 = fine->precision == CMI_PREC_FP32;
  hipLaunchKernelGGL(prolong_xn_kernel, dim3(grid_for(nv, 256, 8192)), dim3(256), 0,
                     fine->stream, coarse->g, (const double *)coarse->f.xn, f,
                     coarse->ncell[0], coarse->ncell[1], coarse->ncell[2], fine->g,
                     (const double *)fine->f.n, fine->f.xn, fine->f.J,
                     fp32 ? fine->f.kappa : (float *)nullptr,
                     fine->sigma * fine->g.cell_size[0]);
  CMI_TRY_HIP(fine, hipGetLastError());
  return CMI_OK;
}

int cmi_net_context_features(cmi_net *n, cmi_ctx *c, double luminosity, uint64_t npackets,
                             float *features) {
  if (!n || !features) return CMI_ERR_ARG;
  int rc = check_context(n, c, luminosity, npackets);
  if (rc != CMI_OK) return rc;
  (void)hipSetDevice(n->device);
  const int D = c->ncell[0], H = c->ncell[1], W = c->ncell[2];
  const size_t nv = (size_t)D * H * W;
  float *d = nullptr;
  NET_TRY(n, hipStreamSynchronize(c->stream));
  NET_TRY(n, hipMalloc(&d, nv * 2 * sizeof(float)));
  hipLaunchKernelGGL(ctx_features_kernel, dim3(grid_for(nv, 256, 8192)), dim3(256), 0,
                     n->stream, c->g, (const double *)c->f.n, (const double *)c->f.J,
                     (const double *)c->f.xn, n->mode, D, H, W,
                     context_jconv(c, luminosity, npackets), 0.f, 1.f, 0.f, 1.f,
                     (__bf16 *)nullptr, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(features, d, nv * 2 * sizeof(float), hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  hipFree(d);
  if (e != hipSuccess) return nfail(n, CMI_ERR_HIP, std::string("context features: ") + hipGetErrorString(e));
  return CMI_OK;
}

}  // extern "C"
