// Learned warm start of the registration: a 3D convolutional encoder-decoder
// that maps a pair of binned images (the fixed and moving image at the
// half-resolution pyramid level) to the C x C x C lowest orthonormal-DCT
// coefficients of each displacement component (C = ncoef <= 16, fixed per
// network), trained on the GPU with its
// own forward and backward kernels.  No counterpart in the reference, which
// always starts the pyramid from the identity
// (BINDER/FastRegistration.py:206-225); the registration driven from the
// predicted field is the reference's EM (BINDER/src/cpu/Registration.cpp:122-217).
//
// Network (input grid G = nx x ny x nz, each a multiple of 8 in [32, 128]):
//   level 0 (G):     conv 8->16, conv 16->16            (input channels 0, 1 =
//                                                        fixed bin/(L-1),
//                                                        moving bin/(K-1);
//                                                        2..7 zero)
//   level 1 (G/2):   avg-pool, conv 16->32, conv 32->32
//   level 2 (G/4):   avg-pool, conv 32->48, conv 48->48 -> skip S3
//   level 3 (G/8):   avg-pool, conv 48->64, 3 x conv 64->64 -> S4
//   level 2 (G/4):   nearest up-sample S4, concat [up(S4), S3] (112 ch),
//                    conv 112->64, conv 64->64 -> S5
//   head:            1x1x1 conv 64->3 (FP32) -> U on the G/4 grid; the
//                    coefficients are the orthonormal DCT-II of U over that
//                    grid, the C x C x C lowest (zero at frequencies the
//                    grid does not have), divided by sqrt(|G/4|).
// Every conv is 3x3x3, stride 1, zero padding, bias, ReLU.
//
// Arithmetic: activations and the compute copies of the weights are BF16;
// the convolutions accumulate in FP32 on v_mfma_f32_16x16x32_bf16; master
// weights, the head, the loss and Adam are FP32 (loss reduced in FP64).
//
// Layouts: activations are channels-last [b][x][y][z][c] with a row stride
// `ld` (channels per voxel, so a concatenation is two channel ranges of one
// tensor); conv weights are [cout][27][cin] (tap t = 9(dx+1) + 3(dy+1) +
// (dz+1)) in the FP32 parameter vector, packed to BF16 [cout][kpad] for the
// forward pass and to [cin][kpadT] with the taps mirrored for the
// input-gradient pass, which is then the same kernel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "warmstart.hpp"

namespace binder {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4_t;

__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
// round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

constexpr int kHeadC = 64;         // channels entering the head
constexpr int kInC = 8;            // stored input channels (2 used)

// ---------------------------------------------------------------------------
// 3x3x3 convolution as an implicit GEMM: rows = voxels, columns = output
// channels, contraction k = tap * CIN + input channel (zero-padded to a
// multiple of 32).  A wavefront owns 32 consecutive voxels and all output
// channels (NT tiles of 16); each lane's A fragment is 8 consecutive channels
// of one neighbour voxel, a single 16-byte load from the channels-last
// tensor.  B fragments (weights, [cout][kpad]) are 16-byte loads that stay
// in the caches.  Epilogue: + bias, ReLU, or x (mask > 0) for the input
// gradient pass, BF16 store.
// ---------------------------------------------------------------------------
struct ConvArgs {
  const uint16_t *in;
  int ld_in, off_in;
  const uint16_t *w;
  int kpad;
  const float *bias;
  uint16_t *out;
  int ld_out, off_out;
  const uint16_t *mask;
  int ld_mask, off_mask;
  int relu;
  int X, Y, Z, nvox;
};

constexpr int kConvVoxPerBlock = 128;  // 4 wavefronts x 32 voxels

template <int CIN, int NT>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int vbase = blockIdx.x * kConvVoxPerBlock + wave * 32;
  int cx[2], cy[2], cz[2], vv[2];
  bool ok[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int v = vbase + 16 * m + r;
    ok[m] = v < a.nvox;
    const int vc = ok[m] ? v : 0;
    const int t = vc / a.Z;
    cz[m] = vc - t * a.Z;
    const int t2 = t / a.Y;
    cy[m] = t - t2 * a.Y;
    cx[m] = t2 % a.X;
    vv[m] = vc;
  }
  const int YZ = a.Y * a.Z;
  f32x4_t acc[2][NT];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  constexpr int KS = (27 * CIN + 31) / 32;
  const uint16_t *wrow = a.w + (size_t)r * a.kpad + 8 * g;
  for (int s = 0; s < KS; ++s) {
    const int kk = s * 32 + 8 * g;
    const int tap = kk / CIN;
    const int c = kk - tap * CIN;
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    bf16x8_t A[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      uint4 q = make_uint4(0u, 0u, 0u, 0u);
      const int px = cx[m] + dx, py = cy[m] + dy, pz = cz[m] + dz;
      if (tap < 27 && ok[m] && (unsigned)px < (unsigned)a.X && (unsigned)py < (unsigned)a.Y &&
          (unsigned)pz < (unsigned)a.Z)
        q = *reinterpret_cast<const uint4 *>(
            a.in + (size_t)(vv[m] + dx * YZ + dy * a.Z + dz) * a.ld_in + a.off_in + c);
      A[m] = __builtin_bit_cast(bf16x8_t, q);
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const uint4 bq =
          *reinterpret_cast<const uint4 *>(wrow + (size_t)16 * n * a.kpad + s * 32);
      const bf16x8_t B = __builtin_bit_cast(bf16x8_t, bq);
#pragma unroll
      for (int m = 0; m < 2; ++m)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][n], 0, 0, 0);
    }
  }
  // lane holds rows 4g..4g+3 (voxels) of column r (channel) of each tile
#pragma unroll
  for (int m = 0; m < 2; ++m) {
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int ch = 16 * n + r;
      const float bv = a.bias ? a.bias[ch] : 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int v = vbase + 16 * m + 4 * g + j;
        if (v >= a.nvox) continue;
        float y = acc[m][n][j] + bv;
        if (a.relu) y = fmaxf(y, 0.f);
        if (a.mask && !(bf2f(a.mask[(size_t)v * a.ld_mask + a.off_mask + ch]) > 0.f)) y = 0.f;
        a.out[(size_t)v * a.ld_out + a.off_out + ch] = f2bf(y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient of one tap: dW[co][tap][ci] = sum_v g[v][co] x[v + off][ci].
// A block owns one tap and one range of voxels; chunks of 64 voxels of g and
// of the shifted input are transposed into LDS (voxel index contiguous), so
// both MFMA operands are 16-byte LDS reads with the voxels as contraction.
// Partials per block; k_reduce_rows sums them in a fixed order.
// ---------------------------------------------------------------------------
struct WgradArgs {
  const uint16_t *x;
  int ld_x, off_x;
  const uint16_t *g;
  int ld_g, off_g;
  float *part;   // [P][COUT][27*CIN]
  float *bpart;  // [P][COUT] bias-gradient partials (written by the centre tap)
  int X, Y, Z, nvox, per_block;
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
  constexpr int CP = CIN < 16 ? 16 : CIN;
  constexpr int MT = COUT / 16, NTI = CP / 16, TILES = MT * NTI, TPW = (TILES + 3) / 4;
  constexpr int PITCH = 72;  // 64 voxels + 8 (16-byte aligned rows)
  __shared__ __align__(16) uint16_t gT[COUT * PITCH];
  __shared__ __align__(16) uint16_t xT[CP * PITCH];
  const int tap = blockIdx.y, p = blockIdx.x;
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int YZ = a.Y * a.Z;
  f32x4_t acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int v_begin = p * a.per_block;
  const int v_end = min(a.nvox, v_begin + a.per_block);
  if (CP > CIN)
    for (int i = threadIdx.x; i < (CP - CIN) * PITCH; i += 256) xT[CIN * PITCH + i] = 0;
  // the centre tap's blocks also sum g per channel: the bias gradient
  const bool bias_block = tap == 13 && a.bpart != nullptr;
  float bsum = 0.f;
  for (int v0 = v_begin; v0 < v_end; v0 += 64) {
    __syncthreads();
    constexpr int GG = COUT / 8;
    for (int it = threadIdx.x; it < 64 * GG; it += 256) {
      const int vox = it / GG, cg = it - vox * GG;
      const int v = v0 + vox;
      uint4 q = make_uint4(0u, 0u, 0u, 0u);
      if (v < v_end)
        q = *reinterpret_cast<const uint4 *>(a.g + (size_t)v * a.ld_g + a.off_g + 8 * cg);
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
      for (int j = 0; j < 8; ++j) gT[(8 * cg + j) * PITCH + vox] = e[j];
    }
    constexpr int XG = CIN / 8;
    for (int it = threadIdx.x; it < 64 * XG; it += 256) {
      const int vox = it / XG, cg = it - vox * XG;
      const int v = v0 + vox;
      uint4 q = make_uint4(0u, 0u, 0u, 0u);
      if (v < v_end) {
        const int t = v / a.Z;
        const int z = v - t * a.Z;
        const int t2 = t / a.Y;
        const int y = t - t2 * a.Y;
        const int x = t2 % a.X;
        const int px = x + dx, py = y + dy, pz = z + dz;
        if ((unsigned)px < (unsigned)a.X && (unsigned)py < (unsigned)a.Y &&
            (unsigned)pz < (unsigned)a.Z)
          q = *reinterpret_cast<const uint4 *>(
              a.x + (size_t)(v + dx * YZ + dy * a.Z + dz) * a.ld_x + a.off_x + 8 * cg);
      }
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
      for (int j = 0; j < 8; ++j) xT[(8 * cg + j) * PITCH + vox] = e[j];
    }
    __syncthreads();
    if (bias_block && threadIdx.x < COUT) {
      const uint16_t *row = &gT[threadIdx.x * PITCH];
#pragma unroll 8
      for (int v = 0; v < 64; ++v) bsum += bf2f(row[v]);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        const int tile = wave + 4 * t;
        if (tile < TILES) {
          const int mt = tile / NTI, nt = tile - mt * NTI;
          const bf16x8_t A = __builtin_bit_cast(
              bf16x8_t, *reinterpret_cast<const uint4 *>(&gT[(16 * mt + r) * PITCH + 32 * s + 8 * g]));
          const bf16x8_t B = __builtin_bit_cast(
              bf16x8_t, *reinterpret_cast<const uint4 *>(&xT[(16 * nt + r) * PITCH + 32 * s + 8 * g]));
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
        }
      }
    }
  }
  if (bias_block && threadIdx.x < COUT) a.bpart[(size_t)p * COUT + threadIdx.x] = bsum;
  // lane holds dW[co = 16 mt + 4g + j][ci = 16 nt + r]
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave + 4 * t;
    if (tile >= TILES) continue;
    const int mt = tile / NTI, nt = tile - mt * NTI;
    const int ci = 16 * nt + r;
    if (ci >= CIN) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = 16 * mt + 4 * g + j;
      a.part[((size_t)p * COUT + co) * (27 * CIN) + tap * CIN + ci] = acc[t][j];
    }
  }
}

// The same for narrow layers (cin, cout <= 32), where one tile per 64
// voxels leaves most of a block idle: every wavefront stages and contracts
// its own 64-voxel chunks in a private LDS region (no block barrier in the
// loop) and writes its own partial row p * 4 + wave.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad_wave(WgradArgs a) {
  constexpr int CP = CIN < 16 ? 16 : CIN;
  constexpr int MT = COUT / 16, NTI = CP / 16, TILES = MT * NTI;
  constexpr int PITCH = 72;
  __shared__ __align__(16) uint16_t gT_all[4][COUT * PITCH];
  __shared__ __align__(16) uint16_t xT_all[4][CP * PITCH];
  const int tap = blockIdx.y, p = blockIdx.x;
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int YZ = a.Y * a.Z;
  uint16_t *gT = gT_all[wave];
  uint16_t *xT = xT_all[wave];
  f32x4_t acc[TILES];
#pragma unroll
  for (int t = 0; t < TILES; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int v_begin = p * a.per_block;
  const int v_end = min(a.nvox, v_begin + a.per_block);
  if (CP > CIN)
    for (int i = lane; i < (CP - CIN) * PITCH; i += 64) xT[CIN * PITCH + i] = 0;
  const bool bias_block = tap == 13 && a.bpart != nullptr;
  float bsum = 0.f;
  for (int v0 = v_begin + 64 * wave; v0 < v_end; v0 += 256) {
    // the previous chunk's LDS reads are issued before these writes (a
    // wavefront's LDS operations complete in order)
    __builtin_amdgcn_wave_barrier();
    constexpr int GG = COUT / 8;
    for (int it = lane; it < 64 * GG; it += 64) {
      const int vox = it / GG, cg = it - vox * GG;
      const int v = v0 + vox;
      uint4 q = make_uint4(0u, 0u, 0u, 0u);
      if (v < v_end)
        q = *reinterpret_cast<const uint4 *>(a.g + (size_t)v * a.ld_g + a.off_g + 8 * cg);
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
      for (int j = 0; j < 8; ++j) gT[(8 * cg + j) * PITCH + vox] = e[j];
    }
    constexpr int XG = CIN / 8;
    for (int it = lane; it < 64 * XG; it += 64) {
      const int vox = it / XG, cg = it - vox * XG;
      const int v = v0 + vox;
      uint4 q = make_uint4(0u, 0u, 0u, 0u);
      if (v < v_end) {
        const int t = v / a.Z;
        const int z = v - t * a.Z;
        const int t2 = t / a.Y;
        const int y = t - t2 * a.Y;
        const int x = t2 % a.X;
        const int px = x + dx, py = y + dy, pz = z + dz;
        if ((unsigned)px < (unsigned)a.X && (unsigned)py < (unsigned)a.Y &&
            (unsigned)pz < (unsigned)a.Z)
          q = *reinterpret_cast<const uint4 *>(
              a.x + (size_t)(v + dx * YZ + dy * a.Z + dz) * a.ld_x + a.off_x + 8 * cg);
      }
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
      for (int j = 0; j < 8; ++j) xT[(8 * cg + j) * PITCH + vox] = e[j];
    }
    __builtin_amdgcn_wave_barrier();
    if (bias_block && lane < COUT) {
      const uint16_t *row = &gT[lane * PITCH];
#pragma unroll 8
      for (int v = 0; v < 64; ++v) bsum += bf2f(row[v]);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int t = 0; t < TILES; ++t) {
        const int mt = t / NTI, nt = t - mt * NTI;
        const bf16x8_t A = __builtin_bit_cast(
            bf16x8_t, *reinterpret_cast<const uint4 *>(&gT[(16 * mt + r) * PITCH + 32 * s + 8 * g]));
        const bf16x8_t B = __builtin_bit_cast(
            bf16x8_t, *reinterpret_cast<const uint4 *>(&xT[(16 * nt + r) * PITCH + 32 * s + 8 * g]));
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
      }
    }
  }
  const size_t q = (size_t)p * 4 + wave;
  if (bias_block && lane < COUT) a.bpart[q * COUT + lane] = bsum;
#pragma unroll
  for (int t = 0; t < TILES; ++t) {
    const int mt = t / NTI, nt = t - mt * NTI;
    const int ci = 16 * nt + r;
    if (ci >= CIN) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = 16 * mt + 4 * g + j;
      a.part[(q * COUT + co) * (27 * CIN) + tap * CIN + ci] = acc[t][j];
    }
  }
}

// dst[i] = sum_p part[p * n + i], p ascending.
__global__ void k_reduce_rows(const float *part, int P, long long n, float *dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int p = 0; p < P; ++p) s += part[(size_t)p * n + i];
  dst[i] = s;
}

// ---------------------------------------------------------------------------
// 2x2x2 average pooling, nearest 2x up-sampling and their adjoints.  One
// thread per (voxel, group of 8 channels); 16-byte loads and stores.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unpack8(const uint4 &q, float f[8]) {
  const uint16_t *e = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
  for (int j = 0; j < 8; ++j) f[j] = bf2f(e[j]);
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
  uint4 q;
  uint16_t *e = reinterpret_cast<uint16_t *>(&q);
#pragma unroll
  for (int j = 0; j < 8; ++j) e[j] = f2bf(f[j]);
  return q;
}

// out (grid X,Y,Z) = mean of the 2x2x2 children in `in` (grid 2X,2Y,2Z)
__global__ void k_pool_fwd(const uint16_t *in, int ld_in, int off_in, uint16_t *out, int ld_out,
                           int off_out, int C, int B, int X, int Y, int Z) {
  const int G = C / 8;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)B * X * Y * Z * G;
  if (idx >= total) return;
  const int cg = (int)(idx % G);
  long long v = idx / G;
  const int z = (int)(v % Z);
  v /= Z;
  const int y = (int)(v % Y);
  v /= Y;
  const int x = (int)(v % X);
  const int b = (int)(v / X);
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const long long vin =
        (((long long)b * 2 * X + 2 * x + (i >> 2)) * 2 * Y + 2 * y + ((i >> 1) & 1)) * 2 * Z +
        2 * z + (i & 1);
    float f[8];
    unpack8(*reinterpret_cast<const uint4 *>(in + vin * ld_in + off_in + 8 * cg), f);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += f[j];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] *= 0.125f;
  const long long vo = idx / G;
  *reinterpret_cast<uint4 *>(out + vo * ld_out + off_out + 8 * cg) = pack8(s);
}

// din (grid 2X..) = (dout[parent] / 8 + add) * (mask > 0); add may be null
__global__ void k_pool_bwd(const uint16_t *dout, int ld_dout, const uint16_t *add, int ld_add,
                           int off_add, const uint16_t *mask, int ld_mask, int off_mask,
                           uint16_t *din, int ld_din, int C, int B, int X, int Y, int Z) {
  const int G = C / 8;
  const int X2 = 2 * X, Y2 = 2 * Y, Z2 = 2 * Z;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)B * X2 * Y2 * Z2 * G;
  if (idx >= total) return;
  const int cg = (int)(idx % G);
  const long long vin = idx / G;
  long long v = vin;
  const int z = (int)(v % Z2);
  v /= Z2;
  const int y = (int)(v % Y2);
  v /= Y2;
  const int x = (int)(v % X2);
  const int b = (int)(v / X2);
  const long long vp = (((long long)b * X + x / 2) * Y + y / 2) * Z + z / 2;
  float d[8];
  unpack8(*reinterpret_cast<const uint4 *>(dout + vp * ld_dout + 8 * cg), d);
  float ad[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (add) unpack8(*reinterpret_cast<const uint4 *>(add + vin * ld_add + off_add + 8 * cg), ad);
  float mk[8];
  unpack8(*reinterpret_cast<const uint4 *>(mask + vin * ld_mask + off_mask + 8 * cg), mk);
  float o[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = mk[j] > 0.f ? d[j] * 0.125f + ad[j] : 0.f;
  *reinterpret_cast<uint4 *>(din + vin * ld_din + 8 * cg) = pack8(o);
}

// out (grid 2X..) channel range [off_out, off_out + C) = in[parent]
__global__ void k_up_fwd(const uint16_t *in, int ld_in, uint16_t *out, int ld_out, int off_out,
                         int C, int B, int X, int Y, int Z) {
  const int G = C / 8;
  const int X2 = 2 * X, Y2 = 2 * Y, Z2 = 2 * Z;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)B * X2 * Y2 * Z2 * G;
  if (idx >= total) return;
  const int cg = (int)(idx % G);
  const long long vo = idx / G;
  long long v = vo;
  const int z = (int)(v % Z2);
  v /= Z2;
  const int y = (int)(v % Y2);
  v /= Y2;
  const int x = (int)(v % X2);
  const int b = (int)(v / X2);
  const long long vp = (((long long)b * X + x / 2) * Y + y / 2) * Z + z / 2;
  *reinterpret_cast<uint4 *>(out + vo * ld_out + off_out + 8 * cg) =
      *reinterpret_cast<const uint4 *>(in + vp * ld_in + 8 * cg);
}

// din (grid X..) = (sum of the 8 children of dout) * (mask > 0)
__global__ void k_up_bwd(const uint16_t *dout, int ld_dout, int off_dout, const uint16_t *mask,
                         uint16_t *din, int C, int B, int X, int Y, int Z) {
  const int G = C / 8;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)B * X * Y * Z * G;
  if (idx >= total) return;
  const int cg = (int)(idx % G);
  const long long vc = idx / G;
  long long v = vc;
  const int z = (int)(v % Z);
  v /= Z;
  const int y = (int)(v % Y);
  v /= Y;
  const int x = (int)(v % X);
  const int b = (int)(v / X);
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const long long vf =
        (((long long)b * 2 * X + 2 * x + (i >> 2)) * 2 * Y + 2 * y + ((i >> 1) & 1)) * 2 * Z +
        2 * z + (i & 1);
    float f[8];
    unpack8(*reinterpret_cast<const uint4 *>(dout + vf * ld_dout + off_dout + 8 * cg), f);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += f[j];
  }
  float mk[8];
  unpack8(*reinterpret_cast<const uint4 *>(mask + vc * C + 8 * cg), mk);
#pragma unroll
  for (int j = 0; j < 8; ++j) s[j] = mk[j] > 0.f ? s[j] : 0.f;
  *reinterpret_cast<uint4 *>(din + vc * C + 8 * cg) = pack8(s);
}

// Network input of `batch` samples: channels bin/(levels-1) of the fixed and
// the moving image, then zeros.  data: [n][2][nv] bins; idx: sample of each
// batch entry (null: entry b is sample b).
__global__ void k_build_input(const uint8_t *data, const int *idx, int batch, int nv,
                              float inv_l, float inv_k, uint16_t *x0) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)batch * nv) return;
  const int b = (int)(i / nv), v = (int)(i - (long long)b * nv);
  const long long smp = idx ? idx[b] : b;
  const uint8_t *d = data + smp * 2 * nv;
  float f[8] = {(float)d[v] * inv_l, (float)d[nv + v] * inv_k, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  *reinterpret_cast<uint4 *>(x0 + i * kInC) = pack8(f);
}

// ---------------------------------------------------------------------------
// Head: 1x1x1 conv to three displacement channels on the coarse grid, then
// the separable orthonormal DCT-II restricted to the 8 lowest frequencies per
// axis (tables c*[a][i], a < 8), scaled by 1/sqrt(coarse voxel count).
// ---------------------------------------------------------------------------
// U[b][n][v] = sum_c wh[n][c] s5[b][v][c] + bh[n]   (wh = [3][64], bh after)
__global__ void k_head_u(const uint16_t *s5, const float *wh, float *U, int B, int nc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)B * nc) return;
  const int b = (int)(i / nc), v = (int)(i - (long long)b * nc);
  float acc[3] = {wh[3 * kHeadC + 0], wh[3 * kHeadC + 1], wh[3 * kHeadC + 2]};
#pragma unroll
  for (int cg = 0; cg < kHeadC / 8; ++cg) {
    float f[8];
    unpack8(*reinterpret_cast<const uint4 *>(s5 + i * kHeadC + 8 * cg), f);
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int n = 0; n < 3; ++n) acc[n] += wh[n * kHeadC + 8 * cg + j] * f[j];
  }
#pragma unroll
  for (int n = 0; n < 3; ++n) U[((size_t)b * 3 + n) * nc + v] = acc[n];
}

// One block per (b, component): coef[b][n][a][bb][k] (a, bb, k < C), with
// the tables c*[a][x] zero for a at or above the grid side.
__global__ __launch_bounds__(256) void k_head_proj(const float *U, const float *cx,
                                                   const float *cy, const float *cz, int hx,
                                                   int hy, int hz, int C, float scale,
                                                   float *coef) {
  extern __shared__ float sm[];
  float *T1 = sm;                 // [hx][hy][C]
  float *T2 = sm + hx * hy * C;   // [hx][C][C]
  const float *u = U + (size_t)blockIdx.x * hx * hy * hz;
  const int t = threadIdx.x, C3 = C * C * C;
  for (int i = t; i < hx * hy * C; i += 256) {
    const int k = i % C, xy = i / C;
    float s = 0.f;
    for (int z = 0; z < hz; ++z) s += cz[k * hz + z] * u[(size_t)xy * hz + z];
    T1[i] = s;
  }
  __syncthreads();
  for (int i = t; i < hx * C * C; i += 256) {
    const int k = i % C, bb = (i / C) % C, x = i / (C * C);
    float s = 0.f;
    for (int y = 0; y < hy; ++y) s += cy[bb * hy + y] * T1[(x * hy + y) * C + k];
    T2[i] = s;
  }
  __syncthreads();
  for (int i = t; i < C3; i += 256) {
    const int k = i % C, bb = (i / C) % C, aa = i / (C * C);
    float s = 0.f;
    for (int x = 0; x < hx; ++x) s += cx[aa * hx + x] * T2[(x * C + bb) * C + k];
    coef[(size_t)blockIdx.x * C3 + i] = scale * s;
  }
}

// Adjoint of k_head_proj: dU[b][n] from dcoef[b][n].
__global__ __launch_bounds__(256) void k_head_expand(const float *dcoef, const float *cx,
                                                     const float *cy, const float *cz, int hx,
                                                     int hy, int hz, int C, float scale,
                                                     float *dU) {
  extern __shared__ float sm[];
  float *E1 = sm;                // [C][C][hz]   (a, bb, z)
  float *E2 = sm + C * C * hz;   // [C][hy][hz]  (a, y, z)
  const float *dc = dcoef + (size_t)blockIdx.x * C * C * C;
  const int t = threadIdx.x;
  for (int i = t; i < C * C * hz; i += 256) {
    const int z = i % hz, ab = i / hz;
    float s = 0.f;
    for (int k = 0; k < C; ++k) s += cz[k * hz + z] * dc[ab * C + k];
    E1[i] = s;
  }
  __syncthreads();
  for (int i = t; i < C * hy * hz; i += 256) {
    const int z = i % hz, y = (i / hz) % hy, aa = i / (hy * hz);
    float s = 0.f;
    for (int bb = 0; bb < C; ++bb) s += cy[bb * hy + y] * E1[(aa * C + bb) * hz + z];
    E2[i] = s;
  }
  __syncthreads();
  float *out = dU + (size_t)blockIdx.x * hx * hy * hz;
  for (int i = t; i < hx * hy * hz; i += 256) {
    const int yz = i % (hy * hz), x = i / (hy * hz);
    float s = 0.f;
    for (int aa = 0; aa < C; ++aa) s += cx[aa * hx + x] * E2[aa * hy * hz + yz];
    out[i] = scale * s;
  }
}

// ds5[b][v][c] = (sum_n wh[n][c] dU[b][n][v]) * (s5 > 0)
__global__ void k_head_bwd_data(const float *dU, const float *wh, const uint16_t *s5,
                                uint16_t *ds5, int B, int nc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)B * nc * (kHeadC / 8);
  if (i >= total) return;
  const int cg = (int)(i % (kHeadC / 8));
  const long long vb = i / (kHeadC / 8);
  const int b = (int)(vb / nc), v = (int)(vb - (long long)b * nc);
  float du[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) du[n] = dU[((size_t)b * 3 + n) * nc + v];
  float mk[8], o[8];
  unpack8(*reinterpret_cast<const uint4 *>(s5 + vb * kHeadC + 8 * cg), mk);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = 8 * cg + j;
    const float s = wh[c] * du[0] + wh[kHeadC + c] * du[1] + wh[2 * kHeadC + c] * du[2];
    o[j] = mk[j] > 0.f ? s : 0.f;
  }
  *reinterpret_cast<uint4 *>(ds5 + vb * kHeadC + 8 * cg) = pack8(o);
}

// part[p][n*64 + c] = sum_v dU[n][v] s5[v][c];  part[p][192 + n] = sum_v dU[n][v]
constexpr int kHeadParams = 3 * kHeadC + 3;
__global__ __launch_bounds__(256) void k_head_wgrad(const float *dU, const uint16_t *s5, int B,
                                                    int nc, int per_block, float *part) {
  const int t = threadIdx.x;
  if (t >= kHeadParams) return;
  const long long nvb = (long long)B * nc;
  const long long v_begin = (long long)blockIdx.x * per_block;
  const long long v_end = min(nvb, v_begin + per_block);
  const int n = t < 3 * kHeadC ? t / kHeadC : t - 3 * kHeadC;
  const int c = t < 3 * kHeadC ? t % kHeadC : -1;
  float s = 0.f;
  for (long long vb = v_begin; vb < v_end; ++vb) {
    const int b = (int)(vb / nc), v = (int)(vb - (long long)b * nc);
    const float du = dU[((size_t)b * 3 + n) * nc + v];
    s += c >= 0 ? du * bf2f(s5[vb * kHeadC + c]) : du;
  }
  part[(size_t)blockIdx.x * kHeadParams + t] = s;
}

// Mean squared error over batch x 3 C^3 coefficients (per = 3 C^3) and its
// gradient, in two steps: each block sums its fixed grid-stride share of
// the squared errors into lpart[block], then one block adds the partials in
// index order (the loss does not depend on scheduling).
constexpr int kMseBlocks = 64;
__global__ __launch_bounds__(256) void k_mse_part(const float *coef, const float *targets,
                                                  const int *idx, int B, int per,
                                                  float *dcoef, double *lpart) {
  __shared__ double red[256];
  const int n = B * per;
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int b = i / per, k = i - b * per;
    const float e = coef[i] - targets[(size_t)idx[b] * per + k];
    s += (double)e * e;
    dcoef[i] = 2.0f * e / (float)n;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

__global__ void k_mse_final(const double *lpart, int P, int n, double *loss) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int p = 0; p < P; ++p) s += lpart[p];
  *loss = s / n;
}

// Adam (Kingma & Ba 2015) with bias correction folded into c1, c2.
__global__ void k_adam(float *w, const float *g, float *m, float *v, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 compute copies of one conv layer: [cout][kpad] and the tap-mirrored
// transpose [cin][kpadT] (entries past 27*cin / 27*cout are zero).
__global__ void k_pack(const float *w, int cin, int cout, int kpad, int kpadT, uint16_t *wb,
                       uint16_t *wtb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n1 = (long long)cout * kpad, n2 = (long long)cin * kpadT;
  if (i < n1) {
    const int co = (int)(i / kpad), kk = (int)(i - (long long)co * kpad);
    wb[i] = f2bf(kk < 27 * cin ? w[(size_t)co * 27 * cin + kk] : 0.f);
  } else if (i < n1 + n2) {
    const long long j = i - n1;
    const int ci = (int)(j / kpadT), kk = (int)(j - (long long)ci * kpadT);
    float val = 0.f;
    if (kk < 27 * cout) {
      const int tp = kk / cout, co = kk - tp * cout;
      val = w[((size_t)co * 27 + (26 - tp)) * cin + ci];
    }
    wtb[j] = f2bf(val);
  }
}

// Low-frequency field of a pyramid level from normalised coefficients: one
// block per (i, j) row of the level grid.  The row's x and y basis values
// fold the C^3 coefficients of each component into C weights of the z basis
// (P[n][q] = sum_{a,b} Cx[a](i) Cy[b](j) coef[n][a][b][q]), so each voxel
// costs C basis values and 3 C products.
__global__ __launch_bounds__(64) void k_dct_field(const float *coef, int C, int fx, int fy,
                                                  int fz, int res, const float *x, float *d,
                                                  int nx, int ny, int nz, float amp) {
  __shared__ float bxy[2 * kMaxWarmCoef];
  __shared__ float P[3 * kMaxWarmCoef];
  const int i = (int)(blockIdx.x / ny), j = (int)(blockIdx.x % ny);
  const int t = threadIdx.x;
  if (t < 2 * C) {
    const int a = t % C;
    const int f = t < C ? fx : fy, I = res * (t < C ? i : j);
    bxy[t] = a ? sqrtf(2.f / f) * cospif((float)((2 * I + 1) * a) / (float)(2 * f))
               : sqrtf(1.f / f);
  }
  __syncthreads();
  if (t < 3 * C) {
    const int n = t / C, q = t % C;
    const float *cn = coef + (size_t)n * C * C * C + q;
    float s = 0.f;
    for (int a = 0; a < C; ++a) {
      float sa = 0.f;
      for (int b = 0; b < C; ++b) sa += bxy[C + b] * cn[(a * C + b) * C];
      s += bxy[a] * sa;
    }
    P[t] = s;
  }
  __syncthreads();
  const long long nv = (long long)nx * ny * nz;
  const float sz0 = sqrtf(1.f / fz), sz = sqrtf(2.f / fz);
  for (int k = t; k < nz; k += blockDim.x) {
    const long long v = (long long)blockIdx.x * nz + k;
    const int K = res * k;
    float u[3] = {0.f, 0.f, 0.f};
    for (int q = 0; q < C; ++q) {
      const float bz = q ? sz * cospif((float)((2 * K + 1) * q) / (float)(2 * fz)) : sz0;
      u[0] += P[q] * bz;
      u[1] += P[C + q] * bz;
      u[2] += P[2 * C + q] * bz;
    }
    const float grid[3] = {(float)i, (float)j, (float)k};
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const float pos = x ? x[n * nv + v] : grid[n];
      d[n * nv + v] = pos + amp * u[n] / (float)res;
    }
  }
}

int blocks_for(long long n, int per = 256) { return (int)((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// Dispatch of the templated kernels over the layer shapes of the network.
// ---------------------------------------------------------------------------
void launch_conv(int cin, int cout, const ConvArgs &a, hipStream_t s) {
  const dim3 grid(blocks_for(a.nvox, kConvVoxPerBlock)), block(256);
#define BINDER_CONV_CASE(CI, CO)                                           \
  if (cin == CI && cout == CO) {                                           \
    hipLaunchKernelGGL((k_conv<CI, CO / 16>), grid, block, 0, s, a);       \
    HIP_LAUNCH_OK();                                                       \
    return;                                                                \
  }
  BINDER_CONV_CASE(8, 16)
  BINDER_CONV_CASE(16, 16)
  BINDER_CONV_CASE(16, 32)
  BINDER_CONV_CASE(32, 32)
  BINDER_CONV_CASE(32, 48)
  BINDER_CONV_CASE(48, 48)
  BINDER_CONV_CASE(48, 64)
  BINDER_CONV_CASE(64, 64)
  BINDER_CONV_CASE(112, 64)
  BINDER_CONV_CASE(32, 16)
  BINDER_CONV_CASE(48, 32)
  BINDER_CONV_CASE(64, 48)
  BINDER_CONV_CASE(64, 112)
#undef BINDER_CONV_CASE
  throw Error(BINDER_EINVAL, "no convolution kernel for this channel pair");
}

// Narrow layers use the per-wavefront kernel, which writes 4 partial rows
// per block; returns the number of partial rows written.
bool wgrad_per_wave(int cin, int cout) { return cin <= 32 && cout <= 32; }

void launch_wgrad(int cin, int cout, const WgradArgs &a, int P, hipStream_t s) {
  const dim3 grid(P, 27), block(256);
  if (wgrad_per_wave(cin, cout)) {
#define BINDER_WGRAD_WAVE_CASE(CI, CO)                                     \
    if (cin == CI && cout == CO) {                                         \
      hipLaunchKernelGGL((k_wgrad_wave<CI, CO>), grid, block, 0, s, a);    \
      HIP_LAUNCH_OK();                                                     \
      return;                                                              \
    }
    BINDER_WGRAD_WAVE_CASE(8, 16)
    BINDER_WGRAD_WAVE_CASE(16, 16)
    BINDER_WGRAD_WAVE_CASE(16, 32)
    BINDER_WGRAD_WAVE_CASE(32, 32)
#undef BINDER_WGRAD_WAVE_CASE
  }
#define BINDER_WGRAD_CASE(CI, CO)                                          \
  if (cin == CI && cout == CO) {                                           \
    hipLaunchKernelGGL((k_wgrad<CI, CO>), grid, block, 0, s, a);           \
    HIP_LAUNCH_OK();                                                       \
    return;                                                                \
  }
  BINDER_WGRAD_CASE(8, 16)
  BINDER_WGRAD_CASE(16, 16)
  BINDER_WGRAD_CASE(16, 32)
  BINDER_WGRAD_CASE(32, 32)
  BINDER_WGRAD_CASE(32, 48)
  BINDER_WGRAD_CASE(48, 48)
  BINDER_WGRAD_CASE(48, 64)
  BINDER_WGRAD_CASE(64, 64)
  BINDER_WGRAD_CASE(112, 64)
#undef BINDER_WGRAD_CASE
  throw Error(BINDER_EINVAL, "no weight-gradient kernel for this channel pair");
}

}  // namespace

void launch_dct_field(const float *coef_dev, int ncoef, int fx, int fy, int fz, int res,
                      const float *x, float *d, int nx, int ny, int nz, hipStream_t s) {
  BINDER_REQUIRE(ncoef >= 1 && ncoef <= kMaxWarmCoef, "ncoef must be in [1, 16]");
  const float amp = (float)std::sqrt((double)fx * fy * fz);
  hipLaunchKernelGGL(k_dct_field, dim3((unsigned)((long long)nx * ny)), dim3(64), 0, s,
                     coef_dev, ncoef, fx, fy, fz, res, x, d, nx, ny, nz, amp);
  HIP_LAUNCH_OK();
}

}  // namespace binder

using namespace binder;

// ---------------------------------------------------------------------------
// The network handle.
// ---------------------------------------------------------------------------
namespace {

struct Tensor {
  uint16_t *p;
  int ld, off;
};

struct ConvLayer {
  int cin, cout, level, real_cin;
  long long woff, boff;  // offsets in the parameter vector
  int kpad, kpadT;
  DBuf<uint16_t> wb, wtb;
};

enum { A1, B1, A2, B2, A3, B3, A4, B4, C4, D4, A5, B5, NCONV };

}  // namespace

struct binder_warmstart {
  hipStream_t stream = nullptr;
  int dims[4][3];     // grid per level
  long long nv[4];    // voxels per sample per level
  int max_batch = 0;
  int ncoef = 8;      // C: coefficients per axis and component
  ConvLayer conv[NCONV];
  long long head_off = 0, nparam = 0;
  DBuf<float> w, grad, m, v;        // FP32 master parameters and Adam state
  long long adam_t = 0;
  // activations (BF16, channels-last)
  DBuf<uint16_t> x0, t1a, s1, p1, t2a, s2, p2, t3a, cat, p3, t4a, t4b, t4c, s4, t5a, s5;
  DBuf<uint16_t> dt1a, ds1, dp1, dt2a, ds2, dp2, dt3a, ds3, dcat, dp3, dt4a, dt4b, dt4c, ds4,
      dt5a, ds5;
  DBuf<float> U, dU, coef, dcoef, part, cpart, hpart;
  DBuf<float> cx, cy, cz;           // head DCT tables [C][h]
  DBuf<double> losses, lpart;
  // training set on the device
  DBuf<uint8_t> data, stage;
  DBuf<float> targets;
  DBuf<int> idx;
  int ndata = 0;
  float inv_l = 1.f, inv_k = 1.f;

  ~binder_warmstart() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }

  Tensor T(DBuf<uint16_t> &b, int ld, int off = 0) { return Tensor{b.get(), ld, off}; }
  int coefs_per_sample() const { return 3 * ncoef * ncoef * ncoef; }

  void conv_fwd(ConvLayer &L, Tensor in, Tensor out, int B) {
    const int *d = dims[L.level];
    ConvArgs a{};
    a.in = in.p;
    a.ld_in = in.ld;
    a.off_in = in.off;
    a.w = L.wb.get();
    a.kpad = L.kpad;
    a.bias = w.get() + L.boff;
    a.out = out.p;
    a.ld_out = out.ld;
    a.off_out = out.off;
    a.relu = 1;
    a.X = d[0];
    a.Y = d[1];
    a.Z = d[2];
    a.nvox = (int)(B * nv[L.level]);
    launch_conv(L.cin, L.cout, a, stream);
  }

  // out = input gradient of layer L from g (its pre-activation gradient),
  // times (mask > 0) when mask is given
  void conv_bwd_data(ConvLayer &L, Tensor g, Tensor out, const Tensor *mask, int B) {
    const int *d = dims[L.level];
    ConvArgs a{};
    a.in = g.p;
    a.ld_in = g.ld;
    a.off_in = g.off;
    a.w = L.wtb.get();
    a.kpad = L.kpadT;
    a.bias = nullptr;
    a.out = out.p;
    a.ld_out = out.ld;
    a.off_out = out.off;
    if (mask) {
      a.mask = mask->p;
      a.ld_mask = mask->ld;
      a.off_mask = mask->off;
    }
    a.relu = 0;
    a.X = d[0];
    a.Y = d[1];
    a.Z = d[2];
    a.nvox = (int)(B * nv[L.level]);
    launch_conv(L.cout, L.cin, a, stream);
  }

  int partitions(long long nvox) const {
    long long P = nvox / 4096;
    return (int)std::max(1ll, std::min(64ll, P));
  }
  // the head's weight gradient: one thread per weight, so many short blocks
  static constexpr int kHeadMaxParts = 1024;
  int head_partitions(long long nvb) const {
    return (int)std::max(1ll, std::min((long long)kHeadMaxParts, nvb / 64));
  }

  void conv_wgrad(ConvLayer &L, Tensor x, Tensor g, int B) {
    const int *d = dims[L.level];
    const long long nvox = B * nv[L.level];
    const int P = partitions(nvox);
    const int per = (int)(((nvox + P - 1) / P + 63) / 64 * 64);
    WgradArgs a{};
    a.x = x.p;
    a.ld_x = x.ld;
    a.off_x = x.off;
    a.g = g.p;
    a.ld_g = g.ld;
    a.off_g = g.off;
    a.part = part.get();
    a.bpart = cpart.get();
    a.X = d[0];
    a.Y = d[1];
    a.Z = d[2];
    a.nvox = (int)nvox;
    a.per_block = per;
    launch_wgrad(L.cin, L.cout, a, P, stream);
    const int rows = wgrad_per_wave(L.cin, L.cout) ? 4 * P : P;
    const long long nw = (long long)L.cout * 27 * L.cin;
    hipLaunchKernelGGL(k_reduce_rows, dim3(blocks_for(nw)), dim3(256), 0, stream, part.get(),
                       rows, nw, grad.get() + L.woff);
    HIP_LAUNCH_OK();
    // bias: the centre tap's blocks summed g per channel
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, stream, cpart.get(), rows,
                       (long long)L.cout, grad.get() + L.boff);
    HIP_LAUNCH_OK();
  }

  void pool_fwd(Tensor in, DBuf<uint16_t> &out, int C, int lvl_out, int B) {
    const int *d = dims[lvl_out];
    const long long n = B * nv[lvl_out] * (C / 8);
    hipLaunchKernelGGL(k_pool_fwd, dim3(blocks_for(n)), dim3(256), 0, stream, in.p, in.ld,
                       in.off, out.get(), C, 0, C, B, d[0], d[1], d[2]);
    HIP_LAUNCH_OK();
  }

  // din (level lvl_out - 1) = (pool adjoint of dout + add) * (mask > 0)
  void pool_bwd(DBuf<uint16_t> &dout, const Tensor *add, Tensor mask, DBuf<uint16_t> &din, int C,
                int lvl_out, int B) {
    const int *d = dims[lvl_out];
    const long long n = B * nv[lvl_out - 1] * (C / 8);
    hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(n)), dim3(256), 0, stream, dout.get(), C,
                       add ? add->p : nullptr, add ? add->ld : 0, add ? add->off : 0, mask.p,
                       mask.ld, mask.off, din.get(), C, C, B, d[0], d[1], d[2]);
    HIP_LAUNCH_OK();
  }

  void forward(int B) {
    const int n2 = (int)nv[2];
    conv_fwd(conv[A1], T(x0, kInC), T(t1a, 16), B);
    conv_fwd(conv[B1], T(t1a, 16), T(s1, 16), B);
    pool_fwd(T(s1, 16), p1, 16, 1, B);
    conv_fwd(conv[A2], T(p1, 16), T(t2a, 32), B);
    conv_fwd(conv[B2], T(t2a, 32), T(s2, 32), B);
    pool_fwd(T(s2, 32), p2, 32, 2, B);
    conv_fwd(conv[A3], T(p2, 32), T(t3a, 48), B);
    conv_fwd(conv[B3], T(t3a, 48), T(cat, 112, 64), B);
    pool_fwd(T(cat, 112, 64), p3, 48, 3, B);
    conv_fwd(conv[A4], T(p3, 48), T(t4a, 64), B);
    conv_fwd(conv[B4], T(t4a, 64), T(t4b, 64), B);
    conv_fwd(conv[C4], T(t4b, 64), T(t4c, 64), B);
    conv_fwd(conv[D4], T(t4c, 64), T(s4, 64), B);
    {
      const int *d = dims[3];
      const long long n = B * nv[2] * 8;
      hipLaunchKernelGGL(k_up_fwd, dim3(blocks_for(n)), dim3(256), 0, stream, s4.get(), 64,
                         cat.get(), 112, 0, 64, B, d[0], d[1], d[2]);
      HIP_LAUNCH_OK();
    }
    conv_fwd(conv[A5], T(cat, 112), T(t5a, 64), B);
    conv_fwd(conv[B5], T(t5a, 64), T(s5, 64), B);
    const float *wh = w.get() + head_off;
    hipLaunchKernelGGL(k_head_u, dim3(blocks_for((long long)B * n2)), dim3(256), 0, stream,
                       s5.get(), wh, U.get(), B, n2);
    HIP_LAUNCH_OK();
    const int *h = dims[2];
    const int C = ncoef;
    const size_t shm = (size_t)(h[0] * h[1] * C + h[0] * C * C) * sizeof(float);
    hipLaunchKernelGGL(k_head_proj, dim3(3 * B), dim3(256), shm, stream, U.get(), cx.get(),
                       cy.get(), cz.get(), h[0], h[1], h[2], C,
                       (float)(1.0 / std::sqrt((double)n2)), coef.get());
    HIP_LAUNCH_OK();
  }

  void backward(int B) {
    const int n2 = (int)nv[2];
    const int *h = dims[2];
    const float scale = (float)(1.0 / std::sqrt((double)n2));
    const int C = ncoef;
    const size_t shm = (size_t)(C * C * h[2] + C * h[1] * h[2]) * sizeof(float);
    hipLaunchKernelGGL(k_head_expand, dim3(3 * B), dim3(256), shm, stream, dcoef.get(),
                       cx.get(), cy.get(), cz.get(), h[0], h[1], h[2], C, scale, dU.get());
    HIP_LAUNCH_OK();
    const float *wh = w.get() + head_off;
    hipLaunchKernelGGL(k_head_bwd_data, dim3(blocks_for((long long)B * n2 * 8)), dim3(256), 0,
                       stream, dU.get(), wh, s5.get(), ds5.get(), B, n2);
    HIP_LAUNCH_OK();
    {
      const long long nvb = (long long)B * n2;
      const int P = head_partitions(nvb);
      const int per = (int)((nvb + P - 1) / P);
      hipLaunchKernelGGL(k_head_wgrad, dim3(P), dim3(256), 0, stream, dU.get(), s5.get(), B, n2,
                         per, hpart.get());
      HIP_LAUNCH_OK();
      hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, stream, hpart.get(), P,
                         (long long)kHeadParams, grad.get() + head_off);
      HIP_LAUNCH_OK();
    }
    // level 2 decoder
    conv_wgrad(conv[B5], T(t5a, 64), T(ds5, 64), B);
    {
      Tensor mk = T(t5a, 64);
      conv_bwd_data(conv[B5], T(ds5, 64), T(dt5a, 64), &mk, B);
    }
    conv_wgrad(conv[A5], T(cat, 112), T(dt5a, 64), B);
    conv_bwd_data(conv[A5], T(dt5a, 64), T(dcat, 112), nullptr, B);
    {
      const int *d = dims[3];
      const long long n = B * nv[3] * 8;
      hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(n)), dim3(256), 0, stream, dcat.get(), 112, 0,
                         s4.get(), ds4.get(), 64, B, d[0], d[1], d[2]);
      HIP_LAUNCH_OK();
    }
    // level 3
    conv_wgrad(conv[D4], T(t4c, 64), T(ds4, 64), B);
    {
      Tensor mk = T(t4c, 64);
      conv_bwd_data(conv[D4], T(ds4, 64), T(dt4c, 64), &mk, B);
    }
    conv_wgrad(conv[C4], T(t4b, 64), T(dt4c, 64), B);
    {
      Tensor mk = T(t4b, 64);
      conv_bwd_data(conv[C4], T(dt4c, 64), T(dt4b, 64), &mk, B);
    }
    conv_wgrad(conv[B4], T(t4a, 64), T(dt4b, 64), B);
    {
      Tensor mk = T(t4a, 64);
      conv_bwd_data(conv[B4], T(dt4b, 64), T(dt4a, 64), &mk, B);
    }
    conv_wgrad(conv[A4], T(p3, 48), T(dt4a, 64), B);
    conv_bwd_data(conv[A4], T(dt4a, 64), T(dp3, 48), nullptr, B);
    // level 2 encoder: S3 feeds the pool and the concatenation
    {
      Tensor add = T(dcat, 112, 64);
      pool_bwd(dp3, &add, T(cat, 112, 64), ds3, 48, 3, B);
    }
    conv_wgrad(conv[B3], T(t3a, 48), T(ds3, 48), B);
    {
      Tensor mk = T(t3a, 48);
      conv_bwd_data(conv[B3], T(ds3, 48), T(dt3a, 48), &mk, B);
    }
    conv_wgrad(conv[A3], T(p2, 32), T(dt3a, 48), B);
    conv_bwd_data(conv[A3], T(dt3a, 48), T(dp2, 32), nullptr, B);
    // level 1
    pool_bwd(dp2, nullptr, T(s2, 32), ds2, 32, 2, B);
    conv_wgrad(conv[B2], T(t2a, 32), T(ds2, 32), B);
    {
      Tensor mk = T(t2a, 32);
      conv_bwd_data(conv[B2], T(ds2, 32), T(dt2a, 32), &mk, B);
    }
    conv_wgrad(conv[A2], T(p1, 16), T(dt2a, 32), B);
    conv_bwd_data(conv[A2], T(dt2a, 32), T(dp1, 16), nullptr, B);
    // level 0
    pool_bwd(dp1, nullptr, T(s1, 16), ds1, 16, 1, B);
    conv_wgrad(conv[B1], T(t1a, 16), T(ds1, 16), B);
    {
      Tensor mk = T(t1a, 16);
      conv_bwd_data(conv[B1], T(ds1, 16), T(dt1a, 16), &mk, B);
    }
    conv_wgrad(conv[A1], T(x0, kInC), T(dt1a, 16), B);
  }

  void pack_all() {
    for (auto &L : conv) {
      const long long n = (long long)L.cout * L.kpad + (long long)L.cin * L.kpadT;
      hipLaunchKernelGGL(k_pack, dim3(blocks_for(n)), dim3(256), 0, stream, w.get() + L.woff,
                         L.cin, L.cout, L.kpad, L.kpadT, L.wb.get(), L.wtb.get());
      HIP_LAUNCH_OK();
    }
  }

  void build_input(const uint8_t *src, const int *didx, int B) {
    const long long n = (long long)B * nv[0];
    hipLaunchKernelGGL(k_build_input, dim3(blocks_for(n)), dim3(256), 0, stream, src, didx, B,
                       (int)nv[0], inv_l, inv_k, x0.get());
    HIP_LAUNCH_OK();
  }

  void mse(const int *didx, int B, double *loss_dev) {
    hipLaunchKernelGGL(k_mse_part, dim3(kMseBlocks), dim3(256), 0, stream, coef.get(),
                       targets.get(), didx, B, coefs_per_sample(), dcoef.get(), lpart.get());
    HIP_LAUNCH_OK();
    hipLaunchKernelGGL(k_mse_final, dim3(1), dim3(64), 0, stream, lpart.get(), kMseBlocks,
                       B * coefs_per_sample(), loss_dev);
    HIP_LAUNCH_OK();
  }

  void adam(float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++adam_t;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(nparam)), dim3(256), 0, stream, w.get(),
                       grad.get(), m.get(), v.get(), nparam, lr, b1, b2, eps, c1, c2);
    HIP_LAUNCH_OK();
    pack_all();
  }

  void require_data(const int *idx_host, int count) {
    BINDER_REQUIRE(ndata > 0, "set_dataset must be called first");
    for (int i = 0; i < count; ++i)
      BINDER_REQUIRE(idx_host[i] >= 0 && idx_host[i] < ndata, "sample index out of range");
  }
};

namespace {

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Orthonormal DCT-II rows a < C of length h, zero for a >= h.
void dct_table(int h, int C, std::vector<float> &t) {
  t.assign((size_t)C * h, 0.f);
  const double pi = 3.14159265358979323846;
  for (int a = 0; a < std::min(C, h); ++a)
    for (int x = 0; x < h; ++x)
      t[(size_t)a * h + x] =
          (float)(std::sqrt((a ? 2.0 : 1.0) / h) * std::cos(pi * (2 * x + 1) * a / (2.0 * h)));
}

}  // namespace

extern "C" {

int binder_warmstart_create(binder_warmstart **out, int nx, int ny, int nz, int max_batch,
                            int ncoef, unsigned long long seed) {
  return guarded([&] {
    BINDER_REQUIRE(out != nullptr, "out handle pointer is null");
    *out = nullptr;
    BINDER_REQUIRE(ncoef >= 1 && ncoef <= kMaxWarmCoef, "ncoef must be in [1, 16]");
    for (int d : {nx, ny, nz})
      BINDER_REQUIRE(d % 8 == 0 && d >= 32 && d <= 128,
                     "network grid sides must be multiples of 8 in [32, 128]");
    BINDER_REQUIRE(max_batch >= 1 && max_batch <= 64, "max_batch must be in [1, 64]");
    auto ws = std::make_unique<binder_warmstart>();
    ws->ncoef = ncoef;
    HIP_OK(hipStreamCreate(&ws->stream));
    for (int l = 0; l < 4; ++l) {
      ws->dims[l][0] = nx >> l;
      ws->dims[l][1] = ny >> l;
      ws->dims[l][2] = nz >> l;
      ws->nv[l] = (long long)ws->dims[l][0] * ws->dims[l][1] * ws->dims[l][2];
    }
    ws->max_batch = max_batch;
    const int spec[NCONV][3] = {{8, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 32, 1},
                                {32, 48, 2}, {48, 48, 2}, {48, 64, 3}, {64, 64, 3},
                                {64, 64, 3}, {64, 64, 3}, {112, 64, 2}, {64, 64, 2}};
    long long off = 0;
    for (int i = 0; i < NCONV; ++i) {
      ConvLayer &L = ws->conv[i];
      L.cin = spec[i][0];
      L.cout = spec[i][1];
      L.level = spec[i][2];
      L.real_cin = i == A1 ? 2 : L.cin;
      L.woff = off;
      off += (long long)L.cout * 27 * L.cin;
      L.boff = off;
      off += L.cout;
      L.kpad = round_up(27 * L.cin, 32);
      L.kpadT = round_up(27 * L.cout, 32);
      L.wb.alloc((size_t)L.cout * L.kpad);
      L.wtb.alloc((size_t)L.cin * L.kpadT);
    }
    ws->head_off = off;
    off += kHeadParams;
    ws->nparam = off;
    // He-normal initialisation of the convolutions (fan-in 27 x used input
    // channels), zero biases, a small head
    std::vector<float> w0((size_t)off, 0.f);
    std::mt19937_64 gen(seed);
    for (int i = 0; i < NCONV; ++i) {
      const ConvLayer &L = ws->conv[i];
      std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (27.0 * L.real_cin)));
      for (int co = 0; co < L.cout; ++co)
        for (int t = 0; t < 27; ++t)
          for (int ci = 0; ci < L.cin; ++ci)
            w0[(size_t)L.woff + ((size_t)co * 27 + t) * L.cin + ci] =
                ci < L.real_cin ? (float)nd(gen) : 0.f;
    }
    {
      std::normal_distribution<double> nd(0.0, 0.01 / std::sqrt((double)kHeadC));
      for (int i = 0; i < 3 * kHeadC; ++i) w0[(size_t)ws->head_off + i] = (float)nd(gen);
    }
    ws->w.upload(w0.data(), w0.size(), ws->stream);
    ws->grad.alloc(off);
    ws->grad.zero(ws->stream);
    ws->m.alloc(off);
    ws->m.zero(ws->stream);
    ws->v.alloc(off);
    ws->v.zero(ws->stream);
    const long long B = max_batch;
    auto act = [&](DBuf<uint16_t> &b, int lvl, int C) { b.alloc((size_t)(B * ws->nv[lvl] * C)); };
    act(ws->x0, 0, kInC);
    act(ws->t1a, 0, 16);
    act(ws->s1, 0, 16);
    act(ws->dt1a, 0, 16);
    act(ws->ds1, 0, 16);
    act(ws->p1, 1, 16);
    act(ws->dp1, 1, 16);
    act(ws->t2a, 1, 32);
    act(ws->s2, 1, 32);
    act(ws->dt2a, 1, 32);
    act(ws->ds2, 1, 32);
    act(ws->p2, 2, 32);
    act(ws->dp2, 2, 32);
    act(ws->t3a, 2, 48);
    act(ws->dt3a, 2, 48);
    act(ws->ds3, 2, 48);
    act(ws->cat, 2, 112);
    act(ws->dcat, 2, 112);
    act(ws->t5a, 2, 64);
    act(ws->dt5a, 2, 64);
    act(ws->s5, 2, 64);
    act(ws->ds5, 2, 64);
    act(ws->p3, 3, 48);
    act(ws->dp3, 3, 48);
    for (auto *b : {&ws->t4a, &ws->t4b, &ws->t4c, &ws->s4, &ws->dt4a, &ws->dt4b, &ws->dt4c,
                    &ws->ds4})
      act(*b, 3, 64);
    ws->U.alloc((size_t)(B * 3 * ws->nv[2]));
    ws->dU.alloc((size_t)(B * 3 * ws->nv[2]));
    ws->coef.alloc((size_t)B * ws->coefs_per_sample());
    ws->dcoef.alloc((size_t)B * ws->coefs_per_sample());
    long long maxpart = 0;
    for (auto &L : ws->conv)
      maxpart = std::max(maxpart, (long long)ws->partitions(B * ws->nv[L.level]) *
                                      (wgrad_per_wave(L.cin, L.cout) ? 4 : 1) * L.cout * 27 *
                                      L.cin);
    ws->part.alloc((size_t)maxpart);
    ws->cpart.alloc((size_t)4 * 64 * 64);
    ws->hpart.alloc((size_t)binder_warmstart::kHeadMaxParts * kHeadParams);
    std::vector<float> t;
    dct_table(ws->dims[2][0], ws->ncoef, t);
    ws->cx.upload(t.data(), t.size(), ws->stream);
    dct_table(ws->dims[2][1], ws->ncoef, t);
    ws->cy.upload(t.data(), t.size(), ws->stream);
    dct_table(ws->dims[2][2], ws->ncoef, t);
    ws->cz.upload(t.data(), t.size(), ws->stream);
    {
      // the head's projection and its adjoint keep two partial transforms in
      // LDS: above 64 KB (large grids with many coefficients) they need the
      // larger per-block limit
      const int *h = ws->dims[2];
      const int C = ws->ncoef;
      const size_t proj = (size_t)(h[0] * h[1] * C + h[0] * C * C) * sizeof(float);
      const size_t expd = (size_t)(C * C * h[2] + C * h[1] * h[2]) * sizeof(float);
      BINDER_REQUIRE(std::max(proj, expd) <= (size_t)160 * 1024,
                     "the head's transforms do not fit in LDS for this grid and ncoef");
      if (std::max(proj, expd) > (size_t)64 * 1024) {
        HIP_OK(hipFuncSetAttribute((const void *)k_head_proj,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        HIP_OK(hipFuncSetAttribute((const void *)k_head_expand,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      }
    }
    ws->losses.alloc(1);
    ws->lpart.alloc(kMseBlocks);
    ws->pack_all();
    HIP_OK(hipStreamSynchronize(ws->stream));
    *out = ws.release();
  });
}

void binder_warmstart_destroy(binder_warmstart *ws) { delete ws; }

int binder_warmstart_num_params(binder_warmstart *ws, long long *n) {
  return guarded([&] {
    BINDER_REQUIRE(ws && n, "null argument");
    *n = ws->nparam;
  });
}

int binder_warmstart_get_params(binder_warmstart *ws, float *out) {
  return guarded([&] {
    BINDER_REQUIRE(ws && out, "null argument");
    ws->w.download(out, (size_t)ws->nparam, ws->stream);
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

int binder_warmstart_set_params(binder_warmstart *ws, const float *in) {
  return guarded([&] {
    BINDER_REQUIRE(ws && in, "null argument");
    HIP_OK(hipMemcpyAsync(ws->w.get(), in, (size_t)ws->nparam * sizeof(float),
                          hipMemcpyHostToDevice, ws->stream));
    ws->m.zero(ws->stream);
    ws->v.zero(ws->stream);
    ws->adam_t = 0;
    ws->pack_all();
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

int binder_warmstart_set_dataset(binder_warmstart *ws, int n, const unsigned char *fixed_bins,
                                 const unsigned char *moving_bins, const float *targets, int L,
                                 int K) {
  return guarded([&] {
    BINDER_REQUIRE(ws && fixed_bins && moving_bins, "null argument");
    BINDER_REQUIRE(n >= 1, "empty dataset");
    BINDER_REQUIRE(L >= 2 && K >= 2 && L <= 256 && K <= 256, "bin counts must be in [2, 256]");
    const size_t nv = (size_t)ws->nv[0];
    ws->data.alloc((size_t)n * 2 * nv);
    for (int i = 0; i < n; ++i) {
      HIP_OK(hipMemcpyAsync(ws->data.get() + (size_t)i * 2 * nv, fixed_bins + (size_t)i * nv, nv,
                            hipMemcpyHostToDevice, ws->stream));
      HIP_OK(hipMemcpyAsync(ws->data.get() + (size_t)i * 2 * nv + nv,
                            moving_bins + (size_t)i * nv, nv, hipMemcpyHostToDevice, ws->stream));
    }
    if (targets)
      ws->targets.upload(targets, (size_t)n * ws->coefs_per_sample(), ws->stream);
    else
      ws->targets.release();
    ws->ndata = n;
    ws->inv_l = 1.f / (float)(L - 1);
    ws->inv_k = 1.f / (float)(K - 1);
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

int binder_warmstart_train(binder_warmstart *ws, const int *idx, int steps, int batch, double lr,
                           double *losses) {
  return guarded([&] {
    BINDER_REQUIRE(ws && idx, "null argument");
    BINDER_REQUIRE(batch >= 1 && batch <= ws->max_batch, "batch must be in [1, max_batch]");
    BINDER_REQUIRE(steps >= 1, "steps must be positive");
    ws->require_data(idx, steps * batch);
    BINDER_REQUIRE(ws->targets.get() != nullptr, "the dataset has no targets");
    ws->idx.upload(idx, (size_t)steps * batch, ws->stream);
    ws->losses.alloc((size_t)std::max(steps, 1));
    for (int s = 0; s < steps; ++s) {
      const int *di = ws->idx.get() + (size_t)s * batch;
      ws->build_input(ws->data.get(), di, batch);
      ws->forward(batch);
      ws->mse(di, batch, ws->losses.get() + s);
      ws->backward(batch);
      ws->adam((float)lr);
    }
    if (losses) ws->losses.download(losses, (size_t)steps, ws->stream);
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

int binder_warmstart_gradients(binder_warmstart *ws, const int *idx, int batch, double *loss,
                               float *grad_out) {
  return guarded([&] {
    BINDER_REQUIRE(ws && idx && grad_out, "null argument");
    BINDER_REQUIRE(batch >= 1 && batch <= ws->max_batch, "batch must be in [1, max_batch]");
    ws->require_data(idx, batch);
    BINDER_REQUIRE(ws->targets.get() != nullptr, "the dataset has no targets");
    ws->idx.upload(idx, (size_t)batch, ws->stream);
    ws->losses.alloc(1);
    ws->build_input(ws->data.get(), ws->idx.get(), batch);
    ws->forward(batch);
    ws->mse(ws->idx.get(), batch, ws->losses.get());
    ws->backward(batch);
    ws->grad.download(grad_out, (size_t)ws->nparam, ws->stream);
    if (loss) ws->losses.download(loss, 1, ws->stream);
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

int binder_warmstart_predict_dataset(binder_warmstart *ws, const int *idx, int batch,
                                     float *coef_out) {
  return guarded([&] {
    BINDER_REQUIRE(ws && idx && coef_out, "null argument");
    BINDER_REQUIRE(batch >= 1 && batch <= ws->max_batch, "batch must be in [1, max_batch]");
    ws->require_data(idx, batch);
    ws->idx.upload(idx, (size_t)batch, ws->stream);
    ws->build_input(ws->data.get(), ws->idx.get(), batch);
    ws->forward(batch);
    ws->coef.download(coef_out, (size_t)batch * ws->coefs_per_sample(), ws->stream);
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

int binder_warmstart_predict(binder_warmstart *ws, int batch, const unsigned char *fixed_bins,
                             const unsigned char *moving_bins, int L, int K, float *coef_out) {
  return guarded([&] {
    BINDER_REQUIRE(ws && fixed_bins && moving_bins && coef_out, "null argument");
    BINDER_REQUIRE(batch >= 1 && batch <= ws->max_batch, "batch must be in [1, max_batch]");
    BINDER_REQUIRE(L >= 2 && K >= 2 && L <= 256 && K <= 256, "bin counts must be in [2, 256]");
    const size_t nv = (size_t)ws->nv[0];
    ws->stage.alloc((size_t)ws->max_batch * 2 * nv);
    for (int b = 0; b < batch; ++b) {
      HIP_OK(hipMemcpyAsync(ws->stage.get() + (size_t)b * 2 * nv, fixed_bins + (size_t)b * nv, nv,
                            hipMemcpyHostToDevice, ws->stream));
      HIP_OK(hipMemcpyAsync(ws->stage.get() + (size_t)b * 2 * nv + nv,
                            moving_bins + (size_t)b * nv, nv, hipMemcpyHostToDevice,
                            ws->stream));
    }
    const float il = ws->inv_l, ik = ws->inv_k;
    ws->inv_l = 1.f / (float)(L - 1);
    ws->inv_k = 1.f / (float)(K - 1);
    ws->build_input(ws->stage.get(), nullptr, batch);
    ws->inv_l = il;
    ws->inv_k = ik;
    ws->forward(batch);
    ws->coef.download(coef_out, (size_t)batch * ws->coefs_per_sample(), ws->stream);
    HIP_OK(hipStreamSynchronize(ws->stream));
  });
}

}  // extern "C"
