// Per-axis DCT contractions for the deformation M-step / d-sample, on MFMA.
//
// Every pass contracts the contiguous axis of its input and writes its output
// transposed:  Out[k'][r] = sum_k In[r][k] M[k][k']  for r over all R rows.  A
// volume laid out [X][Y][Z] therefore rotates to [Z][X][Y], [Y][Z][X] and back
// to [X][Y][Z] after one pass per axis, so both MFMA operands are always
// contiguous in the contraction index and no separate transpose is needed.
//
// The products use split BF16: a = a_hi + a_lo with a_hi = bf16(a) and
// a_lo = bf16(a - a_hi) for the data and the (precomputed) matrix, and
// a*b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi accumulated in FP32 by
// v_mfma_f32_16x16x32_bf16 (about 2^-16 relative per product).
//
// The forward pass that completes the transform applies the prior filter
// p = c / (1 + gamma sigma^2 D) (+ sqrt(sigma^2/(1+gamma sigma^2 D)) * eps when
// sampling) and accumulates the prior cost 0.5 gamma D p^2 (and, for the nu
// draw, eta^2 rho D p^2) per block; the last inverse pass adds x back.
// Reference semantics: BINDER/src/cpu/NonLinearTransformation.cpp:58-126,
// :128-223, :344-364, :366-430, :433-451; ortho convention
// BINDER/src/cpu/DCT.cpp:83-175.
#include <atomic>
#include <cmath>
#include <cstring>

#include "autotune.hpp"
#include "smoothing.hpp"

namespace binder {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4_t;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4_t;

constexpr int kRowsPerBlock = 64;  // 4 wavefronts x 16 rows
constexpr int kMaxKSteps = 8;      // contraction length <= 256
constexpr int kSymMaxTiles = 6;    // even/odd split: half length 32..96

enum : int { EPI_STORE = 0, EPI_FILTER = 1, EPI_ADDX = 2, EPI_ADDGRID = 3 };
// prologue: 0 none, 1 subtract xpro, 2 subtract the generated index grid
enum : int { PRO_NONE = 0, PRO_X = 1, PRO_GRID = 2 };

struct PassArgs {
  const float *in;       // [comp][R][n_in]
  float *out;            // [comp][n_out][R]
  const float *xpro;     // prologue: in - xpro (same layout as in), may be null
  const float *xepi;     // ADDX epilogue: + xepi (same layout as out)
  // index grid in place of xpro / xepi (PRO_GRID, EPI_ADDGRID): component
  // `comp` of the voxel at (row, k) is k when comp == xg_axis, else
  // (row / xg_div[comp]) (mod xg_mod[comp] unless 0)
  int xg_axis;
  int xg_div[3], xg_mod[3];
  const uint32_t *bhi;   // swizzled matrix, bf16 hi / lo
  const uint32_t *blo;
  long long R;
  // output row blocks: row r = b * rin + i writes Out[b][k'][i]; rin == 0
  // writes Out[k'][r] (the whole row range rotates behind the new axis)
  long long rin;
  long long cin, cout;   // component strides of in / out
  int n_in, n_out, ks;   // ks = ceil(n_in / 32)
  // filter epilogue
  const float *D;
  const double *gamma;
  float sigma2;
  double etarho[3];
  int noise, raw;
  uint32_t k0, k1, sweep, stream;
  const uint32_t *keyx;
  double *cost_part, *raw_part;
};

// An optional sum of an earlier launch's partials that a contraction launch
// runs before its own work, in its first block's first wavefront, in place
// of a launch of its own: *out = sum of part[0 .. count), and with host,
// (stats[0], stats[1], sum) to page-locked host memory.  (A kernel argument
// of its own: the batched contractions' PassArgs stay as small as they are.)
struct SumJob {
  const double *part;
  long long count;
  double *out;
  const double *stats;
  double *host;
};

// Lane i adds partials i, i + 64, ... in index order, then a fixed
// butterfly; lane 0 writes.  Deterministic (the same assignment and order
// for every launch shape).
__device__ __forceinline__ void maybe_sum_job(const SumJob &j) {
  if (!j.out || blockIdx.x != 0 || blockIdx.y != 0 || threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  double s = 0.0;
  for (long long i = lane; i < j.count; i += 64) s += j.part[i];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0) {
    *j.out = s;
    if (j.host) {
      j.host[0] = j.stats[0];
      j.host[1] = j.stats[1];
      j.host[2] = s;
      __threadfence_system();
    }
  }
}


// Split 8 floats into bf16 hi and lo vectors: hi = bf16(v), lo = bf16(v - hi),
// both rounded to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ void split8(const float v[8], bf16x8_t &hi,
                                       bf16x8_t &lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 h = (__bf16)v[j];
    hi[j] = h;
    lo[j] = (__bf16)(v[j] - (float)h);
  }
}

// Index-grid coordinate of component `comp` at `row` (comp != xg_axis).
__device__ __forceinline__ float grid_coord(const PassArgs &a, int comp, long long row) {
  uint32_t v = (uint32_t)row / (uint32_t)a.xg_div[comp];
  if (a.xg_mod[comp]) v %= (uint32_t)a.xg_mod[comp];
  return (float)v;
}

// Offset of output element (k' = kc, row) in one component (see rin).
__device__ __forceinline__ long long out_offset(const PassArgs &a, int kc, long long row) {
  if (a.rin == 0) return (long long)kc * a.R + row;
  const uint32_t b = (uint32_t)row / (uint32_t)a.rin;
  const uint32_t i = (uint32_t)row - b * (uint32_t)a.rin;
  return ((long long)b * a.n_out + kc) * a.rin + i;
}

// Four standard normals from one Philox block (Box-Muller, FP32).
__device__ __forceinline__ void normals4(u32x4 r, float n[4]) {
  const float u1 = ((float)(r.x >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = (float)(r.y >> 8) * (1.0f / 16777216.0f);
  const float u3 = ((float)(r.z >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float u4 = (float)(r.w >> 8) * (1.0f / 16777216.0f);
  const float a = sqrtf(-2.0f * __logf(u1)), b = sqrtf(-2.0f * __logf(u3));
  float s1, c1, s2, c2;
  __sincosf(6.283185307f * u2, &s1, &c1);
  __sincosf(6.283185307f * u4, &s2, &c2);
  n[0] = a * c1;
  n[1] = a * s1;
  n[2] = b * c2;
  n[3] = b * s2;
}

// KS = number of 32-wide k-steps (compile time, so every fragment loop is
// straight-line code and the B loads of a whole output tile issue together).
// Block (bx, by) of a grid of gx x comps blocks of the non-persistent
// contraction: rows bx * 64 .. + 63 of component by.  Its outputs and, for
// the filter epilogue, its cost partial (slot by * gx + bx) depend only on
// (a, bx, by, gx).
template <int KS, bool VEC, int PRO, int EPI>
__device__ __forceinline__ void contract_block(const PassArgs &a, int bx, int by, int gx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int comp = by;
  const long long row0 = (long long)bx * kRowsPerBlock + wave * 16;
  const float *in = a.in + comp * a.cin;
  const float *xp = PRO == PRO_X ? a.xpro + comp * a.cin : nullptr;
  float *out = a.out + comp * a.cout;

  // A fragments: this lane's row and 8 consecutive k per 32-wide k-step
  const long long arow = row0 + (lane & 15);
  const bool rowok = arow < a.R;
  const int kq = 8 * (lane >> 4);
  bf16x8_t ahi[KS], alo[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    float v[8];
    const int k0 = s * 32 + kq;
    if (VEC && rowok && k0 + 8 <= a.n_in) {
      const float4 *p = reinterpret_cast<const float4 *>(in + arow * a.n_in + k0);
      const float4 x0 = p[0], x1 = p[1];
      v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
      v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
      if (PRO == PRO_X) {
        const float4 *q = reinterpret_cast<const float4 *>(xp + arow * a.n_in + k0);
        const float4 y0 = q[0], y1 = q[1];
        v[0] -= y0.x; v[1] -= y0.y; v[2] -= y0.z; v[3] -= y0.w;
        v[4] -= y1.x; v[5] -= y1.y; v[6] -= y1.z; v[7] -= y1.w;
      } else if (PRO == PRO_GRID) {
        const bool along = comp == a.xg_axis;
        const float rc = along ? 0.f : grid_coord(a, comp, arow);
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] -= along ? (float)(k0 + j) : rc;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + j;
        float t = 0.f;
        if (rowok && k < a.n_in) {
          t = in[arow * a.n_in + k];
          if (PRO == PRO_X) t -= xp[arow * a.n_in + k];
          if (PRO == PRO_GRID)
            t -= comp == a.xg_axis ? (float)k : grid_coord(a, comp, arow);
        }
        v[j] = t;
      }
    }
    split8(v, ahi[s], alo[s]);
  }

  double g = 0.0;
  float gs2 = 0.f;
  if (EPI == EPI_FILTER) {
    g = a.gamma[comp];
    gs2 = (float)g * a.sigma2;
  }
  double cost = 0.0, raw = 0.0;
  const int nt_count = (a.n_out + 15) >> 4;
  const long long orow = row0 + 4 * (lane >> 4);  // first of this lane's 4 rows
  float grc[4] = {0.f, 0.f, 0.f, 0.f};  // EPI_ADDGRID: this lane's rows' coordinate
  if (EPI == EPI_ADDGRID && comp != a.xg_axis) {
#pragma unroll
    for (int i = 0; i < 4; ++i) grc[i] = grid_coord(a, comp, orow + i);
  }
  const uint4 *bh = reinterpret_cast<const uint4 *>(a.bhi) + lane;
  const uint4 *bl = reinterpret_cast<const uint4 *>(a.blo) + lane;
  // B fragments of the current output tile; the next tile's are prefetched
  // while the current tile's MFMAs run
  uint4 ch[KS], cl[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    ch[s] = bh[s * 64];
    cl[s] = bl[s * 64];
  }
  for (int nt = 0; nt < nt_count; ++nt) {
    uint4 nh[KS], nl[KS];
    const int nn = (nt + 1 < nt_count) ? nt + 1 : nt;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      nh[s] = bh[((long long)nn * KS + s) * 64];
      nl[s] = bl[((long long)nn * KS + s) * 64];
    }
    // epilogue operands of this tile (D or x), loaded before the MFMAs
    const int kc = nt * 16 + (lane & 15);
    const long long obase = out_offset(a, kc, orow);
    const bool colok = kc < a.n_out;
    float4 epv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (EPI == EPI_FILTER || (EPI == EPI_ADDX && a.xepi)) {
      const float *src = EPI == EPI_FILTER ? a.D : a.xepi + comp * a.cout;
      if (VEC && colok && orow + 3 < a.R) {
        epv = *reinterpret_cast<const float4 *>(src + obase);
      } else if (colok) {
        float t4[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (orow + i < a.R) t4[i] = src[out_offset(a, kc, orow + i)];
        epv = make_float4(t4[0], t4[1], t4[2], t4[3]);
      }
    }
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, ch[s]);
      const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, cl[s]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[s], bhi, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[s], blo, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[s], bhi, acc, 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      ch[s] = nh[s];
      cl[s] = nl[s];
    }
    // lane holds rows orow..orow+3 of output column kc
    if (!colok) continue;
    float v[4] = {acc[0], acc[1], acc[2], acc[3]};
    const float ep[4] = {epv.x, epv.y, epv.z, epv.w};
    if (EPI == EPI_FILTER) {
      float eps[4] = {0.f, 0.f, 0.f, 0.f};
      if (a.noise) {
        // one block of four normals per (output column, 4-row group)
        const u32x4 r = philox4x32_10(
            u32x4{(uint32_t)orow, (uint32_t)kc, a.sweep,
                  ((uint32_t)comp << 16) | a.stream},
            a.keyx ? a.k0 ^ a.keyx[0] : a.k0, a.keyx ? a.k1 ^ a.keyx[1] : a.k1);
        normals4(r, eps);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (orow + i >= a.R) break;
        const float Dv = ep[i];
        const float den = 1.0f + gs2 * Dv;
        float p = v[i] * (1.0f / den);
        if (a.noise) p += eps[i] * sqrtf(a.sigma2 / den);
        const double pd = (double)p;
        cost += 0.5 * g * (double)Dv * pd * pd;
        if (a.raw) raw += a.etarho[comp] * (double)Dv * pd * pd;
        v[i] = p;
      }
    } else if (EPI == EPI_ADDX) {
      if (a.xepi) {
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] += ep[i];
      }
    } else if (EPI == EPI_ADDGRID) {
      const bool along = comp == a.xg_axis;
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] += along ? (float)kc : grc[i];
    }
    if (VEC && orow + 3 < a.R) {
      *reinterpret_cast<float4 *>(out + obase) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (orow + i < a.R) out[out_offset(a, kc, orow + i)] = v[i];
    }
  }

  if (EPI == EPI_FILTER) {
    __shared__ double red[2][256];
    const int t = threadIdx.x;
    red[0][t] = cost;
    red[1][t] = raw;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        red[0][t] += red[0][t + s];
        red[1][t] += red[1][t + s];
      }
      __syncthreads();
    }
    if (t == 0) {
      const long long b = (long long)by * gx + bx;
      a.cost_part[b] = red[0][0];
      if (a.raw) a.raw_part[b] = red[1][0];
    }
  }
}

template <int KS, bool VEC, int PRO, int EPI>
__global__ __launch_bounds__(256) void k_contract(PassArgs a, SumJob j) {
  maybe_sum_job(j);
  contract_block<KS, VEC, PRO, EPI>(a, blockIdx.x, blockIdx.y, gridDim.x);
}

// Pair of block `b` in a batched launch: prefix[p] <= b < prefix[p + 1]
// (prefix[0] = 0, prefix[npairs] = total blocks).
__device__ __forceinline__ int batch_pair(const int *prefix, int npairs, int b) {
  int lo = 0, hi = npairs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The non-persistent contraction of several problems in one launch: pair p
// owns blocks prefix[p] .. prefix[p + 1] - 1, which it runs as its own grid
// of ceil(R / 64) x comps blocks.
template <int KS, bool VEC, int PRO, int EPI>
__global__ __launch_bounds__(256) void k_contract_batch(const PassArgs *args, const int *prefix,
                                                        int npairs) {
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int p = batch_pair(prefix, npairs, b);
  const PassArgs a = args[p];
  const int lb = b - prefix[p];
  const int gx = (int)((a.R + kRowsPerBlock - 1) / kRowsPerBlock);
  contract_block<KS, VEC, PRO, EPI>(a, lb % gx, lb / gx, gx);
}

// Persistent variant for contraction lengths up to 192: one 1024-thread block
// per CU holds the whole pass matrix (bf16 hi/lo, MFMA order) in LDS and its
// 16 wavefronts stream 16-row tiles; the many wavefronts per CU hide the HBM
// latency of the tile loads.
constexpr int kPersistThreads = 1024;
constexpr size_t kPersistMaxLds = 144 * 1024;

template <int KS, bool VEC, int PRO, int EPI>
__global__ __launch_bounds__(kPersistThreads) void k_contract_lds(PassArgs a, int comps,
                                                                 SumJob j) {
  maybe_sum_job(j);
  extern __shared__ __align__(16) uint4 bsh[];  // [nt][KS][64] hi, then lo
  const int nt_count = (a.n_out + 15) >> 4;
  const int nfrag = nt_count * KS * 64;
  {
    const uint4 *gh = reinterpret_cast<const uint4 *>(a.bhi);
    const uint4 *gl = reinterpret_cast<const uint4 *>(a.blo);
    for (int i = threadIdx.x; i < nfrag; i += blockDim.x) {
      bsh[i] = gh[i];
      bsh[nfrag + i] = gl[i];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int waves_per_block = blockDim.x >> 6;
  const long long tiles_per_comp = (a.R + 15) >> 4;
  const long long total = tiles_per_comp * comps;
  const int kq = 8 * (lane >> 4);
  double cost = 0.0, raw = 0.0;
  for (long long tile = (long long)blockIdx.x * waves_per_block + wave; tile < total;
       tile += (long long)gridDim.x * waves_per_block) {
    const int comp = (int)(tile / tiles_per_comp);
    const long long row0 = (tile % tiles_per_comp) * 16;
    const float *in = a.in + comp * a.cin;
    const float *xp = PRO == PRO_X ? a.xpro + comp * a.cin : nullptr;
    float *out = a.out + comp * a.cout;
    const long long arow = row0 + (lane & 15);
    const bool rowok = arow < a.R;
    bf16x8_t ahi[KS], alo[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      float v[8];
      const int k0 = s * 32 + kq;
      if (VEC && rowok && k0 + 8 <= a.n_in) {
        const float4 *p = reinterpret_cast<const float4 *>(in + arow * a.n_in + k0);
        const float4 x0 = p[0], x1 = p[1];
        v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
        v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
        if (PRO == PRO_X) {
          const float4 *q = reinterpret_cast<const float4 *>(xp + arow * a.n_in + k0);
          const float4 y0 = q[0], y1 = q[1];
          v[0] -= y0.x; v[1] -= y0.y; v[2] -= y0.z; v[3] -= y0.w;
          v[4] -= y1.x; v[5] -= y1.y; v[6] -= y1.z; v[7] -= y1.w;
        } else if (PRO == PRO_GRID) {
          const bool along = comp == a.xg_axis;
          const float rc = along ? 0.f : grid_coord(a, comp, arow);
#pragma unroll
          for (int j = 0; j < 8; ++j) v[j] -= along ? (float)(k0 + j) : rc;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = k0 + j;
          float t = 0.f;
          if (rowok && k < a.n_in) {
            t = in[arow * a.n_in + k];
            if (PRO == PRO_X) t -= xp[arow * a.n_in + k];
            if (PRO == PRO_GRID)
              t -= comp == a.xg_axis ? (float)k : grid_coord(a, comp, arow);
          }
          v[j] = t;
        }
      }
      split8(v, ahi[s], alo[s]);
    }
    double g = 0.0;
    float gs2 = 0.f;
    if (EPI == EPI_FILTER) {
      g = a.gamma[comp];
      gs2 = (float)g * a.sigma2;
    }
    const long long orow = row0 + 4 * (lane >> 4);
    float grc[4] = {0.f, 0.f, 0.f, 0.f};  // EPI_ADDGRID: this lane's rows' coordinate
    if (EPI == EPI_ADDGRID && comp != a.xg_axis) {
#pragma unroll
      for (int i = 0; i < 4; ++i) grc[i] = grid_coord(a, comp, orow + i);
    }
    for (int nt = 0; nt < nt_count; ++nt) {
      const int kc = nt * 16 + (lane & 15);
      const long long obase = out_offset(a, kc, orow);
      const bool colok = kc < a.n_out;
      float4 epv = make_float4(0.f, 0.f, 0.f, 0.f);
      if (EPI == EPI_FILTER || (EPI == EPI_ADDX && a.xepi)) {
        const float *src = EPI == EPI_FILTER ? a.D : a.xepi + comp * a.cout;
        if (VEC && colok && orow + 3 < a.R) {
          epv = *reinterpret_cast<const float4 *>(src + obase);
        } else if (colok) {
          float t4[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (orow + i < a.R) t4[i] = src[out_offset(a, kc, orow + i)];
          epv = make_float4(t4[0], t4[1], t4[2], t4[3]);
        }
      }
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      const uint4 *bt = bsh + (nt * KS) * 64 + lane;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, bt[s * 64]);
        const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, bt[nfrag + s * 64]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[s], bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[s], blo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[s], bhi, acc, 0, 0, 0);
      }
      if (!colok) continue;
      float v[4] = {acc[0], acc[1], acc[2], acc[3]};
      const float ep[4] = {epv.x, epv.y, epv.z, epv.w};
      if (EPI == EPI_FILTER) {
        float eps[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.noise) {
          const u32x4 r = philox4x32_10(
              u32x4{(uint32_t)orow, (uint32_t)kc, a.sweep,
                    ((uint32_t)comp << 16) | a.stream},
              a.keyx ? a.k0 ^ a.keyx[0] : a.k0, a.keyx ? a.k1 ^ a.keyx[1] : a.k1);
          normals4(r, eps);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (orow + i >= a.R) break;
          const float Dv = ep[i];
          const float den = 1.0f + gs2 * Dv;
          float p = v[i] * (1.0f / den);
          if (a.noise) p += eps[i] * sqrtf(a.sigma2 / den);
          const double pd = (double)p;
          cost += 0.5 * g * (double)Dv * pd * pd;
          if (a.raw) raw += a.etarho[comp] * (double)Dv * pd * pd;
          v[i] = p;
        }
      } else if (EPI == EPI_ADDX) {
        if (a.xepi) {
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] += ep[i];
        }
      } else if (EPI == EPI_ADDGRID) {
        const bool along = comp == a.xg_axis;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] += along ? (float)kc : grc[i];
      }
      if (VEC && orow + 3 < a.R) {
        *reinterpret_cast<float4 *>(out + obase) = make_float4(v[0], v[1], v[2], v[3]);
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (orow + i < a.R) out[out_offset(a, kc, orow + i)] = v[i];
      }
    }
  }
  if (EPI == EPI_FILTER) {
    __syncthreads();
    double *red = reinterpret_cast<double *>(bsh);  // matrix no longer needed
    const int t = threadIdx.x;
    red[t] = cost;
    red[kPersistThreads + t] = raw;
    __syncthreads();
    for (int st = kPersistThreads / 2; st > 0; st >>= 1) {
      if (t < st) {
        red[t] += red[t + st];
        red[kPersistThreads + t] += red[kPersistThreads + t + st];
      }
      __syncthreads();
    }
    if (t == 0) {
      a.cost_part[blockIdx.x] = red[0];
      if (a.raw) a.raw_part[blockIdx.x] = red[kPersistThreads];
    }
  }
}

// Persistent contraction of a square orthonormal DCT-II (INV = false) or
// DCT-III (INV = true) of even length N = 2H on its even / odd halves.  The
// matrix satisfies C[k][N-1-x] = (-1)^k C[k][x], so
//   forward: out[2m]   = sum_{x<H} C[2m][x]   (in[x] + in[N-1-x]),
//            out[2m+1] = sum_{x<H} C[2m+1][x] (in[x] - in[N-1-x]);
//   inverse: E[x] = sum_m C[2m][x] in[2m], O[x] = sum_m C[2m+1][x] in[2m+1]
//            (x < H), out[x] = E + O, out[N-1-x] = E - O.
// Two H x H matrices replace one N x N: half the MFMA work and half the LDS
// reads.  Tiles, rows and the output layout (rin) are those of
// k_contract_lds; a lane's two output columns are 2m, 2m + 1 (forward) or
// x, N - 1 - x (inverse).  NT = ceil(H / 16) output tiles per parity stay in
// registers while the k-steps stream through: the next k-step's input is
// loaded before the current one's MFMAs, so the HBM loads overlap the
// matrix work.  The forward takes the prologues and a plain store, the
// inverse a plain store or the x / index-grid epilogue.

// Raw A data of k-step s (j < 8): forward the sums (p) and differences (q)
// of x = x0 + j and N - 1 - x (prologue applied), inverse the coefficients
// 2m (p) and 2m + 1 (q) of m = x0 + j.  H % 8 == 0 and N % 16 == 0, so a
// group of 8 lies wholly inside [0, H) or wholly past it (zero), and rows
// are 16-byte aligned.
template <bool INV, int PRO>
__device__ __forceinline__ void sym_load(const PassArgs &a, const float *src, const float *xr,
                                         bool rowok, int x0, int N, int H, int comp, float rc,
                                         float p[8], float q[8]) {
  // loads into vectors that stay zero when the group lies past H (no early
  // return: it would make the compiler keep p and q in scratch memory)
  const bool ok = rowok && x0 < H;
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!INV) {
    // v[j] = in[x0 + j], w[j] = in[N - 1 - x0 - j]
    float4 v0 = z4, v1 = z4, w0 = z4, w1 = z4;
    if (ok) {
      const float4 *pv = reinterpret_cast<const float4 *>(src + x0);
      const float4 *pw = reinterpret_cast<const float4 *>(src + (N - 8 - x0));
      v0 = pv[0];
      v1 = pv[1];
      w0 = pw[0];
      w1 = pw[1];
      if (PRO == PRO_X) {
        const float4 *qv = reinterpret_cast<const float4 *>(xr + x0);
        const float4 *qw = reinterpret_cast<const float4 *>(xr + (N - 8 - x0));
        const float4 a0 = qv[0], a1 = qv[1], b0 = qw[0], b1 = qw[1];
        v0.x -= a0.x; v0.y -= a0.y; v0.z -= a0.z; v0.w -= a0.w;
        v1.x -= a1.x; v1.y -= a1.y; v1.z -= a1.z; v1.w -= a1.w;
        w0.x -= b0.x; w0.y -= b0.y; w0.z -= b0.z; w0.w -= b0.w;
        w1.x -= b1.x; w1.y -= b1.y; w1.z -= b1.z; w1.w -= b1.w;
      }
    }
    float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
    float w[8] = {w1.w, w1.z, w1.y, w1.x, w0.w, w0.z, w0.y, w0.x};
    if (PRO == PRO_GRID && ok) {
      const bool along = comp == a.xg_axis;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[j] -= along ? (float)(x0 + j) : rc;
        w[j] -= along ? (float)(N - 1 - x0 - j) : rc;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      p[j] = v[j] + w[j];
      q[j] = v[j] - w[j];
    }
  } else {
    float4 c0 = z4, c1 = z4, c2 = z4, c3 = z4;
    if (ok) {
      const float4 *pc = reinterpret_cast<const float4 *>(src + 2 * x0);
      c0 = pc[0];
      c1 = pc[1];
      c2 = pc[2];
      c3 = pc[3];
    }
    p[0] = c0.x; q[0] = c0.y; p[1] = c0.z; q[1] = c0.w;
    p[2] = c1.x; q[2] = c1.y; p[3] = c1.z; q[3] = c1.w;
    p[4] = c2.x; q[4] = c2.y; p[5] = c2.z; q[5] = c2.w;
    p[6] = c3.x; q[6] = c3.y; p[7] = c3.z; q[7] = c3.w;
  }
}

// Two 512-thread blocks per CU (each holding its own copy of the [E | O]
// matrices, at most 74 KB): the same 16 wavefronts per CU as one 1024-thread
// block, but the two blocks' matrix loads and tile streams are independent
// (measured 10-18 % faster per pass than one 1024-thread block).
constexpr int kSymThreads = 512;

template <int NT, bool INV, int PRO, int EPI>
__global__ __launch_bounds__(kSymThreads) void k_contract_sym(PassArgs a, int comps,
                                                             SumJob j) {
  maybe_sum_job(j);
  constexpr int KS = (NT + 1) / 2;
  extern __shared__ __align__(16) uint4 bsh[];  // [E | O] hi, then [E | O] lo
  const int N = a.n_in, H = N >> 1;
  constexpr int nfrag = NT * KS * 64;  // uint4 per matrix
  {
    const uint4 *gh = reinterpret_cast<const uint4 *>(a.bhi);
    const uint4 *gl = reinterpret_cast<const uint4 *>(a.blo);
    for (int i = threadIdx.x; i < 2 * nfrag; i += blockDim.x) {
      bsh[i] = gh[i];
      bsh[2 * nfrag + i] = gl[i];
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const int waves_per_block = blockDim.x >> 6;
  const long long tiles_per_comp = (a.R + 15) >> 4;
  const long long total = tiles_per_comp * comps;
  const long long stride = (long long)gridDim.x * waves_per_block;
  long long tile = (long long)blockIdx.x * waves_per_block + wave;
  // input of a tile: its component, first row and this lane's row
  struct TileIn {
    int comp;
    long long row0;
    bool rowok;
    const float *src, *xr;
    float rc;
  };
  auto tile_in = [&](long long t, int lane) {
    TileIn ti;
    ti.comp = (int)(t / tiles_per_comp);
    ti.row0 = (t % tiles_per_comp) * 16;
    const long long arow = ti.row0 + (lane & 15);
    ti.rowok = arow < a.R;
    ti.src = a.in + ti.comp * a.cin + arow * N;
    ti.xr = PRO == PRO_X ? a.xpro + ti.comp * a.cin + arow * N : nullptr;
    ti.rc = 0.f;
    if (PRO == PRO_GRID && ti.comp != a.xg_axis) ti.rc = grid_coord(a, ti.comp, arow);
    return ti;
  };
  // raw data of the next k-step to split: loaded one k-step ahead, across
  // tile boundaries (the last k-step of a tile loads the first of the next),
  // so the HBM loads stream while the MFMAs run
  float p[8], q[8];
  TileIn cur{};
  if (tile < total) {
    int lane = threadIdx.x & 63;
    cur = tile_in(tile, lane);
    sym_load<INV, PRO>(a, cur.src, cur.xr, cur.rowok, 8 * (lane >> 4), N, H, cur.comp, cur.rc,
                       p, q);
  }
  for (; tile < total; tile += stride) {
    // the lane index is re-read per tile: values derived from it alone (the
    // index-grid coordinates of every k-step, the output columns) are then
    // recomputed per tile instead of held in registers across the loop
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int kq = 8 * (lane >> 4);
    const long long next = tile + stride;
    TileIn nxt{};
    f32x4_t ae[NT], ao[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      ae[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      ao[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      bf16x8_t phi, plo, qhi, qlo;
      split8(p, phi, plo);
      split8(q, qhi, qlo);
      if (s + 1 < KS) {
        sym_load<INV, PRO>(a, cur.src, cur.xr, cur.rowok, (s + 1) * 32 + kq, N, H, cur.comp,
                           cur.rc, p, q);
      } else if (next < total) {
        nxt = tile_in(next, lane);
        sym_load<INV, PRO>(a, nxt.src, nxt.xr, nxt.rowok, kq, N, H, nxt.comp, nxt.rc, p, q);
      }
      const uint4 *bt = bsh + s * 64 + lane;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint4 *b = bt + t * KS * 64;
        const bf16x8_t eh = __builtin_bit_cast(bf16x8_t, b[0]);
        const bf16x8_t oh = __builtin_bit_cast(bf16x8_t, b[nfrag]);
        const bf16x8_t el = __builtin_bit_cast(bf16x8_t, b[2 * nfrag]);
        const bf16x8_t ol = __builtin_bit_cast(bf16x8_t, b[3 * nfrag]);
        ae[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(plo, eh, ae[t], 0, 0, 0);
        ao[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qlo, oh, ao[t], 0, 0, 0);
        ae[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(phi, el, ae[t], 0, 0, 0);
        ao[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qhi, ol, ao[t], 0, 0, 0);
        ae[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(phi, eh, ae[t], 0, 0, 0);
        ao[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qhi, oh, ao[t], 0, 0, 0);
        // one tile's B operands live at a time (the accumulators of all
        // tiles already hold 8 NT registers)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // lane holds rows orow..orow+3 of its two output columns per tile (all
    // four exist when orow < R: R % 4 == 0)
    const int comp = cur.comp;
    const long long orow = cur.row0 + 4 * (lane >> 4);
    const bool rows4 = orow < a.R;
    float *out = a.out + comp * a.cout;
    const float *xs = EPI == EPI_ADDX && a.xepi ? a.xepi + comp * a.cout : nullptr;
    float grc[4] = {0.f, 0.f, 0.f, 0.f};
    if (EPI == EPI_ADDGRID && comp != a.xg_axis) {
#pragma unroll
      for (int i = 0; i < 4; ++i) grc[i] = grid_coord(a, comp, orow + i);
    }
    // output element (kc, orow) at obase + kc * ostride (out_offset); the
    // stride is re-read per tile so the compiler does not keep the 2 NT
    // column offsets live across the tile loop
    long long obase;
    uint32_t ostride;
    if (a.rin == 0) {
      obase = orow;
      ostride = (uint32_t)a.R;
    } else {
      const uint32_t b = (uint32_t)orow / (uint32_t)a.rin;
      obase = (long long)b * a.n_out * a.rin + (orow - (long long)b * a.rin);
      ostride = (uint32_t)a.rin;
    }
    asm volatile("" : "+v"(ostride));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int c = t * 16 + (lane & 15);  // m (forward) or x (inverse)
      if (c >= H) continue;
      const int kc0 = INV ? c : 2 * c, kc1 = INV ? N - 1 - c : 2 * c + 1;
      float v0[4], v1[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v0[i] = INV ? ae[t][i] + ao[t][i] : ae[t][i];
        v1[i] = INV ? ae[t][i] - ao[t][i] : ao[t][i];
        if (EPI == EPI_ADDGRID) {
          const bool along = comp == a.xg_axis;
          v0[i] += along ? (float)kc0 : grc[i];
          v1[i] += along ? (float)kc1 : grc[i];
        }
      }
      // rows orow..orow+3 are consecutive in the output (rin % 4 == 0)
      const long long ob0 = obase + (long long)((uint32_t)kc0 * ostride);
      const long long ob1 = obase + (long long)((uint32_t)kc1 * ostride);
      if (rows4) {
        if (EPI == EPI_ADDX && xs) {
          const float4 t0 = *reinterpret_cast<const float4 *>(xs + ob0);
          const float4 t1 = *reinterpret_cast<const float4 *>(xs + ob1);
          v0[0] += t0.x; v0[1] += t0.y; v0[2] += t0.z; v0[3] += t0.w;
          v1[0] += t1.x; v1[1] += t1.y; v1[2] += t1.z; v1[3] += t1.w;
        }
        *reinterpret_cast<float4 *>(out + ob0) = make_float4(v0[0], v0[1], v0[2], v0[3]);
        *reinterpret_cast<float4 *>(out + ob1) = make_float4(v1[0], v1[1], v1[2], v1[3]);
      }
    }
    cur = nxt;
  }
}

// ---------------------------------------------------------------------------
// x axis in one pass: forward DCT-II, prior filter (+ noise, cost), inverse
// DCT-III.  The data is [X][R] per component (x outermost, the R = Ny*Nz
// frequency columns contiguous); a wavefront takes 16 columns and all x.
//
// The ortho DCT matrix has C[k][N-1-x] = (-1)^k C[k][x], so with N even and
// H = N/2 the forward splits into even rows on S[x] = U[x] + U[N-1-x] and odd
// rows on D[x] = U[x] - U[N-1-x] (x < H), and the inverse into
// Ve[x] = sum_m C[2m][x] P[2m], Vo[x] = sum_m C[2m+1][x] P[2m+1] with
// V[x] = Ve + Vo, V[N-1-x] = Ve - Vo: four H x H matrices, half the MFMA
// work, and all four fit in LDS for N <= 192.
//
// Forward: F = E . S with the matrix as the A operand (rows m, 16 per
// tile), the data columns as B, so each lane ends with 4 consecutive m of
// its own column.  Two such tiles are exactly the B operand of the inverse
// (V = E^T . P) once the inverse matrix's k order is permuted to match:
// k-slot 8g + j of k-step s holds m = 32 s + 4 g + j (j < 4) or
// 32 s + 16 + 4 g + j - 4 (j >= 4).  The filtered coefficients never leave
// registers.
// LDS: [AfE | AfO | AiE | AiO] hi, then the same lo; each [tile][kstep][lane]
// uint4 (8 bf16).
struct FusedArgs {
  const float *in;  // [comp][N][R]
  float *out;       // [comp][N][R]
  long long R, cin, cout;
  int n;  // N = contraction length (even)
  const uint32_t *mhi, *mlo;
  const float *D;  // [N][R]
  const double *gamma;
  float sigma2;
  double etarho[3];
  int noise, raw;
  uint32_t k0, k1, sweep, stream;
  const uint32_t *keyx;
  double *cost_part, *raw_part;
};

// A zero the compiler cannot see through.  Added to the LDS matrix base in
// every tile iteration, it keeps the fragment reads inside the loop instead
// of being hoisted out of it (all of them at once would not fit the
// register file).
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}

// Float at element idx of base, addressed by a 32-bit byte offset from the
// (uniform) base: one VGPR per address and the scalar-base load form.
__device__ __forceinline__ float ldf(const float *base, uint32_t idx) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + idx * 4u);
}
__device__ __forceinline__ void stf(float *base, uint32_t idx, float v) {
  *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + idx * 4u) = v;
}

__device__ __forceinline__ f32x4_t mfma3(const uint4 &ah, const uint4 &al, const bf16x8_t &bh,
                                         const bf16x8_t &bl, f32x4_t acc) {
  const bf16x8_t h = __builtin_bit_cast(bf16x8_t, ah);
  const bf16x8_t l = __builtin_bit_cast(bf16x8_t, al);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h, bh, acc, 0, 0, 0);
  return acc;
}

// Forward rows of one parity (PAR 0: even k on sums, 1: odd k on
// differences), prior filter, and the split of the filtered coefficients
// into the inverse's B operand.
template <int NTH, int KSH, int PAR, bool EXACT>
__device__ __forceinline__ void fused_forward(const FusedArgs &a, const uint4 *mh,
                                              const uint4 *ml, const bf16x8_t *bh,
                                              const bf16x8_t *bl, const float (*dv)[4],
                                              bf16x8_t *ph, bf16x8_t *pl, int lane, int comp,
                                              uint32_t c, bool colok, int H, double gm,
                                              float gs2, double &cost, double &raw) {
  constexpr int NF = NTH * KSH * 64;
  const int g = lane >> 4;
  const float sigma = sqrtf(a.sigma2);
#pragma unroll
  for (int sp = 0; sp < KSH; ++sp) {
    float pv[8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = 2 * sp + h;
      if (t < NTH) {
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        const int fb = PAR * NF + t * KSH * 64 + lane;
#pragma unroll
        for (int s = 0; s < KSH; ++s)
          acc = mfma3(mh[fb + s * 64], ml[fb + s * 64], bh[s], bl[s], acc);
        const int m0 = 16 * t + 4 * g;
        float eps[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.noise) {
          // one block of four normals per (column, 4 consecutive k of one parity)
          const u32x4 r = philox4x32_10(
              u32x4{c, ((uint32_t)m0 << 1) | (uint32_t)PAR, a.sweep,
                    ((uint32_t)comp << 16) | a.stream},
              a.keyx ? a.k0 ^ a.keyx[0] : a.k0, a.keyx ? a.k1 ^ a.keyx[1] : a.k1);
          normals4(r, eps);
        }
        // D p^2 summed over the 4 rows in FP32, then carried in FP64
        float q4 = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const bool ok = EXACT || (colok && m0 + r < H);
          // 1 / den and sigma / sqrt(den) from the hardware reciprocal and
          // reciprocal square root (1 ulp; the contraction itself carries
          // about 2^-16 relative error).  Rows and columns out of range
          // have F = D = 0 and no noise.
          const float Dv = dv[t][r];
          const float den = fmaf(gs2, Dv, 1.0f);
          float p = acc[r] * __builtin_amdgcn_rcpf(den);
          if (a.noise) p = fmaf(ok ? eps[r] * sigma : 0.f, __builtin_amdgcn_rsqf(den), p);
          q4 = fmaf(Dv * p, p, q4);
          pv[4 * h + r] = p;
        }
        cost += (0.5 * gm) * (double)q4;
        if (a.raw) raw += a.etarho[comp] * (double)q4;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) pv[4 * h + r] = 0.f;
      }
    }
    split8(pv, ph[sp], pl[sp]);
  }
}

constexpr int kFusedThreads = 512;  // 2 wavefronts per SIMD: 256 VGPRs each

// EXACT: H and R are multiples of 16, so every row and column a lane owns
// is in range.
// Block `bid` of `nblocks` of the fused x pass: tiles bid * 8 + wave,
// strided by nblocks * 8, and cost partial slot bid.
template <int NTH, bool EXACT>
__device__ __forceinline__ void fused_block(const FusedArgs &a, int comps, int bid,
                                            int nblocks) {
  constexpr int KSH = (NTH + 1) / 2;
  constexpr int NF = NTH * KSH * 64;
  extern __shared__ __align__(16) uint4 msh[];
  {
    const uint4 *gh = reinterpret_cast<const uint4 *>(a.mhi);
    const uint4 *gl = reinterpret_cast<const uint4 *>(a.mlo);
    for (int i = threadIdx.x; i < 4 * NF; i += blockDim.x) {
      msh[i] = gh[i];
      msh[4 * NF + i] = gl[i];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
  const int waves_per_block = blockDim.x >> 6;
  const int N = a.n, H = N >> 1;
  const long long tiles_per_comp = (a.R + 15) >> 4;
  const long long total = tiles_per_comp * comps;
  double cost = 0.0, raw = 0.0;
  for (long long tile = (long long)bid * waves_per_block + wave; tile < total;
       tile += (long long)nblocks * waves_per_block) {
    const int comp = (int)(tile / tiles_per_comp);
    const uint32_t c = (uint32_t)((tile % tiles_per_comp) * 16) + (uint32_t)(lane & 15);
    const bool colok = EXACT || c < (uint32_t)a.R;
    const float *in = a.in + comp * a.cin;
    float *out = a.out + comp * a.cout;
    const double gm = a.gamma[comp];
    const float gs2 = (float)gm * a.sigma2;
    // per-iteration (opaque) copies of the loop invariants: the row offsets
    // x * R are formed in 32 bits each tile instead of being hoisted out of
    // the loop as dozens of live 64-bit values; within a tile they advance
    // by additions of R
    const int z0 = opaque_zero();
    const uint4 *mh = msh + z0, *ml = msh + 4 * NF + z0;
    const uint32_t Rz = (uint32_t)a.R + (uint32_t)z0;
    // B operands of the forward: sums and differences of mirrored rows
    bf16x8_t sh[KSH], sl[KSH], dh[KSH], dl[KSH];
#pragma unroll
    for (int s = 0; s < KSH; ++s) {
      float sv[8], dv[8];
      const int x0 = 32 * s + 8 * g;
      uint32_t o1 = (uint32_t)x0 * Rz + c, o2 = (uint32_t)(N - 1 - x0) * Rz + c;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float u1 = 0.f, u2 = 0.f;
        if (colok && x0 + j < H) {
          u1 = ldf(in, o1);
          u2 = ldf(in, o2);
        }
        sv[j] = u1 + u2;
        dv[j] = u1 - u2;
        o1 += Rz;
        o2 -= Rz;
      }
      split8(sv, sh[s], sl[s]);
      split8(dv, dh[s], dl[s]);
    }
    // the prior eigenvalues of this lane's coefficients, loaded with the
    // data so that the filter does not wait on memory tile by tile
    float dve[NTH][4], dvo[NTH][4];
#pragma unroll
    for (int t = 0; t < NTH; ++t) {
      uint32_t od = (uint32_t)(2 * (16 * t + 4 * g)) * Rz + c;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool ok = EXACT || (colok && 16 * t + 4 * g + r < H);
        dve[t][r] = ok ? ldf(a.D, od) : 0.f;
        dvo[t][r] = ok ? ldf(a.D, od + Rz) : 0.f;
        od += 2 * Rz;
      }
    }
    bf16x8_t peh[KSH], pel[KSH], poh[KSH], pol[KSH];
    fused_forward<NTH, KSH, 0, EXACT>(a, mh, ml, sh, sl, dve, peh, pel, lane, comp, c, colok,
                                      H, gm, gs2, cost, raw);
    fused_forward<NTH, KSH, 1, EXACT>(a, mh, ml, dh, dl, dvo, poh, pol, lane, comp, c, colok,
                                      H, gm, gs2, cost, raw);
    // inverse: Ve and Vo per 16-row x tile, then the mirrored pair of rows
#pragma unroll
    for (int tp = 0; tp < NTH; ++tp) {
      f32x4_t ae = {0.f, 0.f, 0.f, 0.f}, ao = {0.f, 0.f, 0.f, 0.f};
      const int fe = 2 * NF + tp * KSH * 64 + lane, fo = 3 * NF + tp * KSH * 64 + lane;
#pragma unroll
      for (int s = 0; s < KSH; ++s) {
        ae = mfma3(mh[fe + s * 64], ml[fe + s * 64], peh[s], pel[s], ae);
        ao = mfma3(mh[fo + s * 64], ml[fo + s * 64], poh[s], pol[s], ao);
      }
      const int x0 = 16 * tp + 4 * g;
      uint32_t o1 = (uint32_t)x0 * Rz + c, o2 = (uint32_t)(N - 1 - x0) * Rz + c;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (EXACT || (colok && x0 + r < H)) {
          stf(out, o1, ae[r] + ao[r]);
          stf(out, o2, ae[r] - ao[r]);
        }
        o1 += Rz;
        o2 -= Rz;
      }
    }
  }
  __syncthreads();
  double *red = reinterpret_cast<double *>(msh);  // matrices no longer needed
  const int t = threadIdx.x;
  red[t] = cost;
  red[kFusedThreads + t] = raw;
  __syncthreads();
  for (int st = kFusedThreads / 2; st > 0; st >>= 1) {
    if (t < st) {
      red[t] += red[t + st];
      red[kFusedThreads + t] += red[kFusedThreads + t + st];
    }
    __syncthreads();
  }
  if (t == 0) {
    a.cost_part[bid] = red[0];
    if (a.raw) a.raw_part[bid] = red[kFusedThreads];
  }
}

template <int NTH, bool EXACT>
__global__ __launch_bounds__(kFusedThreads) void k_dct_x_fused(FusedArgs a, int comps) {
  fused_block<NTH, EXACT>(a, comps, blockIdx.x, gridDim.x);
}

// The fused x pass of several problems in one launch: pair p's blocks
// prefix[p] .. prefix[p + 1] - 1 are its own launch's blocks 0 .. n - 1.
template <int NTH, bool EXACT>
__global__ __launch_bounds__(kFusedThreads) void k_dct_x_fused_batch(const FusedArgs *args,
                                                                    const int *prefix,
                                                                    int npairs, int comps) {
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int p = batch_pair(prefix, npairs, b);
  const FusedArgs a = args[p];
  fused_block<NTH, EXACT>(a, comps, b - prefix[p], prefix[p + 1] - prefix[p]);
}

int g_num_cus = 0;

// Deterministic sum of `count` partials (fixed assignment and tree order).
__device__ __forceinline__ void sum_partials_block(const double *part, long long count,
                                                   double *out) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long i = threadIdx.x; i < count; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *part, long long count,
                                                      double *out) {
  sum_partials_block(part, count, out);
}

// k_sum_partials that also writes (stats[0], stats[1], sum) to page-locked
// host memory, in place of a copy after it.
__global__ __launch_bounds__(256) void k_sum_partials_host(const double *part, long long count,
                                                           double *out, const double *stats,
                                                           double *host) {
  sum_partials_block(part, count, out);
  if (threadIdx.x == 0) {
    host[0] = stats[0];
    host[1] = stats[1];
    host[2] = *out;
    __threadfence_system();
  }
}

struct SumArgs {
  const double *part;
  long long count;
  double *out;
};

__global__ __launch_bounds__(256) void k_sum_partials_batch(const SumArgs *args) {
  const SumArgs a = args[blockIdx.x];
  sum_partials_block(a.part, a.count, a.out);
}

// nu ~ Gamma(alpha0 + N nv / 2, scale 1 / (beta0 + nv raw / 2)), then
// gamma_k = nu rho_x rho_y rho_z eta_k^2 nv
// (NonLinearTransformation.cpp:116-124, :140-148).  raw is the sum over
// components of eta_n^2 rho D p^2 over the sampled (noisy) coefficients p.
__global__ void k_draw_nu(const double *raw, double *gamma, int N, double nv,
                          double rho_prod, double e0, double e1, double e2,
                          double alpha0, double beta0,
                          uint32_t k0, uint32_t k1, const uint32_t *keyx, uint32_t sweep) {
  if (threadIdx.x != 0) return;
  if (keyx) {
    k0 ^= keyx[0];
    k1 ^= keyx[1];
  }
  const double a_post = alpha0 + 0.5 * (double)N * nv;
  const double b_post = beta0 + 0.5 * nv * (*raw);
  const double gm = gamma_mt(a_post, u32x4{0u, sweep, 0u, kStreamNu}, k0, k1);
  const double nu = gm / b_post;
  gamma[0] = nu * rho_prod * e0 * e0 * nv;
  gamma[1] = nu * rho_prod * e1 * e1 * nv;
  gamma[2] = nu * rho_prod * e2 * e2 * nv;
}

template <int KS, bool VEC, int PRO, int EPI>
void launch_ks(dim3 grid, hipStream_t s, const PassArgs &a, const SumJob &j) {
  hipLaunchKernelGGL((k_contract<KS, VEC, PRO, EPI>), grid, dim3(256), 0, s, a, j);
}

template <int KS, bool VEC, int PRO, int EPI>
void launch_persist(int blocks, size_t lds, int comps, hipStream_t s, const PassArgs &a,
                    const SumJob &j) {
  auto k = k_contract_lds<KS, VEC, PRO, EPI>;
  static std::atomic<bool> attr_set{false};
  if (!attr_set) {
    HIP_OK(hipFuncSetAttribute((const void *)k,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPersistMaxLds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kPersistThreads), lds, s, a, comps, j);
}

template <bool VEC, int PRO, int EPI>
void launch_lds(int blocks, size_t lds, int comps, hipStream_t s, const PassArgs &a,
                const SumJob &j) {
  switch (a.ks) {
    case 1: launch_persist<1, VEC, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 2: launch_persist<2, VEC, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 3: launch_persist<3, VEC, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 4: launch_persist<4, VEC, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 5: launch_persist<5, VEC, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    default: launch_persist<6, VEC, PRO, EPI>(blocks, lds, comps, s, a, j); break;
  }
}

template <int NT, bool INV, int PRO, int EPI>
void launch_sym_nt(int blocks, size_t lds, int comps, hipStream_t s, const PassArgs &a,
                   const SumJob &j) {
  auto k = k_contract_sym<NT, INV, PRO, EPI>;
  static std::atomic<bool> attr_set{false};
  if (!attr_set) {
    HIP_OK(hipFuncSetAttribute((const void *)k,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPersistMaxLds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kSymThreads), lds, s, a, comps, j);
}

// NT = ceil(H / 16) output tiles per parity, 2..kSymMaxTiles
template <bool INV, int PRO, int EPI>
void launch_sym(int blocks, size_t lds, int comps, hipStream_t s, const PassArgs &a,
                const SumJob &j) {
  switch ((a.n_in / 2 + 15) / 16) {
    case 2: launch_sym_nt<2, INV, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 3: launch_sym_nt<3, INV, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 4: launch_sym_nt<4, INV, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    case 5: launch_sym_nt<5, INV, PRO, EPI>(blocks, lds, comps, s, a, j); break;
    default: launch_sym_nt<6, INV, PRO, EPI>(blocks, lds, comps, s, a, j); break;
  }
}

template <bool VEC, int PRO, int EPI>
void launch(dim3 grid, hipStream_t s, const PassArgs &a, const SumJob &j) {
  switch (a.ks) {
    case 1: launch_ks<1, VEC, PRO, EPI>(grid, s, a, j); break;
    case 2: launch_ks<2, VEC, PRO, EPI>(grid, s, a, j); break;
    case 3: launch_ks<3, VEC, PRO, EPI>(grid, s, a, j); break;
    case 4: launch_ks<4, VEC, PRO, EPI>(grid, s, a, j); break;
    case 5: launch_ks<5, VEC, PRO, EPI>(grid, s, a, j); break;
    case 6: launch_ks<6, VEC, PRO, EPI>(grid, s, a, j); break;
    case 7: launch_ks<7, VEC, PRO, EPI>(grid, s, a, j); break;
    default: launch_ks<8, VEC, PRO, EPI>(grid, s, a, j); break;
  }
}

uint16_t host_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

float host_bf16_to_float(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

}  // namespace

// Swizzle M (n_in x n_out, row-major, double) into the MFMA B-operand order:
// for n-tile t, k-step s, lane l the 8 bf16 values M[32 s + 8 (l/16) + j][16 t + l%16].
void build_swizzled(const std::vector<double> &M, int n_in, int n_out,
                    std::vector<uint32_t> &hi, std::vector<uint32_t> &lo) {
  const int ks = (n_in + 31) / 32, nt = (n_out + 15) / 16;
  hi.assign((size_t)nt * ks * 64 * 4, 0u);
  lo.assign((size_t)nt * ks * 64 * 4, 0u);
  for (int t = 0; t < nt; ++t)
    for (int s = 0; s < ks; ++s)
      for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
          const int k = 32 * s + 8 * (l / 16) + j, c = 16 * t + (l % 16);
          const float v = (k < n_in && c < n_out) ? (float)M[(size_t)k * n_out + c] : 0.f;
          const uint16_t h = host_bf16(v);
          const uint16_t r = host_bf16(v - host_bf16_to_float(h));
          const size_t w = (((size_t)t * ks + s) * 64 + l) * 4 + j / 2;
          hi[w] |= (uint32_t)h << (16 * (j & 1));
          lo[w] |= (uint32_t)r << (16 * (j & 1));
        }
}

// Orthonormal DCT-II matrix entry C[k][j] = s_k cos(pi (2j+1) k / (2n)).
static double dct_entry(int n, int k, int j) {
  const double s = k == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
  return s * std::cos(M_PI * (2.0 * j + 1.0) * k / (2.0 * n));
}

// The four H x H matrices of the fused x pass (N = n even, H = N / 2), in
// the A-operand order of k_dct_x_fused: E[m][x] = C_N[2m][x] and
// O[m][x] = C_N[2m+1][x] for the forward (row m = 16 t + lane%16, k = x =
// 32 s + 8 (lane/16) + j), then E^T and O^T for the inverse (row x, k = m in
// the permuted order 32 s + 4 (lane/16) + j, j < 4, or
// 32 s + 16 + 4 (lane/16) + j - 4).  Entries outside H x H are zero.
static void build_fused_x(int n, std::vector<uint32_t> &hi, std::vector<uint32_t> &lo) {
  const int H = n / 2, nt = (H + 15) / 16, ks = (nt + 1) / 2;
  const size_t nfrag = (size_t)nt * ks * 64;
  hi.assign(4 * nfrag * 4, 0u);
  lo.assign(4 * nfrag * 4, 0u);
  auto E = [&](int m, int x) { return m < H && x < H ? dct_entry(n, 2 * m, x) : 0.0; };
  auto O = [&](int m, int x) { return m < H && x < H ? dct_entry(n, 2 * m + 1, x) : 0.0; };
  auto put = [&](int arr, int t, int s, int l, int j, double v) {
    const float f = (float)v;
    const uint16_t h = host_bf16(f);
    const uint16_t r = host_bf16(f - host_bf16_to_float(h));
    const size_t w = ((size_t)arr * nfrag + ((size_t)t * ks + s) * 64 + l) * 4 + j / 2;
    hi[w] |= (uint32_t)h << (16 * (j & 1));
    lo[w] |= (uint32_t)r << (16 * (j & 1));
  };
  for (int t = 0; t < nt; ++t)
    for (int s = 0; s < ks; ++s)
      for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
          const int row = 16 * t + (l % 16), g = l / 16;
          const int xf = 32 * s + 8 * g + j;
          const int mp = 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
          put(0, t, s, l, j, E(row, xf));
          put(1, t, s, l, j, O(row, xf));
          put(2, t, s, l, j, E(mp, row));
          put(3, t, s, l, j, O(mp, row));
        }
}

const PassMatrix &MatrixCache::get(int kind, int n_in, int n_out,
                                   hipStream_t stream) {
  const long long key = ((long long)kind << 40) | ((long long)n_in << 20) | n_out;
  auto it = cache_.find(key);
  if (it != cache_.end()) return *it->second;
  if (kind == kDctFusedX) {
    std::vector<uint32_t> hi, lo;
    build_fused_x(n_in, hi, lo);
    auto pm = std::make_unique<PassMatrix>();
    pm->n_in = n_in;
    pm->n_out = n_out;
    pm->hi.upload(hi.data(), hi.size(), stream);
    pm->lo.upload(lo.data(), lo.size(), stream);
    HIP_OK(hipStreamSynchronize(stream));
    auto &ref = *pm;
    cache_[key] = std::move(pm);
    return ref;
  }
  std::vector<double> M((size_t)n_in * n_out);
  for (int k = 0; k < n_in; ++k)
    for (int c = 0; c < n_out; ++c) {
      double v;
      if (kind == kDctForward) {  // Out[c] = sum_k In[k] C_n[c][k]
        v = dct_entry(n_in, c, k);
      } else {  // inverse at size n_out from the first n_in coefficients
        v = dct_entry(n_out, k, c);
      }
      M[(size_t)k * n_out + c] = v;
    }
  std::vector<uint32_t> hi, lo;
  build_swizzled(M, n_in, n_out, hi, lo);
  auto pm = std::make_unique<PassMatrix>();
  pm->n_in = n_in;
  pm->n_out = n_out;
  pm->hi.upload(hi.data(), hi.size(), stream);
  pm->lo.upload(lo.data(), lo.size(), stream);
  if (n_in == n_out && n_in % 16 == 0 && n_in / 2 > 16 && n_in / 2 <= 16 * kSymMaxTiles) {
    // even / odd halves: C_n[k][n-1-x] = (-1)^k C_n[k][x]
    const int H = n_in / 2;
    const bool inv = kind == kDctInverse;
    std::vector<double> E((size_t)H * H), O((size_t)H * H);
    for (int r = 0; r < H; ++r)
      for (int c = 0; c < H; ++c) {
        // forward: row r = x, column c = m; inverse: row r = m, column c = x
        const int x = inv ? c : r, m = inv ? r : c;
        E[(size_t)r * H + c] = dct_entry(n_in, 2 * m, x);
        O[(size_t)r * H + c] = dct_entry(n_in, 2 * m + 1, x);
      }
    std::vector<uint32_t> ehi, elo, ohi, olo;
    build_swizzled(E, H, H, ehi, elo);
    build_swizzled(O, H, H, ohi, olo);
    ehi.insert(ehi.end(), ohi.begin(), ohi.end());
    elo.insert(elo.end(), olo.begin(), olo.end());
    pm->shi.upload(ehi.data(), ehi.size(), stream);
    pm->slo.upload(elo.data(), elo.size(), stream);
    pm->sym_inverse = inv ? 1 : 0;
    pm->sym_ks = (H + 31) / 32;
  }
  HIP_OK(hipStreamSynchronize(stream));
  auto &ref = *pm;
  cache_[key] = std::move(pm);
  return ref;
}

namespace {

// A contraction pass ready to launch: its arguments and kernel variant.
struct PassPlan {
  PassArgs a{};
  int pro = PRO_NONE, epi = EPI_STORE, comps = 0;
  bool vec = false, persist = false;
  // persistent even / odd contraction (k_contract_sym) of a forward or
  // inverse square DCT
  bool sym = false, sym_inv = false;
  dim3 grid;
  int blocks = 0;
  size_t lds = 0;
  // optional partial sum run by the launch (SumJob), none by default
  SumJob sum{};
};

PassPlan plan_contract(const PassMatrix &m, const float *in, float *out, int comps,
                       long long cin, long long cout, long long R, const float *xpro,
                       const float *xepi, int epi, const FilterArgs *f, long long rin,
                       const GridCoords *xg) {
  BINDER_REQUIRE(m.n_in <= 32 * kMaxKSteps, "contraction length above 256");
  PassPlan P;
  PassArgs &a = P.a;
  a.in = in;
  a.out = out;
  a.xpro = xpro;
  a.xepi = xepi;
  a.bhi = m.hi.get();
  a.blo = m.lo.get();
  a.R = R;
  a.rin = rin;
  a.cin = cin;
  a.cout = cout;
  a.n_in = m.n_in;
  a.n_out = m.n_out;
  a.ks = (m.n_in + 31) / 32;
  if (f) {
    a.D = f->D;
    a.gamma = f->gamma;
    a.sigma2 = (float)f->sigma2;
    for (int i = 0; i < 3; ++i) a.etarho[i] = f->etarho[i];
    a.noise = f->noise;
    a.raw = f->raw;
    a.k0 = f->k0;
    a.k1 = f->k1;
    a.keyx = f->keyx;
    a.sweep = f->sweep;
    a.stream = f->stream;
    a.cost_part = f->cost_part;
    a.raw_part = f->raw_part;
  }
  P.vec = (m.n_in % 8 == 0) && (R % 4 == 0) && (cin % 4 == 0) && (cout % 4 == 0) &&
          (rin % 4 == 0);
  P.grid = dim3((unsigned)((R + kRowsPerBlock - 1) / kRowsPerBlock), comps);
  P.comps = comps;
  int pro = xpro != nullptr ? PRO_X : PRO_NONE;
  if (xg) {
    a.xg_axis = xg->axis;
    for (int i = 0; i < 3; ++i) {
      a.xg_div[i] = xg->div[i] > 0 ? xg->div[i] : 1;
      a.xg_mod[i] = xg->mod[i];
    }
    if (pro == PRO_X) pro = PRO_GRID;
    if (epi == EPI_ADDX) epi = EPI_ADDGRID;
  }
  BINDER_REQUIRE(pro == PRO_NONE || epi == EPI_STORE || epi == EPI_FILTER,
                 "a pass either subtracts x in the prologue or adds it in the epilogue");
  P.pro = pro;
  P.epi = epi;
  const size_t mat_lds = (size_t)((m.n_out + 15) / 16) * a.ks * 64 * 16 * 2;
  // the filter epilogue reuses the matrix area for a 2 x 1024 double reduction
  P.lds = mat_lds > 16384 ? mat_lds : 16384;
  const long long tiles = ((R + 15) / 16) * comps;
  P.persist = a.ks <= 6 && mat_lds <= kPersistMaxLds && tiles >= 2048;
  if (P.persist && P.vec && m.sym_ks > 0) {
    // the even / odd halves: forward passes with a prologue or a plain
    // store, inverse passes with a plain store or the x epilogue
    const bool ok = m.sym_inverse ? (pro == PRO_NONE && epi != EPI_FILTER) : epi == EPI_STORE;
    const size_t sym_lds = (size_t)4 * ((m.n_in / 2 + 15) / 16) * m.sym_ks * 64 * 16;
    if (ok && sym_lds <= kPersistMaxLds) {
      P.sym = true;
      P.sym_inv = m.sym_inverse != 0;
      a.bhi = m.shi.get();
      a.blo = m.slo.get();
      a.ks = m.sym_ks;
      P.lds = sym_lds;
    }
  }
  P.blocks = P.persist ? persistent_blocks() : (int)(P.grid.x * P.grid.y);
  return P;
}

void launch_contract_blocks(const PassPlan &P, hipStream_t stream, int blocks);

long long launch_contract(const PassPlan &P, hipStream_t stream, KernelTimer *timer,
                          const char *name) {
  // the even / odd kernel runs two blocks per CU by default (kSymThreads)
  int blocks = persistent_blocks() * (P.sym ? 2 : 1);
  if (P.persist && P.epi != EPI_FILTER && 2 * P.lds <= 160 * 1024) {
    // persistent blocks: every tile's outputs are computed by one wavefront
    // whatever block runs it, so the default, twice and (for the even / odd
    // kernel, whose wavefronts stream tiles through a one-k-step prefetch)
    // half as many blocks give the same results; the autotuner picks the
    // faster
    const int cands[3] = {blocks, 2 * blocks, blocks / 2};
    const int ncand = P.sym && blocks >= 2 ? 3 : 2;
    const long long key = ((long long)P.a.n_in << 52) ^ ((long long)P.a.n_out << 42) ^
                          ((long long)P.a.R << 12) ^ ((long long)P.comps << 8) ^
                          ((long long)P.pro << 5) ^ ((long long)P.epi << 2) ^ (P.vec ? 1 : 0) ^
                          (P.sym ? 2 : 0);
    blocks = Autotuner::get().choose("contract", key, cands, ncand, blocks, stream,
                                     [&](int b) { launch_contract_blocks(P, stream, b); });
  }
  TimedLaunch tl(timer, name, stream);
  launch_contract_blocks(P, stream, blocks);
  return P.blocks;
}

void launch_contract_blocks(const PassPlan &P, hipStream_t stream, int blocks) {
  const PassArgs &a = P.a;
  const SumJob &j = P.sum;
  const size_t lds = P.lds;
  const int comps = P.comps;
  const dim3 grid = P.grid;
  if (P.sym) {
    if (P.sym_inv) {
      if (P.epi == EPI_ADDX)
        launch_sym<true, PRO_NONE, EPI_ADDX>(blocks, lds, comps, stream, a, j);
      else if (P.epi == EPI_ADDGRID)
        launch_sym<true, PRO_NONE, EPI_ADDGRID>(blocks, lds, comps, stream, a, j);
      else launch_sym<true, PRO_NONE, EPI_STORE>(blocks, lds, comps, stream, a, j);
    } else {
      if (P.pro == PRO_X)
        launch_sym<false, PRO_X, EPI_STORE>(blocks, lds, comps, stream, a, j);
      else if (P.pro == PRO_GRID)
        launch_sym<false, PRO_GRID, EPI_STORE>(blocks, lds, comps, stream, a, j);
      else launch_sym<false, PRO_NONE, EPI_STORE>(blocks, lds, comps, stream, a, j);
    }
    HIP_LAUNCH_OK();
    return;
  }
#define BINDER_GO(V, PR, E)                                            \
  do {                                                                 \
    if (P.persist) launch_lds<V, PR, E>(blocks, lds, comps, stream, a, j); \
    else launch<V, PR, E>(grid, stream, a, j);                         \
  } while (0)
#define BINDER_DISPATCH(V)                                                     \
  if (P.pro == PRO_X) {                                                        \
    if (P.epi == EPI_FILTER) BINDER_GO(V, PRO_X, EPI_FILTER);                  \
    else BINDER_GO(V, PRO_X, EPI_STORE);                                       \
  } else if (P.pro == PRO_GRID) {                                              \
    if (P.epi == EPI_FILTER) BINDER_GO(V, PRO_GRID, EPI_FILTER);               \
    else BINDER_GO(V, PRO_GRID, EPI_STORE);                                    \
  } else if (P.epi == EPI_FILTER) BINDER_GO(V, PRO_NONE, EPI_FILTER);          \
  else if (P.epi == EPI_ADDX) BINDER_GO(V, PRO_NONE, EPI_ADDX);                \
  else if (P.epi == EPI_ADDGRID) BINDER_GO(V, PRO_NONE, EPI_ADDGRID);          \
  else BINDER_GO(V, PRO_NONE, EPI_STORE);
  if (P.vec) {
    BINDER_DISPATCH(true)
  } else {
    BINDER_DISPATCH(false)
  }
#undef BINDER_DISPATCH
#undef BINDER_GO
  HIP_LAUNCH_OK();
}

}  // namespace

long long contract_pass(hipStream_t stream, const PassMatrix &m, const float *in,
                        float *out, int comps, long long cin, long long cout,
                        long long R, const float *xpro, const float *xepi, int epi,
                        const FilterArgs *f, KernelTimer *timer, const char *name,
                        long long rin, const GridCoords *xg) {
  return launch_contract(plan_contract(m, in, out, comps, cin, cout, R, xpro, xepi, epi, f, rin,
                                       xg),
                         stream, timer, name);
}

bool fused_x_supported(int n) { return n >= 2 && n % 2 == 0 && n <= 192; }

template <int NTH, bool EXACT>
void launch_fused(int blocks, size_t lds, int comps, hipStream_t s, const FusedArgs &a) {
  auto k = k_dct_x_fused<NTH, EXACT>;
  static std::atomic<bool> attr_set{false};
  if (!attr_set) {
    HIP_OK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPersistMaxLds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kFusedThreads), lds, s, a, comps);
}

namespace {

// The fused x pass ready to launch.
struct FusedPlan {
  FusedArgs a{};
  int nt = 0, comps = 0;
  bool exact = false;
  int blocks = 0;
  size_t lds = 0;
};

FusedPlan plan_fused(const PassMatrix &m, const float *in, float *out, int comps,
                     long long cstride, long long R, const FilterArgs &f) {
  const int n = m.n_in;
  BINDER_REQUIRE(fused_x_supported(n), "fused x pass needs an even length <= 192");
  FusedPlan P;
  FusedArgs &a = P.a;
  a.in = in;
  a.out = out;
  a.R = R;
  a.cin = cstride;
  a.cout = cstride;
  a.n = n;
  a.mhi = m.hi.get();
  a.mlo = m.lo.get();
  a.D = f.D;
  a.gamma = f.gamma;
  a.sigma2 = (float)f.sigma2;
  for (int i = 0; i < 3; ++i) a.etarho[i] = f.etarho[i];
  a.noise = f.noise;
  a.raw = f.raw;
  a.k0 = f.k0;
  a.k1 = f.k1;
  a.keyx = f.keyx;
  a.sweep = f.sweep;
  a.stream = f.stream;
  a.cost_part = f.cost_part;
  a.raw_part = f.raw_part;
  const int H = n / 2, nt = (H + 15) / 16, ks = (nt + 1) / 2;
  const size_t mat = (size_t)8 * nt * ks * 64 * 16;
  P.lds = mat > 16384 ? mat : 16384;  // the cost reduction reuses it
  const long long tiles = ((R + 15) / 16) * comps;
  long long blocks = (tiles + kFusedThreads / 64 - 1) / (kFusedThreads / 64);
  if (blocks > persistent_blocks()) blocks = persistent_blocks();
  P.blocks = (int)blocks;
  P.nt = nt < 6 ? nt : 6;
  P.exact = H % 16 == 0 && R % 16 == 0;
  P.comps = comps;
  return P;
}

long long launch_fused_plan(const FusedPlan &P, hipStream_t stream, KernelTimer *timer,
                            const char *name) {
  TimedLaunch tl(timer, name, stream);
#define BINDER_FUSED(NT)                                                                 \
  case NT:                                                                               \
    if (P.exact) launch_fused<NT, true>(P.blocks, P.lds, P.comps, stream, P.a);           \
    else launch_fused<NT, false>(P.blocks, P.lds, P.comps, stream, P.a);                  \
    break;
  switch (P.nt) {
    BINDER_FUSED(1)
    BINDER_FUSED(2)
    BINDER_FUSED(3)
    BINDER_FUSED(4)
    BINDER_FUSED(5)
    default:
      if (P.exact) launch_fused<6, true>(P.blocks, P.lds, P.comps, stream, P.a);
      else launch_fused<6, false>(P.blocks, P.lds, P.comps, stream, P.a);
  }
#undef BINDER_FUSED
  HIP_LAUNCH_OK();
  return P.blocks;
}

}  // namespace

long long fused_x_pass(hipStream_t stream, const PassMatrix &m, const float *in, float *out,
                       int comps, long long cstride, long long R, const FilterArgs &f,
                       KernelTimer *timer, const char *name) {
  return launch_fused_plan(plan_fused(m, in, out, comps, cstride, R, f), stream, timer, name);
}

// The five passes of the EM smoothing (3D, fused x): z forward (subtracts
// x), y forward, fused x (filter, cost partials), the cost sum, z inverse,
// y inverse (adds x) -- exactly the sequence run() launches.
struct EmSmoothPlan {
  PassPlan z_fwd, y_fwd, z_inv, y_inv;
  FusedPlan x;
  SumArgs sum{};
};

int persistent_blocks() {
  if (g_num_cus == 0) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&g_num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (g_num_cus <= 0) g_num_cus = 256;
  }
  return g_num_cus;
}

long long contract_blocks(long long R, int comps) {
  const long long nonpersist = ((R + kRowsPerBlock - 1) / kRowsPerBlock) * comps;
  const long long p = persistent_blocks();
  return nonpersist > p ? nonpersist : p;
}

// D from its separable FP64 terms, in the operation order of the host loop
// (binder_dct_prior_eigenvalues): ((tx + ty) + tz), squared by one multiply.
__global__ __launch_bounds__(256) void k_dct_prior(float *__restrict__ D,
                                                   const double *__restrict__ t, int nx, int ny,
                                                   int nz, int order) {
  const long long nv = (long long)nx * ny * nz;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int z = (int)(i % nz);
  const int y = (int)((i / nz) % ny);
  const int x = (int)(i / ((long long)ny * nz));
  const double sxy = t[x] + t[nx + y];
  const double s = nz > 1 ? sxy + t[nx + ny + z] : sxy;
  D[i] = (float)(order == 1 ? s : s * s);
}

Smoother::Smoother(int nx, int ny, int nz, const double gamma[3],
                   const double rho[3], const double eta[3],
                   const float *D_host, double sigma, hipStream_t stream)
    : nx_(nx), ny_(ny), nz_(nz), N_(nz == 1 ? 2 : 3),
      nv_((long long)nx * ny * nz), sigma_(sigma), stream_(stream) {
  BINDER_REQUIRE(nx >= 2 && ny >= 2 && nz >= 1,
                 "deformation grid must be at least 2 x 2 (x 1)");
  BINDER_REQUIRE(nx <= 256 && ny <= 256 && nz <= 256,
                 "each grid axis must be at most 256 voxels");
  D_.upload(D_host, nv_, stream_);
  init_buffers(gamma, rho, eta);
}

Smoother::Smoother(int nx, int ny, int nz, const double gamma[3], const double rho[3],
                   const double eta[3], const double *axis_terms, int order, double sigma,
                   hipStream_t stream)
    : nx_(nx), ny_(ny), nz_(nz), N_(nz == 1 ? 2 : 3),
      nv_((long long)nx * ny * nz), sigma_(sigma), stream_(stream) {
  BINDER_REQUIRE(nx >= 2 && ny >= 2 && nz >= 1,
                 "deformation grid must be at least 2 x 2 (x 1)");
  BINDER_REQUIRE(nx <= 256 && ny <= 256 && nz <= 256,
                 "each grid axis must be at most 256 voxels");
  BINDER_REQUIRE(order == 1 || order == 2, "device-made D supports order 1 or 2");
  DBuf<double> t;
  t.upload(axis_terms, (size_t)nx + ny + nz, stream_);
  D_.alloc(nv_);
  hipLaunchKernelGGL(k_dct_prior, dim3((unsigned)((nv_ + 255) / 256)), dim3(256), 0, stream_,
                     D_.get(), t.get(), nx, ny, nz, order);
  HIP_LAUNCH_OK();
  init_buffers(gamma, rho, eta);  // synchronizes before t is freed
}

void Smoother::init_buffers(const double gamma[3], const double rho[3], const double eta[3]) {
  const int ny = ny_, nz = nz_;
  for (int i = 0; i < 3; ++i) {
    rho_[i] = rho[i];
    eta_[i] = eta[i];
  }
  nu_ = gamma[0] / (rho[0] * rho[1] * rho[2] * eta[0] * eta[0] * (double)nv_);
  gamma_.upload(gamma, 3, stream_);
  bufA_.alloc((size_t)N_ * nv_);
  bufB_.alloc((size_t)N_ * nv_);
  // the filter pass contracts x over R = ny*nz rows
  const long long nparts = contract_blocks((long long)ny * nz, N_);
  part_cost_.alloc(nparts);
  part_raw_.alloc(nparts);
  scal_.alloc(4);
  HIP_OK(hipStreamSynchronize(stream_));
}

void Smoother::set_gamma(const double g[3]) {
  HIP_OK(hipMemcpyAsync(gamma_.get(), g, 3 * sizeof(double),
                        hipMemcpyHostToDevice, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
}

void Smoother::get_gamma(double g[3]) {
  HIP_OK(hipMemcpyAsync(g, gamma_.get(), 3 * sizeof(double),
                        hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
}

double Smoother::flops() const {
  const double dims[3] = {(double)nx_, (double)ny_, (double)nz_};
  double per_voxel = 0.0;
  for (int a = 0; a < 3; ++a)
    if (dims[a] > 1) per_voxel += 2.0 * dims[a];
  return 2.0 * per_voxel * (double)nv_ * N_;  // forward + inverse
}

void Smoother::run(const float *dream, const float *x, float *out,
                   const SmoothOptions &opt, KernelTimer *timer) {
  hipStream_t st = opt.exec_stream ? opt.exec_stream : stream_;
  FilterArgs f{};
  f.D = D_.get();
  f.gamma = gamma_.get();
  f.sigma2 = sigma_ * sigma_;
  const double rp = rho_[0] * rho_[1] * rho_[2];
  for (int i = 0; i < 3; ++i) f.etarho[i] = eta_[i] * eta_[i] * rp;
  f.noise = opt.sample ? 1 : 0;
  f.raw = opt.sample_gamma ? 1 : 0;
  f.k0 = opt.key0;
  f.k1 = opt.key1;
  f.keyx = opt.keyx;
  f.sweep = opt.sweep;
  f.stream = opt.stream;
  f.cost_part = part_cost_.get();
  f.raw_part = part_raw_.get();

  float *bufs[2] = {bufA_.get(), bufB_.get()};
  long long nparts = 0;  // partial-sum slots written by the filter pass
  BINDER_REQUIRE(!opt.host_stats || (opt.want_prior && opt.stats_in),
                 "host statistics need the prior cost and the device statistics");
  auto sum_kernel = [&](const double *part, long long count, double *out) {
    TimedLaunch tl(timer, "sum_partials", st);
    if (opt.host_stats)
      hipLaunchKernelGGL(k_sum_partials_host, dim3(1), dim3(256), 0, st, part, count, out,
                         opt.stats_in, opt.host_stats);
    else
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, count, out);
    HIP_LAUNCH_OK();
  };
  // The EM's prior-cost sum rides on the pass after the filter (its first
  // wavefront sums the filter pass's partials) instead of a launch of its
  // own; the sampler keeps the separate launches.
  auto fold_sum = [&](PassPlan &P, const double *part, long long count, double *out) {
    P.sum = SumJob{part, count, out, opt.host_stats ? opt.stats_in : nullptr, opt.host_stats};
  };
  auto sum_partials = [&]() {
    if (opt.want_prior)
      sum_kernel(part_cost_.get(), nparts, opt.prior_out ? opt.prior_out : scal_.get());
    if (opt.sample && opt.sample_gamma) {
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st,
                         part_raw_.get(), nparts, scal_.get() + 1);
      HIP_LAUNCH_OK();
    }
  };
  using MC = MatrixCache;
  if (fused_x_supported(nx_) && nz_ > 1 && !opt.sample) {
    // the EM smoothing: the same plan a SmoothBatch launches
    EmSmoothPlan P;
    plan_em(dream, x, out, opt.want_prior ? (opt.prior_out ? opt.prior_out : scal_.get())
                                          : nullptr, P, st);
    launch_contract(P.z_fwd, st, timer, "dct_z_fwd");
    launch_contract(P.y_fwd, st, timer, "dct_y_fwd");
    launch_fused_plan(P.x, st, timer, "dct_x_fused");
    if (P.sum.out) fold_sum(P.z_inv, P.sum.part, P.sum.count, P.sum.out);
    launch_contract(P.z_inv, st, timer, "dct_z_inv");
    launch_contract(P.y_inv, st, timer, "dct_y_inv");
    return;
  }
  if (fused_x_supported(nx_)) {
    // Five passes (3D) or three (2D).  The y and z passes contract their
    // contiguous axis and move it behind x ([X][Y][Z] -> [X][KZ][Y] ->
    // [X][KY][KZ] and back), so x stays outermost for the fused x pass,
    // which transforms, filters and transforms back in one read and write.
    const PassMatrix &mx = mats_.get(MC::kDctFusedX, nx_, nx_, st);
    const PassMatrix &fy = mats_.get(MC::kDctForward, ny_, ny_, st);
    const PassMatrix &iy = mats_.get(MC::kDctInverse, ny_, ny_, st);
    const bool gen = index_grid_ && x != nullptr;
    if (nz_ > 1) {
      const PassMatrix &fz = mats_.get(MC::kDctForward, nz_, nz_, st);
      const PassMatrix &iz = mats_.get(MC::kDctInverse, nz_, nz_, st);
      // z forward: row = x * Ny + y, k = z; y inverse: row = x * Nz + z, k' = y
      const GridCoords gz{2, {ny_, 1, 1}, {0, ny_, 0}};
      const GridCoords gy{1, {nz_, 1, 1}, {0, 0, nz_}};
      contract_pass(st, fz, dream, bufs[0], N_, nv_, nv_, nv_ / nz_, x, nullptr,
                    EPI_STORE, nullptr, timer, "dct_z_fwd", ny_, gen ? &gz : nullptr);
      contract_pass(st, fy, bufs[0], bufs[1], N_, nv_, nv_, nv_ / ny_, nullptr, nullptr,
                    EPI_STORE, nullptr, timer, "dct_y_fwd", nz_);
      nparts = fused_x_pass(st, mx, bufs[1], bufs[0], N_, nv_, (long long)ny_ * nz_, f,
                            timer, "dct_x_fused");
      PassPlan zi = plan_contract(iz, bufs[0], bufs[1], N_, nv_, nv_, nv_ / nz_, nullptr,
                                  nullptr, EPI_STORE, nullptr, ny_, nullptr);
      if (!opt.sample && opt.want_prior)
        fold_sum(zi, part_cost_.get(), nparts, opt.prior_out ? opt.prior_out : scal_.get());
      else
        sum_partials();
      launch_contract(zi, st, timer, "dct_z_inv");
      contract_pass(st, iy, bufs[1], out, N_, nv_, nv_, nv_ / ny_, nullptr, x, EPI_ADDX,
                    nullptr, timer, "dct_y_inv", nz_, gen ? &gy : nullptr);
    } else {
      // rows are x, k (k') is y
      const GridCoords g2{1, {1, 1, 1}, {0, 0, 0}};
      contract_pass(st, fy, dream, bufs[0], N_, nv_, nv_, nx_, x, nullptr, EPI_STORE,
                    nullptr, timer, "dct_y_fwd", 1, gen ? &g2 : nullptr);
      nparts = fused_x_pass(st, mx, bufs[0], bufs[1], N_, nv_, ny_, f, timer,
                            "dct_x_fused");
      PassPlan yi = plan_contract(iy, bufs[1], out, N_, nv_, nv_, nx_, nullptr, x, EPI_ADDX,
                                  nullptr, 1, gen ? &g2 : nullptr);
      if (!opt.sample && opt.want_prior)
        fold_sum(yi, part_cost_.get(), nparts, opt.prior_out ? opt.prior_out : scal_.get());
      else
        sum_partials();
      launch_contract(yi, st, timer, "dct_y_inv");
    }
  } else {
    // Six rotating passes: axes contracted innermost first, each pass
    // rotating the layout, so after one pass per axis the data is back in
    // [X][Y][Z] order; the filter rides on the last forward pass.
    int dims[3], na = 0;
    const char *fwd_names[3], *inv_names[3];
    if (nz_ > 1) {
      dims[na] = nz_;
      fwd_names[na] = "dct_z_fwd";
      inv_names[na++] = "dct_z_inv";
    }
    dims[na] = ny_;
    fwd_names[na] = "dct_y_fwd";
    inv_names[na++] = "dct_y_inv";
    dims[na] = nx_;
    fwd_names[na] = "dct_x_fwd";
    inv_names[na++] = "dct_x_inv";
    int bi = 0;
    const float *src = dream;
    for (int i = 0; i < na; ++i) {
      const int n = dims[i];
      const PassMatrix &m = mats_.get(MC::kDctForward, n, n, st);
      const bool last = (i == na - 1);
      const long long written =
          contract_pass(st, m, src, bufs[bi], N_, nv_, nv_, nv_ / n,
                        i == 0 ? x : nullptr, nullptr, last ? EPI_FILTER : EPI_STORE,
                        last ? &f : nullptr, timer, fwd_names[i]);
      if (last) nparts = written;
      src = bufs[bi];
      bi ^= 1;
    }
    sum_partials();
    for (int i = 0; i < na; ++i) {
      const int n = dims[i];
      const PassMatrix &m = mats_.get(MC::kDctInverse, n, n, st);
      const bool last = (i == na - 1);
      float *dst = last ? out : bufs[bi];
      contract_pass(st, m, src, dst, N_, nv_, nv_, nv_ / n, nullptr,
                    last ? x : nullptr, last ? EPI_ADDX : EPI_STORE, nullptr, timer,
                    inv_names[i]);
      src = dst;
      bi ^= 1;
    }
  }
  if (opt.sample && opt.sample_gamma) {
    hipLaunchKernelGGL(k_draw_nu, dim3(1), dim3(64), 0, st,
                       scal_.get() + 1, gamma_.get(), N_, (double)nv_,
                       rho_[0] * rho_[1] * rho_[2], eta_[0], eta_[1], eta_[2],
                       kNuAlpha0, kNuBeta0, opt.key0, opt.key1, opt.keyx, opt.sweep);
    HIP_LAUNCH_OK();
  }
}

void Smoother::plan_em(const float *dream, const float *x, float *out, double *prior_out,
                       EmSmoothPlan &P, hipStream_t st) {
  BINDER_REQUIRE(fused_x_supported(nx_) && nz_ > 1, "EM plan needs the 3D fused-x schedule");
  using MC = MatrixCache;
  FilterArgs f{};
  f.D = D_.get();
  f.gamma = gamma_.get();
  f.sigma2 = sigma_ * sigma_;
  const double rp = rho_[0] * rho_[1] * rho_[2];
  for (int i = 0; i < 3; ++i) f.etarho[i] = eta_[i] * eta_[i] * rp;
  f.cost_part = part_cost_.get();
  f.raw_part = part_raw_.get();
  const PassMatrix &mx = mats_.get(MC::kDctFusedX, nx_, nx_, st);
  const PassMatrix &fy = mats_.get(MC::kDctForward, ny_, ny_, st);
  const PassMatrix &iy = mats_.get(MC::kDctInverse, ny_, ny_, st);
  const PassMatrix &fz = mats_.get(MC::kDctForward, nz_, nz_, st);
  const PassMatrix &iz = mats_.get(MC::kDctInverse, nz_, nz_, st);
  const bool gen = index_grid_ && x != nullptr;
  // z forward: row = x * Ny + y, k = z; y inverse: row = x * Nz + z, k' = y
  const GridCoords gz{2, {ny_, 1, 1}, {0, ny_, 0}};
  const GridCoords gy{1, {nz_, 1, 1}, {0, 0, nz_}};
  float *b0 = bufA_.get(), *b1 = bufB_.get();
  P.z_fwd = plan_contract(fz, dream, b0, N_, nv_, nv_, nv_ / nz_, x, nullptr, EPI_STORE, nullptr,
                          ny_, gen ? &gz : nullptr);
  P.y_fwd = plan_contract(fy, b0, b1, N_, nv_, nv_, nv_ / ny_, nullptr, nullptr, EPI_STORE,
                          nullptr, nz_, nullptr);
  P.x = plan_fused(mx, b1, b0, N_, nv_, (long long)ny_ * nz_, f);
  P.sum = SumArgs{part_cost_.get(), (long long)P.x.blocks, prior_out};
  P.z_inv = plan_contract(iz, b0, b1, N_, nv_, nv_, nv_ / nz_, nullptr, nullptr, EPI_STORE,
                          nullptr, ny_, nullptr);
  P.y_inv = plan_contract(iy, b1, out, N_, nv_, nv_, nv_ / ny_, nullptr, x, EPI_ADDX, nullptr,
                          nz_, gen ? &gy : nullptr);
}

// Whether a contraction of length n over R rows (3 components) runs the
// non-persistent kernel (the rule of plan_contract).
static bool small_pass(int n, long long R, int comps) {
  const int ks = (n + 31) / 32;
  const size_t mat_lds = (size_t)((n + 15) / 16) * ks * 64 * 16 * 2;
  const long long tiles = ((R + 15) / 16) * comps;
  return !(ks <= 6 && mat_lds <= kPersistMaxLds && tiles >= 2048);
}

bool Smoother::batchable_em() const {
  return fused_x_supported(nx_) && nz_ > 1 && nz_ <= 192 && ny_ <= 192 &&
         small_pass(nz_, nv_ / nz_, N_) && small_pass(ny_, nv_ / ny_, N_);
}

// ---------------------------------------------------------------------------
// SmoothBatch: the EM smoothing of several smoothers, pass by pass, one
// launch per pass and kernel variant.  Each smoother's pass runs as the
// same grid of blocks it runs alone (a block table maps the batched grid to
// (smoother, block)), so its outputs and cost partials are those of its own
// run().
// ---------------------------------------------------------------------------
struct SmoothBatch::Launch {
  int kind = 0;  // 0 contraction, 1 fused x, 2 cost sums
  int ks = 0, pro = 0, epi = 0, nt = 0, comps = 0;
  bool vec = false, exact = false;
  int n = 0, blocks = 0;
  size_t lds = 0, args_off = 0, prefix_off = 0;
};

SmoothBatch::SmoothBatch() = default;
SmoothBatch::~SmoothBatch() {
  if (pin_) (void)hipHostFree(pin_);
}

void SmoothBatch::build(const std::vector<Job> &jobs, hipStream_t stream) {
  launches_.clear();
  std::vector<EmSmoothPlan> plans(jobs.size());
  for (size_t j = 0; j < jobs.size(); ++j) {
    BINDER_REQUIRE(jobs[j].sm->batchable_em(), "smoother cannot join a batched EM smoothing");
    jobs[j].sm->plan_em(jobs[j].dream, jobs[j].x, jobs[j].out, jobs[j].prior_out, plans[j],
                        stream);
  }
  std::vector<unsigned char> hb;
  auto append = [&](const void *p, size_t n) {
    const size_t off = (hb.size() + 255) & ~(size_t)255;
    hb.resize(off + n);
    std::memcpy(hb.data() + off, p, n);
    return off;
  };
  auto contract_group = [&](PassPlan EmSmoothPlan::*member) {
    std::vector<int> done(jobs.size(), 0);
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (done[j]) continue;
      const PassPlan &r = plans[j].*member;
      std::vector<PassArgs> args;
      std::vector<int> prefix{0};
      for (size_t i = j; i < jobs.size(); ++i) {
        const PassPlan &q = plans[i].*member;
        if (done[i] || q.a.ks != r.a.ks || q.vec != r.vec || q.pro != r.pro || q.epi != r.epi)
          continue;
        BINDER_REQUIRE(!q.persist, "batched contraction of a persistent pass");
        done[i] = 1;
        args.push_back(q.a);
        prefix.push_back(prefix.back() + q.blocks);
      }
      Launch L;
      L.kind = 0;
      L.ks = r.a.ks;
      L.vec = r.vec;
      L.pro = r.pro;
      L.epi = r.epi;
      L.n = (int)args.size();
      L.blocks = prefix.back();
      L.args_off = append(args.data(), sizeof(PassArgs) * args.size());
      L.prefix_off = append(prefix.data(), sizeof(int) * prefix.size());
      launches_.push_back(L);
    }
  };
  contract_group(&EmSmoothPlan::z_fwd);
  contract_group(&EmSmoothPlan::y_fwd);
  {
    std::vector<int> done(jobs.size(), 0);
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (done[j]) continue;
      const FusedPlan &r = plans[j].x;
      std::vector<FusedArgs> args;
      std::vector<int> prefix{0};
      size_t lds = 0;
      for (size_t i = j; i < jobs.size(); ++i) {
        const FusedPlan &q = plans[i].x;
        if (done[i] || q.nt != r.nt || q.exact != r.exact || q.comps != r.comps) continue;
        done[i] = 1;
        args.push_back(q.a);
        prefix.push_back(prefix.back() + q.blocks);
        lds = std::max(lds, q.lds);
      }
      Launch L;
      L.kind = 1;
      L.nt = r.nt;
      L.exact = r.exact;
      L.comps = r.comps;
      L.lds = lds;
      L.n = (int)args.size();
      L.blocks = prefix.back();
      L.args_off = append(args.data(), sizeof(FusedArgs) * args.size());
      L.prefix_off = append(prefix.data(), sizeof(int) * prefix.size());
      launches_.push_back(L);
    }
  }
  {
    std::vector<SumArgs> args;
    for (auto &p : plans)
      if (p.sum.out) args.push_back(p.sum);
    if (!args.empty()) {
      Launch L;
      L.kind = 2;
      L.n = (int)args.size();
      L.blocks = L.n;
      L.args_off = append(args.data(), sizeof(SumArgs) * args.size());
      launches_.push_back(L);
    }
  }
  contract_group(&EmSmoothPlan::z_inv);
  contract_group(&EmSmoothPlan::y_inv);
  if (pin_cap_ < hb.size()) {
    if (pin_) HIP_OK(hipHostFree(pin_));
    pin_ = nullptr;
    HIP_OK(hipHostMalloc((void **)&pin_, hb.size(), 0));
    pin_cap_ = hb.size();
  }
  std::memcpy(pin_, hb.data(), hb.size());
  if (dev_.n < hb.size()) dev_.alloc(hb.size());
  HIP_OK(hipMemcpyAsync(dev_.get(), pin_, hb.size(), hipMemcpyHostToDevice, stream));
}

namespace {

template <int KS, bool VEC, int PRO, int EPI>
void launch_cb(int blocks, hipStream_t s, const PassArgs *a, const int *prefix, int n) {
  hipLaunchKernelGGL((k_contract_batch<KS, VEC, PRO, EPI>), dim3(blocks), dim3(256), 0, s, a,
                     prefix, n);
}

template <bool VEC, int PRO, int EPI>
void launch_cb_ks(int ks, int blocks, hipStream_t s, const PassArgs *a, const int *prefix,
                  int n) {
  switch (ks) {
    case 1: launch_cb<1, VEC, PRO, EPI>(blocks, s, a, prefix, n); break;
    case 2: launch_cb<2, VEC, PRO, EPI>(blocks, s, a, prefix, n); break;
    case 3: launch_cb<3, VEC, PRO, EPI>(blocks, s, a, prefix, n); break;
    case 4: launch_cb<4, VEC, PRO, EPI>(blocks, s, a, prefix, n); break;
    case 5: launch_cb<5, VEC, PRO, EPI>(blocks, s, a, prefix, n); break;
    default: launch_cb<6, VEC, PRO, EPI>(blocks, s, a, prefix, n); break;
  }
}

template <int NT, bool EXACT>
void launch_fb(const SmoothBatchLaunchView &v, hipStream_t s) {
  auto k = k_dct_x_fused_batch<NT, EXACT>;
  static std::atomic<bool> attr_set{false};
  if (!attr_set) {
    HIP_OK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kPersistMaxLds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k, dim3(v.blocks), dim3(kFusedThreads), v.lds, s,
                     reinterpret_cast<const FusedArgs *>(v.args),
                     reinterpret_cast<const int *>(v.prefix), v.n, v.comps);
}

}  // namespace

void SmoothBatch::launch(hipStream_t s) const {
  const unsigned char *dev = dev_.get();
  for (const Launch &L : launches_) {
    if (L.blocks <= 0) continue;
    if (L.kind == 0) {
      const auto *a = reinterpret_cast<const PassArgs *>(dev + L.args_off);
      const auto *pf = reinterpret_cast<const int *>(dev + L.prefix_off);
#define BINDER_CB(V, PR, E) launch_cb_ks<V, PR, E>(L.ks, L.blocks, s, a, pf, L.n)
#define BINDER_CBV(V)                                                           \
  if (L.pro == PRO_X) BINDER_CB(V, PRO_X, EPI_STORE);                           \
  else if (L.pro == PRO_GRID) BINDER_CB(V, PRO_GRID, EPI_STORE);                \
  else if (L.epi == EPI_ADDX) BINDER_CB(V, PRO_NONE, EPI_ADDX);                 \
  else if (L.epi == EPI_ADDGRID) BINDER_CB(V, PRO_NONE, EPI_ADDGRID);           \
  else BINDER_CB(V, PRO_NONE, EPI_STORE);
      BINDER_REQUIRE(L.epi != EPI_FILTER, "batched contraction without filter epilogue");
      if (L.vec) {
        BINDER_CBV(true)
      } else {
        BINDER_CBV(false)
      }
#undef BINDER_CBV
#undef BINDER_CB
    } else if (L.kind == 1) {
      const SmoothBatchLaunchView v{dev + L.args_off, dev + L.prefix_off, L.n, L.blocks, L.lds,
                                    L.comps};
#define BINDER_FB(NT)                                       \
  case NT:                                                  \
    if (L.exact) launch_fb<NT, true>(v, s);                 \
    else launch_fb<NT, false>(v, s);                        \
    break;
      switch (L.nt) {
        BINDER_FB(1)
        BINDER_FB(2)
        BINDER_FB(3)
        BINDER_FB(4)
        BINDER_FB(5)
        default:
          if (L.exact) launch_fb<6, true>(v, s);
          else launch_fb<6, false>(v, s);
      }
#undef BINDER_FB
    } else {
      hipLaunchKernelGGL(k_sum_partials_batch, dim3(L.n), dim3(256), 0, s,
                         reinterpret_cast<const SumArgs *>(dev + L.args_off));
    }
    HIP_LAUNCH_OK();
  }
}

// DCT-domain zero-pad resampling of an N-component field
// (NonLinearTransformation.cpp:366-430): orthonormal DCT-II at the input size,
// then DCT-III at the output size from the first n_in coefficients per axis.
void up_sample_field(const float *in_soa, int nx, int ny, int nz, int nxo,
                     int nyo, int nzo, float *out_soa, hipStream_t stream,
                     MatrixCache &mats) {
  const int N = nzo == 1 ? 2 : 3;
  const long long nvi = (long long)nx * ny * nz, nvo = (long long)nxo * nyo * nzo;
  // largest intermediate: the inverse passes change one axis at a time
  long long big = nvi;
  for (long long v : {(long long)nx * ny * nzo, (long long)nx * nyo * nzo, nvo})
    big = v > big ? v : big;
  DBuf<float> A, B;
  A.alloc((size_t)N * big);
  B.alloc((size_t)N * big);
  // forward: [X][Y][Z] -> [Z][X][Y] -> [Y][Z][X] -> [X][Y][Z] (coefficients)
  int din[3], dout[3], na = 0;
  if (nz > 1 || nzo > 1) {
    din[na] = nz;
    dout[na++] = nzo;
  }
  din[na] = ny;
  dout[na++] = nyo;
  din[na] = nx;
  dout[na++] = nxo;
  const float *src = in_soa;
  float *bufs[2] = {A.get(), B.get()};
  int bi = 0;
  for (int i = 0; i < na; ++i) {
    const PassMatrix &m = mats.get(MatrixCache::kDctForward, din[i], din[i], stream);
    contract_pass(stream, m, src, bufs[bi], N, nvi, nvi, nvi / din[i], nullptr,
                  nullptr, EPI_STORE, nullptr, nullptr, "upsample_fwd");
    src = bufs[bi];
    bi ^= 1;
  }
  // inverse at the output size: each pass maps one axis n_in -> n_out
  long long cur = nvi;  // elements per component of the current layout
  for (int i = 0; i < na; ++i) {
    const PassMatrix &m = mats.get(MatrixCache::kDctInverse, din[i], dout[i], stream);
    const long long R = cur / din[i];
    const long long next = R * dout[i];
    float *dst = (i == na - 1) ? out_soa : bufs[bi];
    contract_pass(stream, m, src, dst, N, cur, (i == na - 1) ? nvo : next, R, nullptr,
                  nullptr, EPI_STORE, nullptr, nullptr, "upsample_inv");
    src = dst;
    bi ^= 1;
    cur = next;
  }
  HIP_OK(hipStreamSynchronize(stream));
}

}  // namespace binder
