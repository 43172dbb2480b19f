// E-step kernels for the EM and Gibbs paths.
//
// One thread per fixed-image voxel i (grid-stride).  From the current
// position d_i it forms the cubic B-spline prior over the 4x4x4 moving-grid
// nodes around floor(d_i), weights each node by the likelihood
// theta[v_j, u_i] (v_j = moving bin of node j, u_i = fixed bin of voxel i;
// nodes outside the moving grid use moving bin 0), and
//   EM:      writes the posterior-mean ("dream") location, adds -log Z_i to
//            the block's mll and the posteriors w_ij / Z_i to the block's
//            joint histogram;
//   sampling: draws one node from the posterior with a Philox uniform, writes
//            its grid coordinates and counts it in the block's histogram.
// Reference: BINDER/src/cpu/Registration.cpp:219-330 (voxel),
// :153-165 / :558-571 (loops), BINDER/src/cpu/MILikelihood.cpp:42-55 (theta
// lookup, out-of-range -> bin 0), :89-124 (histogram), :71-86 (MAP theta),
// :132-169 (Dirichlet draw), :183-206 (Dirichlet prior);
// BINDER/src/cpu/Spline.cpp:44-73 (cubic weights).
//
// Moving bins are read as packed words: for 3D, word (x,y,z) of a padded grid
// holds the bins of z..z+3 (bin 0 outside the image), so one 32-bit load
// serves a whole 4-node z-row and no bounds test is needed after clamping.
// The likelihood column theta[:, u] and the block histogram live in LDS; the
// histogram is a 64-bit fixed-point array (exact, order-independent sums).
// EM histogram, K <= 32 (k_estep_em32): the posterior mass of moving bin v
// in voxel i is theta[v, u_i] * S_v / Z_i, where S_v is the spline weight of
// the neighbours whose bin is v.  Each lane adds its nodes' spline weights,
// as 32-bit fixed point (scale 2^31, truncated, S_v <= 1), into a private per-bin LDS
// column laid out bin-major / lane-minor (bank = lane: the integer adds never
// conflict), marks the bins it touched in a 32-bit mask, and after the voxel
// flushes only those bins (about 6 of 32 on noisy images, where runs of equal
// neighbouring bins are short) into the block histogram.  Larger K uses
// k_estep, which merges runs of equal bins in registers before each LDS
// atomic.  The sampling E-step counts its one drawn
// node per voxel in a 32-bit LDS count array.
#include <atomic>

#include "estep.hpp"

namespace binder {
namespace {

__device__ __forceinline__ void spline3(float d, float w[4]) {
  const float d2 = d * d, d3 = d2 * d;
  const float o = 1.0f - d, o2 = o * o, o3 = o2 * o;
  w[0] = o3 * (1.0f / 6.0f);
  w[1] = (2.0f / 3.0f) - d2 + d3 * 0.5f;
  w[2] = (2.0f / 3.0f) - o2 + o3 * 0.5f;
  w[3] = d3 * (1.0f / 6.0f);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// floor() of a position, kept in a range where int arithmetic is safe; NaN
// maps to the padding and keeps NaN weights.
__device__ __forceinline__ int safe_floor(float p, float &frac) {
  float f = floorf(p);
  frac = p - f;
  f = fminf(fmaxf(f, -1.0e7f), 1.0e7f);
  return (int)f;
}

// Node index and B-spline weights of one axis at the node offsets -1..2 the
// E-step kernels visit (Spline.cpp:44-108, Registration.cpp:36-52): cubic
// on floor(p) (weights at -1..2); quadratic on floor(p + 1/2) (-1..1, the
// slot at +2 weighs 0); linear on floor(p) (0 and 1, the slots at -1 and +2
// weigh 0).  Zero-weight slots add nothing to any sum and are never drawn.
// CUBIC: order 3 only, compiled without the other orders' branches (the hot
// kernels' default instantiation); otherwise `order` (1, 2, or 3 / 0 for
// cubic) is read at run time.
template <bool CUBIC>
__device__ __forceinline__ int spline_node(float p, int order, float w[4]) {
  float d;
  if constexpr (!CUBIC) {
    if (order == 2) {
      const float f = fminf(fmaxf(floorf(p + 0.5f), -1.0e7f), 1.0e7f);
      d = p - f;
      const float a = fmaf(-2.f, d, 1.f), b = fmaf(-2.f, d, -1.f);
      w[0] = 0.125f * (a * a);
      w[1] = 0.75f - d * d;
      w[2] = 0.125f * (b * b);
      w[3] = 0.f;
      return (int)f;
    }
    if (order == 1) {
      const int i = safe_floor(p, d);
      w[0] = 0.f;
      w[1] = 1.f - d;
      w[2] = d;
      w[3] = 0.f;
      return i;
    }
  }
  const int i = safe_floor(p, d);
  spline3(d, w);
  return i;
}

// Adds `val` at hist[idx] for every lane with idx >= 0.  Lanes that share an
// index are first summed across the wavefront (one LDS atomic per distinct
// index for up to two distinct indices); the rest add directly.  Must be
// called by all 64 lanes of the wavefront.
__device__ __forceinline__ void wave_hist_add(unsigned long long *hist, int idx,
                                              unsigned long long val) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int round = 0; round < 2; ++round) {
    const unsigned long long pend = __ballot(idx >= 0);
    if (pend == 0ull) return;
    const int first = __ffsll((long long)pend) - 1;
    const int leader = __shfl(idx, first);
    const bool mine = (idx == leader);
    unsigned long long part = mine ? val : 0ull;
    uint32_t lo = (uint32_t)part, hi = (uint32_t)(part >> 32);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t olo = __shfl_xor(lo, off), ohi = __shfl_xor(hi, off);
      const uint32_t nlo = lo + olo;
      hi = hi + ohi + (nlo < lo ? 1u : 0u);
      lo = nlo;
    }
    if (lane == first) atomicAdd(hist + leader, ((unsigned long long)hi << 32) | lo);
    if (mine) idx = -1;
  }
  if (idx >= 0) atomicAdd(hist + idx, val);
}

__device__ __forceinline__ unsigned long long to_fixed(float v) {
  return (unsigned long long)__float2uint_rn(v * 2147483648.0f);
}

__device__ __forceinline__ float sel4(const float w[4], int k) {
  return k == 0 ? w[0] : (k == 1 ? w[1] : (k == 2 ? w[2] : w[3]));
}

// Packed moving bins of the 4 (x = ix-1+xa, y = iy-1..iy+2) columns: 4 z-bins
// per word (3D) or the single bin (2D).
template <bool IS2D>
__device__ __forceinline__ void load_columns(const EStepArgs &a, int ix, int iy,
                                             int iz, uint32_t q[16]) {
#pragma unroll
  for (int xa = 0; xa < 4; ++xa) {
    const int px = clampi(ix - 1 + xa, -1, a.nxn) + 1;
#pragma unroll
    for (int yb = 0; yb < 4; ++yb) {
      const int py = clampi(iy - 1 + yb, -1, a.nyn) + 1;
      if (IS2D) {
        q[xa * 4 + yb] = a.bins2d[(long long)px * a.Yp + py];
      } else {
        const int pz = clampi(iz - 1, -4, a.nzn) + 4;
        q[xa * 4 + yb] = a.quads[((long long)px * a.Yp + py) * a.Zq + pz];
      }
    }
  }
}

// Same words as load_columns, with 32-bit 24-bit-multiply indexing (the
// host guarantees (Ny+2)(Nz+5) < 2^24 and the padded grid < 2^31 words).
// Element `idx` of a device array addressed with a 32-bit byte offset from
// the (wavefront-uniform) base, so the load takes the scalar-base form; the
// host bounds every array read this way below 2^32 bytes.
template <typename T>
__device__ __forceinline__ T ld_off32(const T *base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_off32(T *base, uint32_t idx, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T)) = v;
}

template <bool IS2D>
__device__ __forceinline__ void load_columns32(const EStepArgs &a, int ix, int iy,
                                               int iz, uint32_t q[16]) {
  const uint32_t SX = IS2D ? (uint32_t)a.Yp : (uint32_t)a.Yp * (uint32_t)a.Zq;
  const uint32_t SY = IS2D ? 1u : (uint32_t)a.Zq;
  const uint32_t pz = IS2D ? 0u : (uint32_t)(clampi(iz - 1, -4, a.nzn) + 4);
  uint32_t ox[4], oy[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ox[k] = __umul24((uint32_t)(clampi(ix - 1 + k, -1, a.nxn) + 1), SX);
    oy[k] = __umul24((uint32_t)(clampi(iy - 1 + k, -1, a.nyn) + 1), SY) + pz;
  }
#pragma unroll
  for (int xa = 0; xa < 4; ++xa)
#pragma unroll
    for (int yb = 0; yb < 4; ++yb) {
      const uint32_t idx = ox[xa] + oy[yb];
      q[xa * 4 + yb] = IS2D ? (uint32_t)ld_off32(a.bins2d, idx) : ld_off32(a.quads, idx);
    }
}

// Uniform in (0, 1): the 24 high bits of a Philox word, centred in their
// interval, so a draw never lands exactly on 0 (a zero-weight node) or 1.
__device__ __forceinline__ float u01_mid(uint32_t v) {
  return ((float)(v >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// q[xa*4 + yb] for a runtime slice index xa, without dynamic register indexing
__device__ __forceinline__ void slice_columns(const uint32_t q[16], int xa,
                                              uint32_t s[4]) {
#pragma unroll
  for (int yb = 0; yb < 4; ++yb)
    s[yb] = xa == 0 ? q[yb] : (xa == 1 ? q[4 + yb] : (xa == 2 ? q[8 + yb] : q[12 + yb]));
}

// SSD likelihood of one voxel (fixed bin u) as node weights relative to its
// closest node: with d_v = a - s v (a = u - shift, s = scale; a = u, s = 1
// without shift and scale), r the positive-weight node bin of smallest |d|
// and delta = v - r, ll(v) - ll(r) = h (d_r^2 - d_v^2) =
// delta (2 h s d_r - h s^2 delta), h = 1/(2 sigma^2).  The per-voxel
// coefficients come from d_r in FP64, so the exponent's rounding is relative
// to the exponent itself, not to ll(v) (which can be ~1e9 when sigma^2 is
// small), and d_r carries no cancellation of a - s r.  The exponent is at
// most 0 up to rounding and is clamped there, so a zero-weight node closer
// than r adds exactly 0.  Both coefficients carry log2(e) for the base-2
// exponential.
constexpr double kLog2e = 1.4426950408889634074;
struct SsdParams {
  float s;                  // scale
  double shift, scale, h, logc;
};
__device__ __forceinline__ SsdParams ssd_params(const EStepArgs &a) {
  SsdParams p;
  const double s2 = a.ssd_par[0];
  p.shift = a.ssd_ss ? a.ssd_par[1] : 0.0;
  p.scale = a.ssd_ss ? a.ssd_par[2] : 1.0;
  p.h = 0.5 / s2;
  p.logc = -log(sqrt(2.0 * M_PI * s2));  // log_constant (SSDLikelihood.cpp:57)
  p.s = (float)p.scale;
  return p;
}
struct SsdVoxel {
  float a2, b2;  // 2 h s d_r log2(e), h s^2 log2(e)
  int r;
  __device__ __forceinline__ float w(int v) const {
    const float dl = (float)(v - r);
    return __builtin_amdgcn_exp2f(fminf(dl * fmaf(-b2, dl, a2), 0.f));
  }
};
// r over the nodes with positive spline weight (columns c = 4 xa + yb of
// packed bins q, z weights wz); also returns ll(r) / ln 2 in FP64.
template <int NZ>
__device__ __forceinline__ SsdVoxel ssd_voxel(const SsdParams &p, int u, const float wx[4],
                                              const float wy[4], const float wz[4],
                                              const uint32_t q[16], double &llr2) {
  const double ad = (double)u - p.shift;
  const float av = (float)ad;
  float best = INFINITY;
  int r = 0;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const bool col = wx[c >> 2] * wy[c & 3] > 0.f;
#pragma unroll
    for (int zc = 0; zc < NZ; ++zc) {
      const int v = (q[c] >> (8 * zc)) & 0xFF;
      const float dist = fabsf(fmaf(-p.s, (float)v, av));
      const bool take = col && wz[zc] > 0.f && dist < best;
      best = take ? dist : best;
      r = take ? v : r;
    }
  }
  const double dr = ad - p.scale * (double)r;
  llr2 = (p.logc - p.h * dr * dr) * kLog2e;  // / ln 2
  return SsdVoxel{(float)(2.0 * p.h * p.scale * dr * kLog2e),
                  (float)(p.h * p.scale * p.scale * kLog2e), r};
}

// The 64 (3D) or 16 (2D) neighbour weights are evaluated twice -- once for
// the normaliser and the mean, once for the histogram / the draw -- one
// x-slice (16 or 4 nodes) at a time, so only a slice's packed bins are live.
template <bool IS2D, bool SAMPLE, bool SSD>
__global__ __launch_bounds__(256) void k_estep(EStepArgs a) {
  // Philox key, with the per-chain device mask when given (captured sweeps)
  const uint32_t key0 = a.keyx ? a.k0 ^ a.keyx[0] : a.k0;
  const uint32_t key1 = a.keyx ? a.k1 ^ a.keyx[1] : a.k1;
  extern __shared__ __align__(16) unsigned char smem[];
  const int KL = a.K * a.L;
  unsigned long long *hist = reinterpret_cast<unsigned long long *>(smem);
  // sampling: one drawn node per voxel, counted in 32 bits (the first half of
  // the same zeroed LDS range)
  unsigned int *cnt = reinterpret_cast<unsigned int *>(smem);
  // likelihood rows theta[u][:] at a padded stride (bank spread)
  float *th = reinterpret_cast<float *>(hist + KL);
  const int Kp = theta_row_stride(a.K, a.L);
  for (int i = threadIdx.x; i < KL; i += blockDim.x) {
    hist[i] = 0ull;
    if (!SSD) th[(i / a.K) * Kp + (i % a.K)] = a.thT[i];
  }
  SsdParams sp{};
  if constexpr (SSD) sp = ssd_params(a);
  __syncthreads();

  constexpr int NZ = IS2D ? 1 : 4;
  double mll = 0.0;
  // each logical block owns one contiguous voxel range (XCD-local gathers);
  // the trip count is uniform across the block, so every lane reaches
  // wave_hist_add
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const long long per =
      ((a.nv + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;
  const long long vbeg = (long long)lb * per;
  const long long vend = vbeg + per < a.nv ? vbeg + per : a.nv;
  // the next voxel's fixed bin and position are loaded while the current one
  // computes (lanes past the range use the volume's last voxel, as below)
  int pu = 0;
  float pdx = 0.f, pdy = 0.f, pdz = 0.f;
  if (vbeg < vend) {
    const long long p0 = vbeg + threadIdx.x;
    const long long pii = p0 < vend ? p0 : a.nv - 1;
    pu = a.bv[pii];
    pdx = a.d[pii];
    pdy = a.d[a.nv + pii];
    if (!IS2D) pdz = a.d[2 * a.nv + pii];
  }
  for (long long base = vbeg; base < vend; base += blockDim.x) {
    const long long i = base + threadIdx.x;
    const bool valid = i < vend;
    const long long ii = valid ? i : a.nv - 1;
    const int u = pu;
    const float *thu = th + u * Kp;
    float wx[4], wy[4], wz[4] = {1.f, 0.f, 0.f, 0.f};
    const int ix = spline_node<false>(pdx, a.order, wx);
    const int iy = spline_node<false>(pdy, a.order, wy);
    int iz = 0;
    if (!IS2D) iz = spline_node<false>(pdz, a.order, wz);

    uint32_t qall[16];
    load_columns<IS2D>(a, ix, iy, iz, qall);
    {
      const long long nx = i + blockDim.x;
      const long long nii = nx < vend ? nx : a.nv - 1;
      pu = a.bv[nii];
      pdx = a.d[nii];
      pdy = a.d[a.nv + nii];
      if (!IS2D) pdz = a.d[2 * a.nv + nii];
    }
    // likelihood of moving bin v at this voxel: the table, or the SSD weight
    SsdVoxel sv{};
    double llr2 = 0.0;
    if constexpr (SSD) sv = ssd_voxel<NZ>(sp, u, wx, wy, wz, qall, llr2);
    auto lik = [&](int v) -> float {
      if constexpr (SSD) return sv.w(v);
      else return thu[v];
    };

    // pass 1: normaliser and first moments; the sampler also keeps the 16
    // (x,y) column sums for its two-level draw, the EM path walks the
    // x-slices in a rolled loop so its second pass cannot reuse (and keep
    // live) all 64 first-pass weights
    float Z = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
    float col[16];
    if constexpr (SAMPLE) {
#pragma unroll
      for (int xa = 0; xa < 4; ++xa) {
        float zx = 0.f;
#pragma unroll
        for (int yb = 0; yb < 4; ++yb) {
          const float wxy = wx[xa] * wy[yb];
          const uint32_t qq = qall[xa * 4 + yb];
          float zy = 0.f;
#pragma unroll
          for (int zc = 0; zc < NZ; ++zc) zy += wxy * wz[zc] * lik((qq >> (8 * zc)) & 0xFF);
          col[xa * 4 + yb] = zy;
          zx += zy;
        }
        Z += zx;
      }
    } else {
#pragma unroll 1
      for (int xa = 0; xa < 4; ++xa) {
        uint32_t q[4];
        slice_columns(qall, xa, q);
        const float wxa = sel4(wx, xa);
        float zx = 0.f;
#pragma unroll
        for (int yb = 0; yb < 4; ++yb) {
          const float wxy = wxa * wy[yb];
          float zy = 0.f;
#pragma unroll
          for (int zc = 0; zc < NZ; ++zc) {
            const float wj = wxy * wz[zc] * lik((q[yb] >> (8 * zc)) & 0xFF);
            zy += wj;
            if (!IS2D) sz += wj * (float)(zc - 1);
          }
          zx += zy;
          sy += zy * (float)(yb - 1);
        }
        Z += zx;
        sx += zx * (float)(xa - 1);
      }
    }

    int hidx = -1;
    unsigned long long hval = 0ull;
    if (!SAMPLE) {
      if (valid) {
        mll += (double)__log2f(Z) + llr2;
        const float invZ = 1.0f / Z;
        a.dream[i] = (float)ix + sx * invZ;
        a.dream[a.nv + i] = (float)iy + sy * invZ;
        if (!IS2D) a.dream[2 * a.nv + i] = (float)iz + sz * invZ;
      }
      if (valid && Z > 0.f) {
        // pass 2: posteriors grouped in runs of equal moving bin
        const float invZ = 1.0f / Z;
        int cur = -1;
        float acc = 0.f;
#pragma unroll 1
        for (int xa = 0; xa < 4; ++xa) {
          uint32_t q[4];
          slice_columns(qall, xa, q);
          const float wxa = sel4(wx, xa);
#pragma unroll
          for (int yb = 0; yb < 4; ++yb) {
            const float wxy = wxa * wy[yb];
#pragma unroll
            for (int zc = 0; zc < NZ; ++zc) {
              const int v = (q[yb] >> (8 * zc)) & 0xFF;
              const float wj = wxy * wz[zc] * lik(v);
              if (v == cur) {
                acc += wj;
              } else {
                if (cur >= 0) {
                  const unsigned long long f = to_fixed(acc * invZ);
                  if (f) atomicAdd(hist + cur * a.L + u, f);
                }
                cur = v;
                acc = wj;
              }
            }
          }
        }
        hidx = cur * a.L + u;
        hval = to_fixed(acc * invZ);
      }
    } else {
      // categorical draw in the reference's node order (x, then y, then z):
      // the first node whose running sum reaches r*Z.  The (x,y) column is
      // found from the column sums, then the node within the column.
      const u32x4 rr = philox4x32_10(
          u32x4{(uint32_t)ii, (uint32_t)(ii >> 32), a.sweep, kStreamEStep}, key0,
          key1);
      const float target = u01f(rr.x) * Z;
      int scol = -1, lastcol = -1;
      float before = 0.f, cum = 0.f, before_last = 0.f;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const float nc = cum + col[c];
        if (col[c] > 0.f) {
          lastcol = c;
          before_last = cum;
        }
        if (scol < 0 && target <= nc && col[c] > 0.f) {
          scol = c;
          before = cum;
        }
        cum = nc;
      }
      // rounding of the running sum against Z can leave target just above
      // the final cumulative value: the draw then falls in the last column
      // with positive weight
      if (scol < 0) {
        scol = lastcol;
        before = before_last;
      }
      int sel = -1, selv = 0;
      if (scol >= 0) {
        const int xa = scol >> 2, yb = scol & 3;
        uint32_t qsel = 0;
        float wxy = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c)
          if (c == scol) {
            qsel = qall[c];
            wxy = wx[c >> 2] * wy[c & 3];
          }
        float cz = before;
        int lastz = -1, lastv = 0;
#pragma unroll
        for (int zc = 0; zc < NZ; ++zc) {
          const int v = (qsel >> (8 * zc)) & 0xFF;
          const float wj = wxy * wz[zc] * lik(v);
          cz += wj;
          if (wj > 0.f) {
            lastz = zc;
            lastv = v;
          }
          if (sel < 0 && wj > 0.f && target <= cz) {
            sel = (xa * 4 + yb) * NZ + zc;
            selv = v;
          }
        }
        if (sel < 0 && lastz >= 0) {
          sel = (xa * 4 + yb) * NZ + lastz;
          selv = lastv;
        }
      }
      if (valid) {
        if (Z > 0.f && sel >= 0) {
          const int xa = sel / (4 * NZ);
          const int yb = (sel / NZ) % 4;
          const int zc = sel % NZ;
          a.dream[i] = (float)(ix - 1 + xa);
          a.dream[a.nv + i] = (float)(iy - 1 + yb);
          if (!IS2D) a.dream[2 * a.nv + i] = (float)(iz - 1 + zc);
          hidx = selv * a.L + u;
          hval = (unsigned long long)kHistScale;
        } else {
          // no draw (all weights zero): the reference leaves the voxel's
          // reset dream location (0) and adds nothing to the histogram
          a.dream[i] = 0.f;
          a.dream[a.nv + i] = 0.f;
          if (!IS2D) a.dream[2 * a.nv + i] = 0.f;
        }
      }
    }
    if (hval == 0ull) hidx = -1;
    if constexpr (SAMPLE) {
      if (hidx >= 0) atomicAdd(cnt + hidx, 1u);
    } else {
      wave_hist_add(hist, hidx, hval);
    }
  }

  __syncthreads();
  unsigned long long *out = a.hpart + (long long)lb * KL;
  for (int i = threadIdx.x; i < KL; i += blockDim.x)
    out[i] = SAMPLE ? (unsigned long long)cnt[i] << kHistShift : hist[i];
  // deterministic block reduction of mll
  __shared__ double red[256];
  red[threadIdx.x] = mll;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.mllpart[lb] = -kLn2 * red[0];
}

// Sampling E-step (any K): one categorical node draw per voxel from the
// posterior w_j / Z over the 64 (3D) or 16 (2D) neighbours, in the node
// order of the reference's CDF scan (x, then y, then z;
// Registration.cpp:278-307).  Column sums col = (w_x w_y) * sum_z w_z theta
// are accumulated in that order into cum_0..cum_15, Z = cum_15; the draw
// target r Z (r in (0, 1)) selects the first column with cum >= target by
// counting columns below it, then the node inside the column with partial
// sums formed by the same operations, so the scan always ends inside the
// selected column.  LDS: theta rows (padded stride) | 32-bit counts [K*L].
template <bool IS2D, bool SSD, bool CUBIC>
__global__ __launch_bounds__(256) void k_esample(EStepArgs a) {
  // Philox key, with the per-chain device mask when given (captured sweeps)
  const uint32_t key0 = a.keyx ? a.k0 ^ a.keyx[0] : a.k0;
  const uint32_t key1 = a.keyx ? a.k1 ^ a.keyx[1] : a.k1;
  extern __shared__ __align__(16) unsigned char smem[];
  const int KL = a.K * a.L;
  const int Kp = theta_row_stride(a.K, a.L);
  float *th = reinterpret_cast<float *>(smem);
  unsigned int *cnt = reinterpret_cast<unsigned int *>(th + ((a.L * Kp + 3) & ~3));
  for (int i = threadIdx.x; i < KL; i += blockDim.x) {
    cnt[i] = 0u;
    if (!SSD) th[(i / a.K) * Kp + (i % a.K)] = a.thT[i];
  }
  SsdParams sp{};
  if constexpr (SSD) sp = ssd_params(a);
  __syncthreads();

  constexpr int NZ = IS2D ? 1 : 4;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  // Voxels go in rows of 64, one row per wavefront step (lane l takes voxel
  // 64 row + l).  Four consecutive rows 4c..4c+3 share one Philox block per
  // lane, counter 64 c + l, whose word (row & 3) is the voxel's uniform: the
  // stream depends on the voxel index only.  Every wavefront owns a
  // contiguous run of rows, balanced to within one row, and a logical
  // block's wavefronts own adjacent runs (XCD-local gathers).
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const long long nrows = (a.nv + 63) >> 6;
  const long long wtotal = (long long)gridDim.x * nwaves;
  const long long wid = (long long)lb * nwaves + wave;
  const long long rbeg = wid * nrows / wtotal, rend = (wid + 1) * nrows / wtotal;
  u32x4 rr{0u, 0u, 0u, 0u};
  // 32-bit offsets: the host keeps N * nv below 2^30.  The next row's fixed
  // bin and position are loaded while the current row computes (indices past
  // the volume reload its last voxel; their values are not used).
  const uint32_t nv32 = (uint32_t)a.nv, vlast = nv32 - 1u;
  uint32_t pi = (uint32_t)(rbeg * 64 + lane);
  pi = pi < vlast ? pi : vlast;
  int pu = ld_off32(a.bv, pi);
  float pdx = ld_off32(a.d, pi), pdy = ld_off32(a.d, nv32 + pi), pdz = 0.f;
  if (!IS2D) pdz = ld_off32(a.d, 2 * nv32 + pi);
  for (long long row = rbeg; row < rend; ++row) {
    {
      const int qv = (int)(row & 3);
      if (qv == 0 || row == rbeg) {  // wavefront-uniform
        const long long g = (row >> 2) * 64 + lane;
        rr = philox4x32_10(u32x4{(uint32_t)g, (uint32_t)(g >> 32), a.sweep, kStreamEStep},
                           key0, key1);
        for (int k = 0; k < qv; ++k) rr = u32x4{rr.y, rr.z, rr.w, rr.x};
      }
      const long long i = row * 64 + lane;
      if (i >= a.nv) break;  // only the last row is partial
      const uint32_t word = rr.x;
      rr = u32x4{rr.y, rr.z, rr.w, rr.x};
      const uint32_t iu = (uint32_t)i;
      const int u = pu;
      const float *thu = th + u * Kp;
      float wx[4], wy[4], wz[4] = {1.f, 0.f, 0.f, 0.f};
      const int ix = spline_node<CUBIC>(pdx, a.order, wx);
      const int iy = spline_node<CUBIC>(pdy, a.order, wy);
      int iz = 0;
      if (!IS2D) iz = spline_node<CUBIC>(pdz, a.order, wz);
      uint32_t q[16];
      load_columns32<IS2D>(a, ix, iy, iz, q);
      {
        // the wavefront's last row reloads its own voxel (no fetch from the
        // next wavefront's run)
        uint32_t ni = row + 1 < rend ? iu + 64u : iu;
        ni = ni < vlast ? ni : vlast;
        pu = ld_off32(a.bv, ni);
        pdx = ld_off32(a.d, ni);
        pdy = ld_off32(a.d, nv32 + ni);
        if (!IS2D) pdz = ld_off32(a.d, 2 * nv32 + ni);
      }
      SsdVoxel sv{};
      if constexpr (SSD) {
        double llr2;
        sv = ssd_voxel<NZ>(sp, u, wx, wy, wz, q, llr2);
      }
      auto lik = [&](int v) -> float {
        if constexpr (SSD) return sv.w(v);
        else return thu[v];
      };

      float cum[16];
      float acc = 0.f;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        float zy = wz[0] * lik(q[c] & 0xFF);
#pragma unroll
        for (int zc = 1; zc < NZ; ++zc) zy = fmaf(wz[zc], lik((q[c] >> (8 * zc)) & 0xFF), zy);
        acc = fmaf(wx[c >> 2] * wy[c & 3], zy, acc);
        cum[c] = acc;
      }
      const float Z = acc;
      const float target = u01_mid(word) * Z;
      // column: the first whose cumulative sum reaches the target.  One pass
      // over the (non-decreasing) cumulative sums counts the columns below the
      // target and carries the last sum below it and the next column's word
      // (running selects: no indexed access to the arrays); the weight of the
      // selected column comes from its (x, y) indices afterwards
      int sel = 0;
      float before = 0.f;
      uint32_t qs = q[0];
#pragma unroll
      for (int c = 0; c < 15; ++c) {
        const bool below = cum[c] < target;
        sel += below ? 1 : 0;
        before = below ? cum[c] : before;
        qs = below ? q[c + 1] : qs;
      }
      const int xa = sel >> 2, yb = sel & 3;
      const float wxa = xa == 0 ? wx[0] : (xa == 1 ? wx[1] : (xa == 2 ? wx[2] : wx[3]));
      const float wyb = yb == 0 ? wy[0] : (yb == 1 ? wy[1] : (yb == 2 ? wy[2] : wy[3]));
      const float wxy = wxa * wyb;
      // node inside the column, partial sums with the column's own operations
      int zsel = 0;
      if (!IS2D) {
        float zy = wz[0] * lik(qs & 0xFF);
#pragma unroll
        for (int zc = 1; zc < NZ; ++zc) {
          zsel += fmaf(wxy, zy, before) < target ? 1 : 0;
          zy = fmaf(wz[zc], lik((qs >> (8 * zc)) & 0xFF), zy);
        }
      }
      if (Z > 0.f) {
        st_off32(a.dream, iu, (float)(ix - 1 + xa));
        st_off32(a.dream, nv32 + iu, (float)(iy - 1 + yb));
        if (!IS2D) st_off32(a.dream, 2 * nv32 + iu, (float)(iz - 1 + zsel));
        const int v = (qs >> (8 * zsel)) & 0xFF;
        atomicAdd(cnt + v * a.L + u, 1u);
      } else {
        // no draw (all weights zero): the reference leaves the voxel's reset
        // dream location (0) and adds nothing to the histogram
        st_off32(a.dream, iu, 0.f);
        st_off32(a.dream, nv32 + iu, 0.f);
        if (!IS2D) st_off32(a.dream, 2 * nv32 + iu, 0.f);
      }
    }
  }

  __syncthreads();
  unsigned long long *out = a.hpart + (long long)lb * KL;
  for (int i = threadIdx.x; i < KL; i += blockDim.x)
    out[i] = (unsigned long long)cnt[i] << kHistShift;
  if (threadIdx.x == 0) a.mllpart[lb] = 0.0;
}

// EM E-step for K <= 32 (same outputs as k_estep<IS2D, false>).  LDS:
// block histogram [L][K] (u64 fixed point) | theta columns | private per-lane bin
// columns priv[wave][bin][lane] (u32 fixed-point spline mass).
constexpr int kEm32Threads = 512;
constexpr int kEm32MaxK = 32;
constexpr float kSplineScale = 2147483648.0f;  // 2^31

// One logical block `lb` of `nb` of the EM E-step of registration `a`: the
// voxel range, the histogram partial row and the mll partial of that block.
// The result depends only on (a, lb, nb), never on which physical block runs
// it, so one pair's blocks give the same partials alone or in a batch.
template <bool IS2D, bool CUBIC>
__device__ __forceinline__ void estep_em32_block(const EStepArgs &a, int lb, int nb) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int KL = a.K * a.L;
  unsigned long long *hist = reinterpret_cast<unsigned long long *>(smem);
  float *th = reinterpret_cast<float *>(hist + KL);
  const int Kp = theta_row_stride(a.K, a.L);
  uint32_t *priv = reinterpret_cast<uint32_t *>(th + ((a.L * Kp + 3) & ~3));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t *mine = priv + wave * (kEm32MaxK * 64) + lane;  // bin v at mine[64 * v]
  // the z weights are kept scaled by 2^31, so each node's weight
  // w_j = B_j theta_j comes out scaled by 2^31 (exactly: a power of two), as
  // the fixed-point value the private column adds, and Z, s_x, s_y, s_z with
  // it.  (Integer private columns: LDS integer adds are several times
  // faster than FP32 ones.)
  for (int i = threadIdx.x; i < KL; i += blockDim.x) {
    hist[i] = 0ull;
    th[(i / a.K) * Kp + (i % a.K)] = a.thT[i];
  }
  for (int v = 0; v < kEm32MaxK; ++v) mine[64 * v] = 0u;
  __syncthreads();

  constexpr int NZ = IS2D ? 1 : 4;
  // sum of log2 Z_i in fixed point (round(log2 Z_i * 2^22), int64): exact
  // and order-free, so the mll does not depend on the block decomposition;
  // non-finite terms are flagged instead (bit 0: Z_i == 0, bit 1: NaN)
  long long l2fix = 0;
  int l2odd = 0;
  const long long per =
      ((a.nv + nb - 1) / nb + blockDim.x - 1) / blockDim.x * blockDim.x;
  const long long vbeg = (long long)lb * per;
  const long long vend = vbeg + per < a.nv ? vbeg + per : a.nv;
  // rotation of this lane's flush order, so lanes of one wave that share a
  // histogram row flush different bins in the same step
  const int rot = lane & 31;
  // 32-bit voxel offsets (the host keeps N * nv below 2^30) and 32-bit
  // moving-grid offsets (checked when the bins are set): no 64-bit address
  // arithmetic in the voxel loop
  const uint32_t nv32 = (uint32_t)a.nv, vend32 = (uint32_t)vend;
  // the next voxel's fixed bin and position are loaded while the current one
  // computes (indices past the block's range reload its last voxel, unused)
  const uint32_t vlast = vend32 > 0u ? vend32 - 1u : 0u;
  uint32_t pi = (uint32_t)vbeg + threadIdx.x;
  pi = pi < vlast ? pi : vlast;
  int pu = 0;
  float pdx = 0.f, pdy = 0.f, pdz = 0.f;
  if ((uint32_t)vbeg < vend32) {
    pu = ld_off32(a.bv, pi);
    pdx = ld_off32(a.d, pi);
    pdy = ld_off32(a.d, nv32 + pi);
    if (!IS2D) pdz = ld_off32(a.d, 2 * nv32 + pi);
  }
  for (uint32_t base = (uint32_t)vbeg; base < vend32; base += blockDim.x) {
    const uint32_t iu = base + threadIdx.x;
    if (iu >= vend32) break;
    const int u = pu;
    const float *thu = th + u * Kp;
    float wx[4], wy[4], wz[4] = {1.f, 0.f, 0.f, 0.f};
    const int ix = spline_node<CUBIC>(pdx, a.order, wx);
    const int iy = spline_node<CUBIC>(pdy, a.order, wy);
    int iz = 0;
    if (!IS2D) iz = spline_node<CUBIC>(pdz, a.order, wz);
    float wzs[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) wzs[c] = wz[c] * kSplineScale;
    uint32_t qall[16];
    load_columns32<IS2D>(a, ix, iy, iz, qall);
    {
      uint32_t ni = iu + blockDim.x;
      ni = ni < vlast ? ni : vlast;
      pu = ld_off32(a.bv, ni);
      pdx = ld_off32(a.d, ni);
      pdy = ld_off32(a.d, nv32 + ni);
      if (!IS2D) pdz = ld_off32(a.d, 2 * nv32 + ni);
    }

    float Z = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
    uint32_t mask = 0u;
#pragma unroll
    for (int xa = 0; xa < 4; ++xa) {
      // one x-slice: all theta reads first, then the private adds, so the
      // reads are not serialised behind the adds (LDS completes in order)
      uint32_t w[4 * NZ];
      int vv[4 * NZ];
      float zx = 0.f;
#pragma unroll
      for (int yb = 0; yb < 4; ++yb) {
        const float wxy = wx[xa] * wy[yb];
        const uint32_t qq = qall[xa * 4 + yb];
        float zy = 0.f;
#pragma unroll
        for (int zc = 0; zc < NZ; ++zc) {
          const int v = (qq >> (8 * zc)) & 0xFF;
          const float wj = (wxy * wzs[zc]) * thu[v];  // 2^31 B_j theta_j
          w[yb * NZ + zc] = (uint32_t)wj;             // truncated
          vv[yb * NZ + zc] = v;
          zy += wj;
          // z offset zc - 1 in {-1, 0, 1, 2}: no product with zero
          if (!IS2D) {
            if (zc == 0) sz -= wj;
            else if (zc == 2) sz += wj;
            else if (zc == 3) sz = fmaf(2.f, wj, sz);
          }
        }
        zx += zy;
        sy += zy * (float)(yb - 1);
      }
      Z += zx;
      sx += zx * (float)(xa - 1);
#pragma unroll
      for (int j = 0; j < 4 * NZ; ++j) {
        // private column: only this lane touches it, so the LDS adds (no
        // return value) land in program order and the column sums do not
        // depend on timing or on the block decomposition
        atomicAdd(mine + 64 * vv[j], w[j]);
        mask |= 1u << vv[j];
      }
    }

    // the sums carry the 2^31 scale: Z itself for the log, while the ratios
    // s / Z and the fixed-point masses below take the scaled values as they
    // are (identical quotients)
    const float Zs = Z;
    Z = Zs * (1.0f / kSplineScale);
    {
      // lz * 2^22 is exact in FP32 and below 2^31 in magnitude for every
      // finite log2 of a positive float: one rounding to an int32 per voxel.
      // Branch-free (a branch here makes the compiler keep the whole x-slice
      // state live across it: 247 VGPRs instead of 104)
      const float lz = __log2f(Z);
      const bool fin = fabsf(lz) <= kMllFixMax;
      l2fix += (long long)__float2int_rn(fin ? lz * kMllFixScaleF : 0.f);
      l2odd |= fin ? 0 : (lz != lz ? 2 : 1);
    }
    const float invZ = 1.0f / Zs;
    st_off32(a.dream, iu, (float)ix + sx * invZ);
    st_off32(a.dream, nv32 + iu, (float)iy + sy * invZ);
    if (!IS2D) st_off32(a.dream, 2 * nv32 + iu, (float)iz + sz * invZ);
    // fixed-point mass S_v / Z of bin v at the histogram's 2^31 scale (S_v,
    // the private column, is sum_{j: v_j = v} w_j at that scale already);
    // a voxel with Z = 0 adds zeros
    const float invZs = Zs > 0.f ? invZ * kSplineScale : 0.f;
    unsigned long long *hrow = hist + u * a.K;  // block histogram is [L][K]
    uint32_t m = __builtin_amdgcn_alignbit(mask, mask, rot);  // rotate right
    while (m) {
      const int b = __builtin_ctz(m);
      m &= m - 1u;
      const int v = (b + rot) & 31;
      const uint32_t sv = mine[64 * v];
      mine[64 * v] = 0u;
      atomicAdd(hrow + v, (unsigned long long)__float2uint_rn((float)sv * invZs));
    }
  }

  __syncthreads();
  // partial row in the [K][L] order of the reduction
  unsigned long long *out = a.hpart + (long long)lb * KL;
  for (int i = threadIdx.x; i < KL; i += blockDim.x) out[i] = hist[(i % a.L) * a.K + i / a.L];
  // the private columns (all zero again) are free now: reuse them for the
  // reduction of the fixed-point log sums and flags, so two blocks fit in
  // one CU
  long long *red = reinterpret_cast<long long *>(priv);
  int *odd = reinterpret_cast<int *>(red + kEm32Threads);
  red[threadIdx.x] = l2fix;
  odd[threadIdx.x] = l2odd;
  __syncthreads();
  for (int s2 = kEm32Threads / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) {
      red[threadIdx.x] += red[threadIdx.x + s2];
      odd[threadIdx.x] |= odd[threadIdx.x + s2];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const long long v = odd[0] & 2 ? kMllFixNaN : (odd[0] & 1 ? kMllFixZero : red[0]);
    reinterpret_cast<long long *>(a.mllpart)[lb] = v;
  }
}

template <bool IS2D, bool CUBIC>
__global__ __launch_bounds__(kEm32Threads) void k_estep_em32(EStepArgs a) {
  estep_em32_block<IS2D, CUBIC>(a, xcd_logical_block(blockIdx.x, gridDim.x), gridDim.x);
}

// The same E-step for several registrations in one launch: work item w
// (pair << 16 | logical block) runs estep_em32_block of that pair, with the
// pair's own block count.  Items are ordered pair-major, and consecutive
// items share an XCD, as one pair's logical blocks do alone.
template <bool IS2D, bool CUBIC>
__global__ __launch_bounds__(kEm32Threads) void k_estep_em32_batch(const EStepArgs *args,
                                                                 const uint32_t *work) {
  const uint32_t w = work[xcd_logical_block(blockIdx.x, gridDim.x)];
  const EStepArgs a = args[w >> 16];
  estep_em32_block<IS2D, CUBIC>(a, (int)(w & 0xFFFFu), a.nblocks);
}

constexpr int kRowsPerReduce = kHistReduceRows;

__device__ __forceinline__ void hist_reduce1_block(const unsigned long long *__restrict__ hpart,
                                                   int rows, int KL,
                                                   unsigned long long *__restrict__ hpart2,
                                                   int bx, int by) {
  const int kl = bx * 256 + threadIdx.x;
  if (kl >= KL) return;
  const int r0 = by * kRowsPerReduce;
  if (r0 >= rows) return;
  // all rows of the group in flight at once (rows past the end load the
  // last row and add zero)
  unsigned long long v[kRowsPerReduce];
#pragma unroll
  for (int r = 0; r < kRowsPerReduce; ++r)
    v[r] = hpart[(long long)(r0 + r < rows ? r0 + r : rows - 1) * KL + kl];
  unsigned long long s = 0;
#pragma unroll
  for (int r = 0; r < kRowsPerReduce; ++r) s += r0 + r < rows ? v[r] : 0ull;
  hpart2[(long long)by * KL + kl] = s;
}

__global__ __launch_bounds__(256) void k_hist_reduce1(
    const unsigned long long *__restrict__ hpart, int rows, int KL,
    unsigned long long *__restrict__ hpart2) {
  hist_reduce1_block(hpart, rows, KL, hpart2, blockIdx.x, blockIdx.y);
}

// 64 entries per block, four row subsets per entry summed in parallel and
// combined in a fixed order (integer sums: exact).
__device__ __forceinline__ void hist_reduce2_block(const unsigned long long *__restrict__ hpart2,
                                                   int rows2, int KL, double *__restrict__ Hd,
                                                   int bx) {
  __shared__ unsigned long long part[4][64];
  const int e = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const int kl = bx * 64 + e;
  unsigned long long s = 0;
  if (kl < KL) {
#pragma unroll 8
    for (int r = sub; r < rows2; r += 4) s += hpart2[(long long)r * KL + kl];
  }
  part[sub][e] = s;
  __syncthreads();
  if (sub == 0 && kl < KL)
    Hd[kl] = (double)(part[0][e] + part[1][e] + part[2][e] + part[3][e]) *
             (1.0 / kHistScale);
}

__global__ __launch_bounds__(256) void k_hist_reduce2(
    const unsigned long long *__restrict__ hpart2, int rows2, int KL,
    double *__restrict__ Hd) {
  hist_reduce2_block(hpart2, rows2, KL, Hd, blockIdx.x);
}

// Both reduction stages for several registrations (blockIdx.z = pair).
__global__ __launch_bounds__(256) void k_hist_reduce1_batch(const HistReduceArgs *args) {
  const HistReduceArgs a = args[blockIdx.z];
  hist_reduce1_block(a.hpart, a.rows, a.KL, a.hpart2, blockIdx.x, blockIdx.y);
}
__global__ __launch_bounds__(256) void k_hist_reduce2_batch(const HistReduceArgs *args) {
  const HistReduceArgs a = args[blockIdx.z];
  hist_reduce2_block(a.hpart2, (a.rows + kRowsPerReduce - 1) / kRowsPerReduce, a.KL, a.Hd,
                     blockIdx.x);
}

// Likelihood-parameter step in one block, working in LDS: sum of the mll
// partials, then the MAP update (H + alpha - 1)/row sum, or a
// Dirichlet(H + alpha) draw, or no change; then the Dirichlet prior cost and
// the transposed FP32 table.
__device__ __forceinline__ void theta_block(const ThetaArgs &a) {
  // Philox key, with the per-chain device mask when given (captured sweeps)
  const uint32_t key0 = a.keyx ? a.k0 ^ a.keyx[0] : a.k0;
  const uint32_t key1 = a.keyx ? a.k1 ^ a.keyx[1] : a.k1;
  const int KL = a.K * a.L;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int kWaves = 1024 / 64;
  __shared__ double sh[4096];   // H, then theta
  __shared__ double rowc[256];  // per-row prior cost / row sums
  __shared__ double wsum[kWaves];
  __shared__ long long wfix[kWaves];
  __shared__ int wodd[kWaves];
  // mll partials: per thread (index stride 1024), then per wavefront (a
  // fixed butterfly), then the 16 wavefronts in order by thread 0 at the end
  if (a.mll_fixed) {
    // exact int64 sum of the fixed-point partials (integer additions in any
    // order give the same sum), then one rounding
    const long long *fx = reinterpret_cast<const long long *>(a.mllpart);
    long long acc = 0;
    int odd = 0;
    for (int i = tid; i < a.nmll; i += 1024) {
      const long long v = fx[i];
      if (v == kMllFixNaN) odd |= 2;
      else if (v == kMllFixZero) odd |= 1;
      else acc += v;
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      acc += __shfl_xor(acc, m, 64);
      odd |= __shfl_xor(odd, m, 64);
    }
    if (lane == 0) {
      wfix[wid] = acc;
      wodd[wid] = odd;
    }
  } else {
    double s = 0.0;
    for (int i = tid; i < a.nmll; i += 1024) s += a.mllpart[i];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) wsum[wid] = s;
  }
  for (int kl = tid; kl < KL; kl += 1024) {
    double h;
    if (a.hsum) {
      // exact u64 sum of the partial rows (any order), then the scaling of
      // k_hist_reduce2; 16 rows in flight per step (rows past the end load
      // the last row and add zero)
      unsigned long long acc = 0;
      for (int r0 = 0; r0 < a.hrows; r0 += 16) {
        unsigned long long v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int r = r0 + j < a.hrows ? r0 + j : a.hrows - 1;
          v[j] = a.hsum[(long long)r * KL + kl];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) acc += r0 + j < a.hrows ? v[j] : 0ull;
      }
      h = (double)acc * (1.0 / kHistScale);
      a.Hd[kl] = h;
    } else {
      h = a.Hd[kl];
    }
    if (a.mode == kThetaDirichlet)
      sh[kl] = gamma_mt(h + a.alpha, u32x4{(uint32_t)kl, a.sweep, 0u, kStreamTheta},
                        key0, key1);
    else if (a.mode == kThetaMap)
      sh[kl] = h + a.alpha - 1.0;
    else
      sh[kl] = a.theta[kl];
  }
  __syncthreads();
  if (a.mode != kThetaKeep) {
    // row sums in index order (MILikelihood.cpp:74-79)
    for (int k = tid; k < a.K; k += 1024) {
      double sum = 0.0;
      for (int l = 0; l < a.L; ++l) sum += sh[k * a.L + l];
      rowc[k] = sum;
    }
    __syncthreads();
    for (int kl = tid; kl < KL; kl += 1024) sh[kl] = sh[kl] / rowc[kl / a.L];
    __syncthreads();
  }
  // -log Dir(theta_k; alpha) per row: the K*L terms (alpha - 1) log theta in
  // parallel (after theta is written out), then each row summed in index
  // order, then the rows in row order
  for (int kl = tid; kl < KL; kl += 1024) {
    const double t = sh[kl];
    a.theta[kl] = t;
    a.thT[(kl % a.L) * a.K + kl / a.L] = (float)t;
    sh[kl] = (a.alpha - 1.0) * log(t);
  }
  __syncthreads();
  for (int k = tid; k < a.K; k += 1024) {
    double lp = a.lg_sum_alpha;
    for (int l = 0; l < a.L; ++l) lp -= a.lg_alpha;
    const double *terms = sh + k * a.L;
#pragma unroll 8
    for (int l = 0; l < a.L; ++l) lp += terms[l];
    rowc[k] = -lp;
  }
  __syncthreads();
  if (tid == 0) {
    double mll;
    if (a.mll_fixed) {
      long long acc = 0;
      int odd = 0;
      for (int w = 0; w < kWaves; ++w) {
        acc += wfix[w];
        odd |= wodd[w];
      }
      mll = odd & 2 ? (double)NAN
                    : (odd & 1 ? (double)INFINITY : -kLn2 * ((double)acc / kMllFixScale));
    } else {
      mll = 0.0;
      for (int w = 0; w < kWaves; ++w) mll += wsum[w];
    }
    double c = 0.0;
    for (int k = 0; k < a.K; ++k) c += rowc[k];
    a.stats[0] = mll;
    a.stats[1] = c;
  }
}

__global__ __launch_bounds__(1024) void k_theta(ThetaArgs a) { theta_block(a); }

// The theta step of several registrations, one block each.
__global__ __launch_bounds__(1024) void k_theta_batch(const ThetaArgs *args) {
  const ThetaArgs a = args[blockIdx.x];
  theta_block(a);
}

// One sample's Welford step (Registration.cpp:593-623): d1 = x - mu_old,
// mu += d1 / n, d2 = x - mu_new, M2 += d1 d2^T.
template <int NC>
__device__ __forceinline__ void welford_step(const float *xv, double inv_n, double *mu,
                                             double *M) {
  double d1[NC], d2[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double val = (double)xv[c];
    d1[c] = val - mu[c];
    mu[c] = mu[c] + d1[c] * inv_n;
    d2[c] = val - mu[c];
  }
  int p = 0;
#pragma unroll
  for (int r = 0; r < NC; ++r)
#pragma unroll
    for (int c = r; c < NC; ++c, ++p) M[p] += d1[r] * d2[c];
}

// CNT > 0: exactly CNT samples, all read before the updates so the reads
// are in flight together (32-bit offsets: the host keeps N * nv below
// 2^30); CNT == 0: b.count samples, one at a time.
template <int NC, int CNT>
__global__ __launch_bounds__(256) void k_welford(WelfordBatch b, double *mean, double *m2,
                                                 long long nv) {
  constexpr int NP = NC * (NC + 1) / 2;
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (b.n0_base) b.n0 += *b.n0_base;
  // no sample folded in yet: the accumulators start from zero in registers
  // instead of being read (the first update then gives mean = x, M2 = 0
  // exactly, as from zeroed memory), so they need no clearing beforehand
  const bool fresh = b.n0 == 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    double mu[NC] = {}, M[NP] = {};
    if (!fresh) {
#pragma unroll
      for (int c = 0; c < NC; ++c) mu[c] = mean[c * nv + i];
#pragma unroll
      for (int p = 0; p < NP; ++p) M[p] = m2[p * nv + i];
    }
    if constexpr (CNT > 0) {
      float xs[CNT][NC];
      const uint32_t iu = (uint32_t)i, nv32 = (uint32_t)nv;
#pragma unroll
      for (int k = 0; k < CNT; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) xs[k][c] = ld_off32(b.x[k], c * nv32 + iu);
#pragma unroll
      for (int k = 0; k < CNT; ++k)
        welford_step<NC>(xs[k], 1.0 / ((double)(b.n0 + k) + 1.0), mu, M);
    } else {
      for (int k = 0; k < b.count; ++k) {
        float xv[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) xv[c] = b.x[k][c * nv + i];
        welford_step<NC>(xv, 1.0 / ((double)(b.n0 + k) + 1.0), mu, M);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) mean[c * nv + i] = mu[c];
#pragma unroll
    for (int p = 0; p < NP; ++p) m2[p * nv + i] = M[p];
  }
}

// The spline order runs the kernels' order-3-only instantiation.
inline bool cubic_order(int order) { return order == 3 || order == 0; }

}  // namespace

void launch_estep(const EStepArgs &a, bool is2d, bool sample, int blocks,
                  hipStream_t s) {
  const int Kp = theta_row_stride(a.K, a.L);
  const size_t lds = (size_t)a.K * a.L * sizeof(unsigned long long) +
                     (size_t)Kp * a.L * sizeof(float);
  dim3 g(blocks), b(256);
  if (sample) {
    const size_t ldss = (size_t)((Kp * a.L + 3) & ~3) * sizeof(float) +
                        (size_t)a.K * a.L * sizeof(unsigned int);
    // [ssd][is2d][cubic]
    static void (*const ks[2][2][2])(EStepArgs) = {
        {{k_esample<false, false, false>, k_esample<false, false, true>},
         {k_esample<true, false, false>, k_esample<true, false, true>}},
        {{k_esample<false, true, false>, k_esample<false, true, true>},
         {k_esample<true, true, false>, k_esample<true, true, true>}}};
    hipLaunchKernelGGL(ks[a.ssd_par != nullptr][is2d][cubic_order(a.order)], g, b, ldss, s, a);
    HIP_LAUNCH_OK();
    return;
  }
  if (a.K <= kEm32MaxK && !a.ssd_par) {
    const size_t lds32 = estep_em32_lds(a.K, a.L);
    static std::atomic<bool> attr_set[2][2] = {};
    static void (*const ks[2][2])(EStepArgs) = {
        {k_estep_em32<false, false>, k_estep_em32<false, true>},
        {k_estep_em32<true, false>, k_estep_em32<true, true>}};
    const bool cubic = cubic_order(a.order);
    auto k = ks[is2d][cubic];
    if (!attr_set[is2d][cubic]) {
      HIP_OK(hipFuncSetAttribute((const void *)k,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 160 * 1024));
      attr_set[is2d][cubic] = true;
    }
    hipLaunchKernelGGL(k, g, dim3(kEm32Threads), lds32, s, a);
    HIP_LAUNCH_OK();
    return;
  }
  if (a.ssd_par) {
    if (is2d) hipLaunchKernelGGL((k_estep<true, false, true>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_estep<false, false, true>), g, b, lds, s, a);
  } else {
    if (is2d) hipLaunchKernelGGL((k_estep<true, false, false>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_estep<false, false, false>), g, b, lds, s, a);
  }
  HIP_LAUNCH_OK();
}

size_t estep_em32_lds(int K, int L) {
  const int Kp = theta_row_stride(K, L);
  return (size_t)K * L * sizeof(unsigned long long) + (size_t)((Kp * L + 3) & ~3) * sizeof(float) +
         (size_t)(kEm32Threads / 64) * kEm32MaxK * 64 * sizeof(float);
}

bool estep_em32_supported(int K) { return K <= kEm32MaxK; }

void launch_estep_em_batch(const EStepArgs *d_args, const uint32_t *d_work, int nwork,
                           bool is2d, bool cubic, int K, int L, hipStream_t s) {
  if (nwork <= 0) return;
  BINDER_REQUIRE(K <= kEm32MaxK, "batched E-step needs K <= 32");
  static std::atomic<bool> attr_set[2][2] = {};
  static void (*const ks[2][2])(const EStepArgs *, const uint32_t *) = {
      {k_estep_em32_batch<false, false>, k_estep_em32_batch<false, true>},
      {k_estep_em32_batch<true, false>, k_estep_em32_batch<true, true>}};
  auto k = ks[is2d][cubic];
  if (!attr_set[is2d][cubic]) {
    HIP_OK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024));
    attr_set[is2d][cubic] = true;
  }
  hipLaunchKernelGGL(k, dim3(nwork), dim3(kEm32Threads), estep_em32_lds(K, L), s, d_args,
                     d_work);
  HIP_LAUNCH_OK();
}

void launch_hist_reduce_batch(const HistReduceArgs *d_args, int npairs, int max_rows, int KL,
                              hipStream_t s) {
  if (npairs <= 0) return;
  const int rows2 = (max_rows + kRowsPerReduce - 1) / kRowsPerReduce;
  hipLaunchKernelGGL(k_hist_reduce1_batch, dim3((KL + 255) / 256, rows2, npairs), dim3(256), 0,
                     s, d_args);
  HIP_LAUNCH_OK();
  hipLaunchKernelGGL(k_hist_reduce2_batch, dim3((KL + 63) / 64, 1, npairs), dim3(256), 0, s,
                     d_args);
  HIP_LAUNCH_OK();
}

void launch_theta_batch(const ThetaArgs *d_args, int npairs, hipStream_t s) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(k_theta_batch, dim3(npairs), dim3(1024), 0, s, d_args);
  HIP_LAUNCH_OK();
}

void launch_hist_reduce1(const unsigned long long *hpart, int rows, int KL,
                         unsigned long long *hpart2, hipStream_t s) {
  const int rows2 = (rows + kRowsPerReduce - 1) / kRowsPerReduce;
  hipLaunchKernelGGL(k_hist_reduce1, dim3((KL + 255) / 256, rows2), dim3(256), 0, s, hpart,
                     rows, KL, hpart2);
  HIP_LAUNCH_OK();
}

void launch_hist_reduce(const unsigned long long *hpart, int rows, int KL,
                        unsigned long long *hpart2, double *Hd, hipStream_t s) {
  const int rows2 = (rows + kRowsPerReduce - 1) / kRowsPerReduce;
  hipLaunchKernelGGL(k_hist_reduce1, dim3((KL + 255) / 256, rows2), dim3(256),
                     0, s, hpart, rows, KL, hpart2);
  HIP_LAUNCH_OK();
  hipLaunchKernelGGL(k_hist_reduce2, dim3((KL + 63) / 64), dim3(256), 0, s,
                     hpart2, rows2, KL, Hd);
  HIP_LAUNCH_OK();
}

void launch_theta(const ThetaArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_theta, dim3(1), dim3(1024), 0, s, a);
  HIP_LAUNCH_OK();
}

void launch_welford(const WelfordBatch &b, double *mean, double *m2,
                    long long nv, int N, hipStream_t s) {
  if (b.count <= 0) return;
  long long blocks = (nv + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  const dim3 g((unsigned)blocks), t(256);
  const bool full = b.count == kWelfordBatch;
  if (N == 3) {
    if (full) hipLaunchKernelGGL((k_welford<3, kWelfordBatch>), g, t, 0, s, b, mean, m2, nv);
    else hipLaunchKernelGGL((k_welford<3, 0>), g, t, 0, s, b, mean, m2, nv);
  } else {
    if (full) hipLaunchKernelGGL((k_welford<2, kWelfordBatch>), g, t, 0, s, b, mean, m2, nv);
    else hipLaunchKernelGGL((k_welford<2, 0>), g, t, 0, s, b, mean, m2, nv);
  }
  HIP_LAUNCH_OK();
}

}  // namespace binder
