// E-step, joint-histogram and likelihood-parameter kernels (declarations).
#pragma once
#include "common.hpp"

namespace binder {

// Fixed-point scale of the LDS/partial histograms: posteriors are added as
// round(w * 2^31) (one 32-bit conversion, w <= 1) into 64-bit integers, so
// block and grid reductions are exact and the histogram is independent of
// scheduling.
constexpr double kHistScale = 2147483648.0;
constexpr int kHistShift = 31;  // kHistScale == 2^kHistShift
constexpr double kLn2 = 0.693147180559945309417232121458;
// The EM E-step's per-block sums of log2 Z_i: int64 fixed point with scale
// 2^22 (|log2 Z_i| <= 150 for every positive float, so each term fits an
// int32 and 2^30 voxels cannot overflow the sum), plus two sentinels for a
// block with a zero Z_i (mll = +inf) or a NaN.
constexpr double kMllFixScale = 4194304.0;
constexpr float kMllFixScaleF = 4194304.0f;
constexpr float kMllFixMax = 150.0f;
constexpr long long kMllFixNaN = (long long)(-9223372036854775807LL - 1);
constexpr long long kMllFixZero = (long long)(-9223372036854775807LL);

// Row stride (in floats) of the likelihood table theta[u][v] held in LDS by
// the E-step kernels.  Lanes of one wavefront read theta at (fixed bin u,
// moving bin v) pairs that cluster along the image's intensity mapping; a
// stride of 7 mod 32 maps such pairs to distinct banks far more often than
// K + 1 (about 1.8x fewer conflict cycles on a 32-bank model of the
// 160x192x160 workload's access pattern).  Wide tables (L > 64) keep K + 1 to
// bound the LDS footprint.
__host__ __device__ inline int theta_row_stride(int K, int L) {
  return L <= 64 ? K + 1 + ((7 - (K + 1)) & 31) : K + 1;
}

struct EStepArgs {
  const float *d;           // [N][nv] current deformed positions (SoA)
  const uint8_t *bv;        // [nv] fixed-image bins
  const uint32_t *quads;    // 3D: padded moving bins, 4 consecutive z per word
  const uint8_t *bins2d;    // 2D: padded moving bins
  const float *thT;         // [L][K] likelihood table, transposed
  int K, L;
  long long nv;
  int nxn, nyn, nzn;        // moving grid
  int Yp, Zq;               // padded strides
  float *dream;             // [N][nv] out
  unsigned long long *hpart;  // [blocks][K*L] out
  double *mllpart;            // [blocks] out
  uint32_t k0, k1, sweep;     // sampling keys
  // optional device-resident key mask: the key used is (k0, k1) ^ keyx[0..1]
  // (lets one captured sweep graph serve chains with different seeds)
  const uint32_t *keyx;
  // logical blocks of the EM E-step of this registration (batched launches;
  // a single launch uses its grid size)
  int nblocks;
  // SSD likelihood (BINDER/src/cpu/SSDLikelihood.cpp:60-86) in place of the
  // table thT: ssd_par = (sigma^2, shift, scale) on the device, ssd_ss the
  // scale_shift flag.  Each voxel's node weights are B_j exp(ll_j - ll_r),
  // ll_r the log-likelihood of its closest (positive-weight) node, with
  // ll_j - ll_r formed as a difference of squares, so posteriors are exact
  // however peaked the likelihood (the reference's log-space E-step,
  // Registration.cpp:332-440); mll adds ll_r back in FP64.  The EM E-step
  // then runs k_estep (FP64 mll partials) for any K.
  const double *ssd_par;
  int ssd_ss;
  // B-spline order of the node weights: 1, 2 or 3 (0 is taken as 3)
  int order;
};

void launch_estep(const EStepArgs &a, bool is2d, bool sample, int blocks,
                  hipStream_t s);

// Sum `rows` partial histograms into Hd (doubles, divided by the scale).
void launch_hist_reduce(const unsigned long long *hpart, int rows, int KL,
                        unsigned long long *hpart2, double *Hd, hipStream_t s);

// Several registrations' EM E-steps in one launch (K <= 32 only): d_work
// holds nwork items (pair << 16 | logical block), pair-major; every pair's
// EStepArgs (d_args[pair]) carries its own block count, so its partials are
// those of its own single launch.  cubic: every pair's spline order is 3
// (the order-3-only kernel); otherwise each pair's order is read at run time.
bool estep_em32_supported(int K);
size_t estep_em32_lds(int K, int L);
void launch_estep_em_batch(const EStepArgs *d_args, const uint32_t *d_work, int nwork,
                           bool is2d, bool cubic, int K, int L, hipStream_t s);

struct HistReduceArgs {
  const unsigned long long *hpart;  // [rows][KL]
  unsigned long long *hpart2;       // [ceil(rows / kHistReduceRows)][KL]
  double *Hd;                       // [KL]
  int rows, KL;
};
// launch_hist_reduce for npairs registrations (max_rows: the largest rows).
void launch_hist_reduce_batch(const HistReduceArgs *d_args, int npairs, int max_rows, int KL,
                              hipStream_t s);

enum ThetaMode { kThetaKeep = 0, kThetaMap = 1, kThetaDirichlet = 2 };

struct ThetaArgs {
  // [K*L] histogram; with hsum, written from the sum of the hrows partial
  // rows hsum[r][K*L] (fixed point, as k_hist_reduce2 forms it), else read
  double *Hd;
  const unsigned long long *hsum;
  int hrows;
  double *theta;        // [K*L] in/out
  float *thT;           // [L*K] out
  double *rowbuf;       // [K*L] scratch
  const double *mllpart;
  int nmll;
  int K, L;
  double alpha, lg_sum_alpha, lg_alpha;
  int mode;
  double *stats;        // stats[0] = mll, stats[1] = prior_theta
  uint32_t k0, k1, sweep;
  const uint32_t *keyx;  // optional key mask, as in EStepArgs
  // mllpart holds the EM E-step's (K <= 32) int64 fixed-point partials
  // instead of doubles
  int mll_fixed;
};

void launch_theta(const ThetaArgs &a, hipStream_t s);
// The first stage of launch_hist_reduce alone: rows partial rows summed
// kHistReduceRows at a time into hpart2 (ceil(rows / kHistReduceRows) rows).
void launch_hist_reduce1(const unsigned long long *hpart, int rows, int KL,
                         unsigned long long *hpart2, hipStream_t s);
// Partial rows a theta launch sums itself instead of a separate reduction,
// and the rows one reduction stage sums into one.
constexpr int kThetaDirectRows = 32;
constexpr int kHistReduceRows = 16;
// launch_theta for npairs registrations, one block each.
void launch_theta_batch(const ThetaArgs *d_args, int npairs, hipStream_t s);

// FP64 Welford update of mean [N][nv] and upper-triangular M2 [N(N+1)/2][nv]
// with `count` samples x[0..count) (each [N][nv]) taken in that order, n0
// samples having been folded in before: per voxel the accumulators are
// loaded once (not at all when n0 == 0: they start from zero), updated
// sample by sample exactly as one update per sweep would, and stored once.
constexpr int kWelfordBatch = 16;
struct WelfordBatch {
  const float *x[kWelfordBatch];
  int count;
  long long n0;
  // optional device-resident count added to n0 (samples folded by earlier
  // chains; lets one captured graph serve every chain)
  const long long *n0_base;
};
void launch_welford(const WelfordBatch &b, double *mean, double *m2,
                    long long nv, int N, hipStream_t s);

}  // namespace binder
