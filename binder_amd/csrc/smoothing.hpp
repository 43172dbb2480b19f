// DCT-domain smoothing of a deformation on the GPU: the separable orthonormal
// DCT-II / DCT-III of each displacement component computed as per-axis
// split-BF16 MFMA contractions, with the prior filter, the sampling noise and
// the prior cost fused into the pass that completes the forward transform.
// Replaces NonLinearTransformation::constrain_transformation /
// sample_transformation / up_sample
// (BINDER/src/cpu/NonLinearTransformation.cpp:58-126, :190-223, :366-430).
#pragma once
#include <map>
#include <memory>
#include <vector>

#include "common.hpp"

namespace binder {

// One pass matrix M (n_in x n_out) in MFMA B-operand order, bf16 hi / lo.
struct PassMatrix {
  int n_in = 0, n_out = 0;
  DBuf<uint32_t> hi, lo;
  // Even/odd split of a square DCT-II / DCT-III of size n % 16 == 0 (H = n / 2,
  // 16 < H <= 96): two H x H matrices [even | odd] in the same B-operand order,
  // used by the persistent contraction when the pass allows it (sym_ks > 0).
  // Forward: E[x][m] = C_n[2m][x], O[x][m] = C_n[2m+1][x] on the sums and
  // differences of mirrored inputs; inverse: E[m][x] = C_n[2m][x],
  // O[m][x] = C_n[2m+1][x] on the even and odd coefficients.
  int sym_inverse = 0, sym_ks = 0;
  DBuf<uint32_t> shi, slo;
};

class MatrixCache {
 public:
  // kDctFusedX: the four half-size matrices of the fused x pass (n_in = n_out = N).
  enum Kind { kDctForward = 1, kDctInverse = 2, kDctFusedX = 3 };
  // kDctForward: M[k][c] = C_{n}[c][k] (n = n_in = n_out);
  // kDctInverse: M[k][c] = C_{n_out}[k][c] (first n_in coefficients).
  const PassMatrix &get(int kind, int n_in, int n_out, hipStream_t stream);

 private:
  std::map<long long, std::unique_ptr<PassMatrix>> cache_;
};

struct FilterArgs {
  const float *D;
  const double *gamma;
  double sigma2;
  double etarho[3];
  int noise, raw;
  uint32_t k0, k1, sweep, stream;
  const uint32_t *keyx;  // optional device key mask: key = (k0, k1) ^ keyx
  double *cost_part, *raw_part;
};

// The index grid (x = voxel indices) as a pass sees it: component c of the
// element at (row, k) is k when c == axis, else (row / div[c]) mod mod[c]
// (no modulo when mod[c] == 0).
struct GridCoords {
  int axis;
  int div[3], mod[3];
};

// Out[c][k'][r] = sum_k In[c][r][k] M[k][k'] for r < R, c < comps; with
// rin > 0 the rows form blocks of rin and row r = b * rin + i goes to
// Out[c][b][k'][i] instead.  epi: 0 store, 1 filter (needs f), 2 add xepi
// (laid out as Out).  xpro: In - xpro prologue.  With xg, the x of the
// prologue / epilogue is generated from the indices instead of read (xpro /
// xepi then only say which end applies it).  Returns the number of partial-
// sum slots written by a filter pass.
long long contract_pass(hipStream_t stream, const PassMatrix &m, const float *in,
                        float *out, int comps, long long cin, long long cout,
                        long long R, const float *xpro, const float *xepi, int epi,
                        const FilterArgs *f, KernelTimer *timer, const char *name,
                        long long rin = 0, const GridCoords *xg = nullptr);
// x axis of a [comp][N][R] field in one pass: forward DCT-II, prior filter
// with noise and cost (f), inverse DCT-III; N even, N <= 192 (m from
// MatrixCache kDctFusedX).  Returns the partial-sum slots written.
bool fused_x_supported(int n);
long long fused_x_pass(hipStream_t stream, const PassMatrix &m, const float *in, float *out,
                       int comps, long long cstride, long long R, const FilterArgs &f,
                       KernelTimer *timer, const char *name);
// Upper bound on the partial-sum slots a filter pass over R rows writes.
long long contract_blocks(long long R, int comps);
int persistent_blocks();

void up_sample_field(const float *in_soa, int nx, int ny, int nz, int nxo,
                     int nyo, int nzo, float *out_soa, hipStream_t stream,
                     MatrixCache &mats);

// Gamma(alpha0, rate beta0) prior of the regularisation weight nu
// (NonLinearTransformation.cpp:41-42).
constexpr double kNuAlpha0 = 1.0;
constexpr double kNuBeta0 = 10000.0;

struct SmoothOptions {
  bool sample = false;         // add DCT-domain Gaussian noise
  bool sample_gamma = false;   // accumulate the raw regulariser and draw nu
  uint32_t key0 = 0, key1 = 0; // Philox key
  const uint32_t *keyx = nullptr;  // optional device key mask (key ^ keyx)
  uint32_t sweep = 0;          // Philox counter word
  uint32_t stream = 0;         // Philox stream word
  double *prior_out = nullptr; // device double: receives sum_n 0.5 g_n D p^2
  bool want_prior = true;      // false: skip the reduction of the prior cost
  // with want_prior: the kernel that sums the prior cost also writes
  // host_stats[0..2] = stats_in[0], stats_in[1], prior cost (host_stats in
  // page-locked memory; stats_in on the device, complete before run())
  double *host_stats = nullptr;
  const double *stats_in = nullptr;
  hipStream_t exec_stream = nullptr;  // stream to run on (null: the smoother's own)
};

struct EmSmoothPlan;

class Smoother {
 public:
  // gamma/rho/eta as in NonLinearTransformation.cpp:15-43; D is [nx,ny,nz].
  Smoother(int nx, int ny, int nz, const double gamma[3], const double rho[3],
           const double eta[3], const float *D_host, double sigma,
           hipStream_t stream);
  // The same with D made on the device from its separable terms:
  // D[x][y][z] = (float)(t[x] + t[nx + y] (+ t[nx + ny + z]))^order (FP64,
  // additions in axis order; order 1 or 2), terms t of length nx + ny + nz.
  Smoother(int nx, int ny, int nz, const double gamma[3], const double rho[3],
           const double eta[3], const double *axis_terms, int order, double sigma,
           hipStream_t stream);

  // dream, x: [N][nv] SoA on the device (x may be null: zero).  out: [N][nv].
  // Computes out = x + IDCT(filter(DCT(dream - x)) [+ noise]).
  void run(const float *dream, const float *x, float *out,
           const SmoothOptions &opt, KernelTimer *timer);

  int N() const { return N_; }
  long long nv() const { return nv_; }
  double *gamma_device() const { return gamma_.get(); }
  void set_gamma(const double g[3]);
  void get_gamma(double g[3]);
  double nu() const { return nu_; }
  // x of run() is the index grid: the passes generate it instead of reading it
  void set_index_grid(bool v) { index_grid_ = v; }

  // Flops of one run (forward + inverse contractions over all components).
  double flops() const;

  // The launches of run() for an EM smoothing (no sampling) of a 3D field,
  // without launching them; prior_out (device, may be null) receives the
  // prior cost.
  void plan_em(const float *dream, const float *x, float *out, double *prior_out,
               EmSmoothPlan &plan, hipStream_t stream);
  // Whether this smoother's EM smoothing can join a SmoothBatch: 3D, fused x
  // pass, and y / z contractions small enough for the non-persistent kernel.
  bool batchable_em() const;

 private:
  int nx_, ny_, nz_, N_;
  long long nv_;
  double sigma_, rho_[3], eta_[3], nu_;
  bool index_grid_ = false;
  hipStream_t stream_;
  void init_buffers(const double gamma[3], const double rho[3], const double eta[3]);
  DBuf<float> D_, bufA_, bufB_;
  DBuf<double> gamma_, part_cost_, part_raw_, scal_;
  MatrixCache mats_;
};

// Device argument arrays of one batched launch (internal to SmoothBatch).
struct SmoothBatchLaunchView {
  const void *args, *prefix;
  int n, blocks;
  size_t lds;
  int comps;
};

// The EM smoothing of several smoothers in lockstep: pass by pass, one
// launch per pass and kernel variant over all of them.  build() plans the
// launches for these jobs (their device arguments go up on `stream`);
// launch() enqueues them.  Each smoother computes exactly what its own
// run() computes.  Every smoother must satisfy batchable_em().
class SmoothBatch {
 public:
  struct Job {
    Smoother *sm;
    const float *dream, *x;
    float *out;
    double *prior_out;
  };
  SmoothBatch();
  ~SmoothBatch();
  SmoothBatch(const SmoothBatch &) = delete;
  SmoothBatch &operator=(const SmoothBatch &) = delete;
  void build(const std::vector<Job> &jobs, hipStream_t stream);
  void launch(hipStream_t stream) const;

 private:
  struct Launch;
  std::vector<Launch> launches_;
  DBuf<unsigned char> dev_;
  unsigned char *pin_ = nullptr;
  size_t pin_cap_ = 0;
};

}  // namespace binder
