// Registration and NonLinearTransformation handles behind the C ABI
// (include/binder_c.h): device-resident state, the EM loop and the Gibbs
// sampler, driven from the host with one HIP stream per handle.
//
// Device layout (all [component][voxel] SoA so every pass is coalesced):
//   d, x, dream      float [N][nv]   positions (grid units)
//   bv               uint8 [nv]      fixed-image bins
//   quads            uint32 [(Nx_n+2)(Ny_n+2)(Nz_n+5)] packed moving bins (3D)
//   bins2d           uint8 [(Nx_n+2)(Ny_n+2)]          padded moving bins (2D)
//   theta (f64), thT (f32, transposed), Hd (f64) K*L tables
//   mean (f64 [N][nv]), m2 (f64 [N(N+1)/2][nv]) sampler accumulators
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>

#include "autotune.hpp"
#include "estep.hpp"
#include "models.hpp"
#include "smoothing.hpp"
#include "topology.hpp"
#include "resample.hpp"
#include "warmstart.hpp"

namespace binder {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

namespace {

// AoS [nv][N] <-> SoA [N][nv]
template <class T, class U>
void aos_to_soa(const T *in, U *out, long long nv, int N) {
  for (long long i = 0; i < nv; ++i)
    for (int c = 0; c < N; ++c) out[c * nv + i] = (U)in[i * N + c];
}
template <class T, class U>
void soa_to_aos(const T *in, U *out, long long nv, int N) {
  for (long long i = 0; i < nv; ++i)
    for (int c = 0; c < N; ++c) out[i * N + c] = (U)in[c * nv + i];
}

// Device-side layout changes: [nv][N] interleaved <-> [N][nv] planes, and
// the index grid (i, j[, k]) as planes.
__global__ __launch_bounds__(256) void k_interleaved_to_planes(const float *__restrict__ in,
                                                               float *__restrict__ out,
                                                               long long nv, int N) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride)
    for (int c = 0; c < N; ++c) out[c * nv + i] = in[i * N + c];
}
__global__ __launch_bounds__(256) void k_planes_to_interleaved(const float *__restrict__ in,
                                                               float *__restrict__ out,
                                                               long long nv, int N) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride)
    for (int c = 0; c < N; ++c) out[i * N + c] = in[c * nv + i];
}
__global__ __launch_bounds__(256) void k_index_grid(float *__restrict__ out, int nx, int ny,
                                                    int nz, int N) {
  const long long nv = (long long)nx * ny * nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
    const long long yz = (long long)ny * nz;
    out[i] = (float)(i / yz);
    out[nv + i] = (float)((i / nz) % ny);
    if (N == 3) out[2 * nv + i] = (float)(i % nz);
  }
}

// flag stays 1 if planes == the index grid (i, j[, k]) exactly
__global__ __launch_bounds__(256) void k_check_index_grid(const float *__restrict__ x, int nx,
                                                          int ny, int nz, int N,
                                                          int *__restrict__ flag) {
  const long long nv = (long long)nx * ny * nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  bool ok = true;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
    const long long yz = (long long)ny * nz;
    ok = ok && x[i] == (float)(i / yz) && x[nv + i] == (float)((i / nz) % ny);
    if (N == 3) ok = ok && x[2 * nv + i] == (float)(i % nz);
  }
  if (!ok) *flag = 0;
}

// Padded, packed moving bins of the 3D E-step: word (x, y, zq) of the
// [nx+2][ny+2][nz+5] array holds the bins of z = zq - 4 .. zq - 1 of node
// (x - 1, y - 1) in bytes 0..3, zero outside the grid.
__global__ __launch_bounds__(256) void k_pack_quads(const uint8_t *__restrict__ bins, int nx,
                                                    int ny, int nz, uint32_t *__restrict__ q) {
  const int Yp = ny + 2, Zq = nz + 5;
  const long long n = (long long)(nx + 2) * Yp * Zq;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int zq = (int)(i % Zq);
    const int yp = (int)((i / Zq) % Yp);
    const int xp = (int)(i / ((long long)Zq * Yp));
    uint32_t w = 0u;
    if (xp >= 1 && xp <= nx && yp >= 1 && yp <= ny) {
      const uint8_t *col = bins + ((long long)(xp - 1) * ny + (yp - 1)) * nz;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int z = zq - 4 + c;
        if (z >= 0 && z < nz) w |= (uint32_t)col[z] << (8 * c);
      }
    }
    q[i] = w;
  }
}

int grid_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// Host [nv][N] floats -> device planes (staged through `tmp`, nv * N floats).
void upload_planes(const float *host, DBuf<float> &dst, DBuf<float> &tmp, long long nv, int N,
                   hipStream_t s) {
  if (!tmp.get() || tmp.n < (size_t)(nv * N)) tmp.alloc((size_t)(nv * N));
  HIP_OK(hipMemcpyAsync(tmp.get(), host, sizeof(float) * nv * N, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_interleaved_to_planes, dim3(grid_blocks(nv)), dim3(256), 0, s,
                     tmp.get(), dst.get(), nv, N);
  HIP_LAUNCH_OK();
}

// Device planes -> host [nv][N] floats (staged through `tmp`); synchronises.
// fn(begin, end) over [0, n) split across up to 16 host threads (one thread
// when n < min_n).
template <class F>
void parallel_ranges(long long n, F &&fn, long long min_n = 1ll << 15) {
  unsigned hw = std::thread::hardware_concurrency();
  long long T = hw ? (long long)std::min(hw, 16u) : 1;
  if (n < min_n) T = 1;
  if (T > n) T = n > 0 ? n : 1;
  if (T == 1) {
    fn(0ll, n);
    return;
  }
  std::vector<std::thread> ts;
  for (long long t = 0; t < T; ++t)
    ts.emplace_back([&, t] { fn(n * t / T, n * (t + 1) / T); });
  for (auto &t : ts) t.join();
}

// Page-locked staging for large downloads, shared by every handle of the
// process (one download at a time) and held until the process ends.
std::mutex g_pinned_mu;
void *g_pinned = nullptr;
size_t g_pinned_bytes = 0;

void download_planes(const float *src, float *host, DBuf<float> &tmp, long long nv, int N,
                     hipStream_t s) {
  if (!tmp.get() || tmp.n < (size_t)(nv * N)) tmp.alloc((size_t)(nv * N));
  hipLaunchKernelGGL(k_planes_to_interleaved, dim3(grid_blocks(nv)), dim3(256), 0, s, src,
                     tmp.get(), nv, N);
  HIP_LAUNCH_OK();
  const size_t bytes = sizeof(float) * (size_t)(nv * N);
  if (bytes < ((size_t)4 << 20)) {
    HIP_OK(hipMemcpyAsync(host, tmp.get(), bytes, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return;
  }
  // a full-size field: one DMA into page-locked memory, then a threaded copy
  // into the caller's (pageable) buffer, instead of the HIP runtime's staged
  // pageable transfer
  std::lock_guard<std::mutex> lock(g_pinned_mu);
  if (g_pinned_bytes < bytes) {
    if (g_pinned) HIP_OK(hipHostFree(g_pinned));
    g_pinned = nullptr;
    g_pinned_bytes = 0;
    HIP_OK(hipHostMalloc(&g_pinned, bytes, hipHostMallocDefault));
    g_pinned_bytes = bytes;
  }
  HIP_OK(hipMemcpyAsync(g_pinned, tmp.get(), bytes, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  const char *from = static_cast<const char *>(g_pinned);
  char *to = reinterpret_cast<char *>(host);
  parallel_ranges((long long)bytes, [&](long long b0, long long b1) {
    std::memcpy(to + b0, from + b0, (size_t)(b1 - b0));
  }, (long long)4 << 20);
}

// Page-locked staging for uploads of bins (the fixed and moving bins of a
// registration, one byte each), shared by every handle of the process and
// held until the process ends; callers hold g_upload_mu until their copies
// have completed.
std::mutex g_upload_mu;
uint8_t *g_upload = nullptr;
size_t g_upload_bytes = 0;

uint8_t *pinned_upload(size_t bytes) {
  if (g_upload_bytes < bytes) {
    if (g_upload) HIP_OK(hipHostFree(g_upload));
    g_upload = nullptr;
    g_upload_bytes = 0;
    HIP_OK(hipHostMalloc((void **)&g_upload, bytes, hipHostMallocDefault));
    g_upload_bytes = bytes;
  }
  return g_upload;
}

struct Stream {
  hipStream_t s = nullptr;
  Stream() { HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct Event {
  hipEvent_t e = nullptr;
  Event() { HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
};

struct PinnedStats {
  double *p = nullptr;
  PinnedStats() { HIP_OK(hipHostMalloc((void **)&p, 16 * sizeof(double), 0)); }
  ~PinnedStats() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace
}  // namespace binder

using namespace binder;

struct binder_registration {
  Stream stream;
  // the sampler's theta chain (histogram reduction, Dirichlet draw) runs on
  // side_stream, beside the deformation draw, which does not depend on it
  Stream side_stream;
  Event estep_done, theta_done;
  KernelTimer timer;
  PinnedStats host_stats;
  int nxv, nyv, nzv, nxn, nyn, nzn, N;
  int order = 3;  // B-spline order of the E-step's node weights
  // the smoothing's sigma for that order (Spline.cpp:20-26)
  double sigma_spline() const {
    return order == 3 ? std::sqrt(9.0 / (8.0 * M_PI))
                      : (order == 2 ? std::sqrt(8.0 / (9.0 * M_PI)) : 0.4);
  }
  long long nv, nn;
  double nnz;
  bool is2d;
  // d: the current field lives in slot dslot of a ring of kWelfordBatch
  // buffers (slot 0 = d, slots 1.. = dextra, allocated by the first sampler
  // call): each sampling sweep writes the next slot, so the last
  // kWelfordBatch samples stay resident and the FP64 accumulators are
  // updated once per batch.
  DBuf<float> d, dextra, x, dream;
  DBuf<float> stage;  // [nv][N] staging for host transfers
  bool x_index_grid = false;  // x holds the grid indices exactly
  int dslot = 0;
  int ring = kWelfordBatch;   // slots in use (a call with fewer sweeps needs fewer)
  float *slot(int k) {
    return k == 0 ? d.get() : dextra.get() + (size_t)(k - 1) * N * nv;
  }
  float *dcur() { return slot(dslot); }
  // likelihood
  bool has_mi = false;
  int K = 0, L = 0;
  double alpha = 0.0;
  DBuf<uint8_t> bv, bins2d, nb8;  // nb8: moving bins as uploaded, before packing
  DBuf<uint32_t> quads;
  DBuf<double> theta, Hd, rowbuf, mllpart, stats;
  DBuf<float> thT;
  DBuf<unsigned long long> hpart, hpart2;
  // second buffers of the EM loop's state (field, theta, dream, objective
  // terms): iteration k reads parity k & 1 and writes parity (k + 1) & 1, so
  // iteration k + 1 can be queued before the host has read iteration k's
  // objective
  DBuf<float> em_d, em_dream, em_thT;
  DBuf<double> em_theta, em_stats;
  Event em_done[2];
  // logical blocks of the EM E-step and grid of the sampling E-step (launch
  // shapes the autotuner may change; neither changes any result), and the
  // histogram partial rows the last E-step wrote
  int estep_blocks = 0, sample_blocks = 0, hist_rows = 0;
  // likelihood model: 0 MI (theta table), 1 SSD (theta[0..3] hold
  // sigma^2, shift, scale and the draw flag, which the E-step kernels read
  // directly; thT is unused)
  int lik = 0;
  int ssd_ss = 0;
  double ssd_a0 = 0.0, ssd_b0 = 0.0;
  double ssd_given[2] = {1.0, 0.0};  // SSD scale, shift as set (fixed when scale_shift == 0)
  // transformation: the DCT smoother, or an affine / rigid model
  std::unique_ptr<Smoother> sm;
  std::unique_ptr<LinearModel> lin;
  // the transformation step: out = constrained / sampled field of `dream`
  void transform(const float *dream_p, float *out, const SmoothOptions &o) {
    if (lin) lin->run(dream_p, x.get(), out, o);
    else sm->run(dream_p, x.get(), out, o, tm());
  }
  // sampler
  DBuf<double> mean, m2;
  int samples = -1;
  std::vector<double> trace;

  KernelTimer *tm() { return timer.enabled() ? &timer : nullptr; }
  hipStream_t s() const { return stream.s; }

  void require_ready() {
    if (!has_mi)
      throw Error(BINDER_ESTATE, "setMILikelihood must be called first");
    if (!sm && !lin)
      throw Error(BINDER_ESTATE, "setNonLinearTransformation must be called first");
  }

  EStepArgs estep_args(bool sample, uint32_t k0, uint32_t k1, uint32_t sweep) {
    EStepArgs a{};
    a.d = dcur();
    a.bv = bv.get();
    a.quads = quads.get();
    a.bins2d = bins2d.get();
    a.thT = thT.get();
    a.K = K;
    a.L = L;
    a.nv = nv;
    a.nxn = nxn;
    a.nyn = nyn;
    a.nzn = nzn;
    a.Yp = nyn + 2;
    a.Zq = nzn + 5;
    a.dream = dream.get();
    a.hpart = hpart.get();
    a.mllpart = mllpart.get();
    a.k0 = k0;
    a.k1 = k1;
    a.sweep = sweep;
    a.ssd_par = lik == 1 ? theta.get() : nullptr;
    a.ssd_ss = ssd_ss;
    a.order = order;
    (void)sample;
    return a;
  }

  void estep_only(bool sample, uint32_t k0, uint32_t k1, uint32_t sweep,
                  const uint32_t *keyx = nullptr) {
    EStepArgs a = estep_args(sample, k0, k1, sweep);
    a.keyx = keyx;
    hist_rows = sample ? sample_blocks : estep_blocks;
    TimedLaunch t(tm(), sample ? "estep_sample" : "estep", s());
    launch_estep(a, is2d, sample, hist_rows, s());
  }

  void hist_reduce_on(hipStream_t st) {
    TimedLaunch t(tm(), "hist_reduce", st);
    launch_hist_reduce(hpart.get(), hist_rows, K * L, hpart2.get(), Hd.get(), st);
  }

  // Histogram reduction and theta step.  Up to kThetaDirectRows partial
  // rows are summed by the theta launch itself, up to kHistReduceRows times
  // that many after one reduction stage, more after both; Hd is the same in
  // every case (exact integer sums) and holds the histogram afterwards.
  void hist_theta(hipStream_t st, int mode, uint32_t k0, uint32_t k1, uint32_t sweep,
                  const uint32_t *keyx = nullptr, double *theta_out = nullptr,
                  float *thT_out = nullptr, double *stats_out = nullptr,
                  const double *theta_in = nullptr) {
    if (lik == 1) {
      ssd_theta(st, mode, k0, k1, sweep, keyx, theta_out, thT_out, stats_out, theta_in);
      return;
    }
    ThetaArgs t = theta_args(mode, k0, k1, sweep, keyx, theta_out, thT_out, stats_out);
    const int rows2 = (hist_rows + kHistReduceRows - 1) / kHistReduceRows;
    {
      TimedLaunch tl(tm(), "hist_reduce", st);
      if (hist_rows <= kThetaDirectRows) {
        t.hsum = hpart.get();
        t.hrows = hist_rows;
      } else if (rows2 <= kThetaDirectRows) {
        launch_hist_reduce1(hpart.get(), hist_rows, K * L, hpart2.get(), st);
        t.hsum = hpart2.get();
        t.hrows = rows2;
      } else {
        launch_hist_reduce(hpart.get(), hist_rows, K * L, hpart2.get(), Hd.get(), st);
      }
    }
    TimedLaunch tl(tm(), mode == kThetaDirichlet ? "dirichlet" : "theta", st);
    launch_theta(t, st);
  }

  // Histogram partial rows (and mll partial slots) for E-steps of up to
  // `rows` blocks.
  void ensure_rows(int rows) {
    const size_t KL = (size_t)K * L;
    if (hpart.n < (size_t)rows * KL) {
      HIP_OK(hipStreamSynchronize(s()));
      hpart.alloc((size_t)rows * KL);
      hpart2.alloc((size_t)((rows + kHistReduceRows - 1) / kHistReduceRows) * KL);
    }
    if (mllpart.n < (size_t)rows) {
      HIP_OK(hipStreamSynchronize(s()));
      mllpart.alloc(rows);
    }
  }

  // Autotuned E-step launch shapes (large grids only).  The EM E-step's
  // partials are exact integers (fixed-point histogram and log sums) and the
  // sampling E-step's draws depend on the voxel index only, so every
  // candidate gives the same results.
  void tune_estep_em() {
    if (!has_mi || lik != 0 || !estep_em32_supported(K) || nv < (1ll << 18)) return;
    // half, the same and twice the default number of blocks
    const int d = estep_blocks;
    int c[3], nc = 0;
    if (d / 2 >= 1) c[nc++] = d / 2;
    c[nc++] = d;
    if ((long long)d * 2 * 512 <= 2 * nv && d * 2 <= 8192) c[nc++] = d * 2;
    const long long key = (nv << 22) ^ ((long long)K << 12) ^ ((long long)L << 2) ^ (is2d ? 1 : 0);
    const int nb = Autotuner::get().choose("estep_em", key, c, nc, estep_blocks, s(), [&](int b) {
      ensure_rows(b);
      EStepArgs a = estep_args(false, 0, 0, 0);
      launch_estep(a, is2d, false, b, s());
    });
    ensure_rows(nb);
    estep_blocks = nb;
  }

  void tune_estep_sample() {
    if (!has_mi || nv < (1ll << 18)) return;
    const int d = sample_blocks;
    int c[3], nc = 0;
    if (d / 2 >= 1) c[nc++] = d / 2;
    c[nc++] = d;
    if ((long long)d * 2 * 256 <= 2 * nv && d * 2 <= 8192) c[nc++] = d * 2;
    const long long key = (nv << 22) ^ ((long long)K << 12) ^ ((long long)L << 2) ^ (is2d ? 1 : 0);
    const int nb = Autotuner::get().choose("estep_sample", key, c, nc, sample_blocks, s(),
                                           [&](int b) {
      ensure_rows(b);
      EStepArgs a = estep_args(true, 0, 0, 0);
      launch_estep(a, is2d, true, b, s());
    });
    ensure_rows(nb);
    sample_blocks = nb;
  }

  // E-step + histogram reduction (+ theta step with the given mode).
  void e_and_theta(bool sample, int theta_mode, uint32_t k0, uint32_t k1,
                   uint32_t sweep) {
    estep_only(sample, k0, k1, sweep);
    if (theta_mode >= 0) hist_theta(s(), theta_mode, k0, k1, sweep);
    else hist_reduce_on(s());
  }

  // The SSD parameter step in place of the theta step: the same histogram
  // reduction, then mll, the parameter update / draw of `mode` (theta modes:
  // keep, MAP update, Dirichlet draw) and the prior cost.
  // With `fresh` false the histogram is Hd as it stands (the last step's,
  // which every path leaves there) instead of the last E-step's partials.
  void ssd_theta(hipStream_t st, int mode, uint32_t k0, uint32_t k1, uint32_t sweep,
                 const uint32_t *keyx, double *theta_out, float *thT_out,
                 double *stats_out, const double *theta_in, bool fresh = true) {
    SsdThetaArgs a{};
    a.Hd = Hd.get();
    const int rows2 = (hist_rows + kHistReduceRows - 1) / kHistReduceRows;
    if (fresh) {
      TimedLaunch tl(tm(), "hist_reduce", st);
      if (hist_rows <= kThetaDirectRows) {
        a.hsum = hpart.get();
        a.hrows = hist_rows;
      } else if (rows2 <= kThetaDirectRows) {
        launch_hist_reduce1(hpart.get(), hist_rows, K * L, hpart2.get(), st);
        a.hsum = hpart2.get();
        a.hrows = rows2;
      } else {
        launch_hist_reduce(hpart.get(), hist_rows, K * L, hpart2.get(), Hd.get(), st);
      }
    }
    a.mllpart = mllpart.get();
    a.nmll = hist_rows;
    a.mll_fixed = 0;  // the normalised E-step's FP64 partials
    a.par_in = theta_in ? theta_in : theta.get();
    a.par_out = theta_out ? theta_out : theta.get();
    (void)thT_out;
    a.bins = K;
    a.shift_scale = ssd_ss;
    a.mode = mode == kThetaMap ? kSsdUpdate : (mode == kThetaDirichlet ? kSsdDraw : kSsdKeep);
    a.a0 = ssd_a0;
    a.b0 = ssd_b0;
    a.stats = stats_out ? stats_out : stats.get();
    a.k0 = k0;
    a.k1 = k1;
    a.sweep = sweep;
    a.keyx = keyx;
    TimedLaunch tl(tm(), mode == kThetaDirichlet ? "dirichlet" : "theta", st);
    launch_ssd_theta(a, st);
  }

  void run_theta(int mode, uint32_t k0, uint32_t k1, uint32_t sweep,
                 hipStream_t st = nullptr, const uint32_t *keyx = nullptr,
                 double *theta_out = nullptr, float *thT_out = nullptr,
                 double *stats_out = nullptr) {
    if (!st) st = s();
    if (lik == 1) {  // the histogram as it stands
      ssd_theta(st, mode, k0, k1, sweep, keyx, theta_out, thT_out, stats_out, nullptr,
                false);
      return;
    }
    const ThetaArgs t = theta_args(mode, k0, k1, sweep, keyx, theta_out, thT_out, stats_out);
    TimedLaunch tl(tm(), mode == kThetaDirichlet ? "dirichlet" : "theta", st);
    launch_theta(t, st);
  }

  ThetaArgs theta_args(int mode, uint32_t k0, uint32_t k1, uint32_t sweep,
                       const uint32_t *keyx, double *theta_out, float *thT_out,
                       double *stats_out) {
    ThetaArgs t{};
    t.Hd = Hd.get();
    t.theta = theta_out ? theta_out : theta.get();
    t.thT = thT_out ? thT_out : thT.get();
    t.rowbuf = rowbuf.get();
    t.mllpart = mllpart.get();
    t.nmll = hist_rows;
    t.K = K;
    t.L = L;
    t.alpha = alpha;
    // dirichlet_logpdf constants (MILikelihood.cpp:183-206): sum_alpha is
    // accumulated L times, as the reference does
    double sa = 0.0;
    for (int l = 0; l < L; ++l) sa += alpha;
    t.lg_sum_alpha = std::lgamma(sa);
    t.lg_alpha = std::lgamma(alpha);
    t.mode = mode;
    t.stats = stats_out ? stats_out : stats.get();
    t.k0 = k0;
    t.k1 = k1;
    t.keyx = keyx;
    t.sweep = sweep;
    // the EM E-step for K <= 32 writes fixed-point mll partials; the
    // sampling E-step writes FP64 zeros (its mll is not used)
    t.mll_fixed = mode != kThetaDirichlet && estep_em32_supported(K);
    return t;
  }

  // Makes the ring ready for a call of `sweeps` sweeps: the current field
  // moves to slot 0 and the ring gets min(kWelfordBatch, sweeps + 1) slots
  // (no sample is overwritten before it is folded in either way).
  void prepare_ring(int sweeps) {
    const size_t nf = (size_t)N * nv;
    if (dslot != 0) {
      HIP_OK(hipMemcpyAsync(d.get(), dcur(), nf * sizeof(float), hipMemcpyDeviceToDevice, s()));
      dslot = 0;
    }
    ring = std::min(kWelfordBatch, sweeps + 1);
    const size_t need = (size_t)(ring - 1) * nf;
    if (need && dextra.n < need) {
      // the full ring at once, so a later call with more sweeps does not
      // reallocate (and synchronize) inside its sweeps
      HIP_OK(hipStreamSynchronize(s()));
      dextra.alloc((size_t)(kWelfordBatch - 1) * nf);
    }
  }

  // Enqueue N_b + N_s sampling sweeps (E-draw, Dirichlet draw, deformation
  // draw) with Philox key (k0, k1) [^ keyx], folding the N_s samples into the
  // FP64 accumulators (mean_p, m2_p) with Welford (they must hold zeros).
  // Nothing here synchronizes with the host or uses hipMemset, so the whole
  // sequence can be captured into a graph.
  void enqueue_sweeps(int N_b, int N_s, int sample_gamma, int every, int sp, int sl, int st,
                      uint32_t k0, uint32_t k1, const uint32_t *keyx, double *mean_p,
                      double *m2_p, const long long *n0_base = nullptr) {
    WelfordBatch wb{};
    wb.n0_base = n0_base;
    long long folded = 0;  // samples already in the accumulators
    const bool overlap = sp && sl && st;
    for (int sweep = 0; sweep < N_b + N_s; ++sweep) {
      if (overlap) {
        // E-draw; then the histogram reduction and the Dirichlet draw on the
        // side stream while the main stream draws the deformation from the
        // E-draw's dream locations; the next E-draw waits for theta
        if (sweep > 0) HIP_OK(hipStreamWaitEvent(s(), theta_done.e, 0));
        estep_only(true, k0, k1, (uint32_t)sweep, keyx);
        HIP_OK(hipEventRecord(estep_done.e, s()));
        HIP_OK(hipStreamWaitEvent(side_stream.s, estep_done.e, 0));
        hist_theta(side_stream.s, kThetaDirichlet, k0, k1, (uint32_t)sweep, keyx);
        HIP_OK(hipEventRecord(theta_done.e, side_stream.s));
      } else {
        if (sp) estep_only(true, k0, k1, (uint32_t)sweep, keyx);
        if (sp && sl) hist_theta(s(), kThetaDirichlet, k0, k1, (uint32_t)sweep, keyx);
        else if (sp) hist_reduce_on(s());
        else if (sl) run_theta(kThetaDirichlet, k0, k1, (uint32_t)sweep, nullptr, keyx);
      }
      if (st) {
        SmoothOptions o;
        o.sample = true;
        o.sample_gamma = sample_gamma && every > 0 && (sweep % every == 0);
        o.key0 = k0;
        o.key1 = k1;
        o.keyx = keyx;
        o.sweep = (uint32_t)sweep;
        o.stream = kStreamDeform;
        o.want_prior = false;  // the sampler reports no objective
        const int next = (dslot + 1) % ring;
        transform(dream.get(), slot(next), o);
        dslot = next;
      }
      if (sweep >= N_b) {
        // the sample stays in its ring slot until the batch is folded in,
        // which happens before the ring wraps onto it
        wb.x[wb.count++] = dcur();
        if (wb.count == kWelfordBatch || sweep == N_b + N_s - 1) {
          wb.n0 = folded;
          TimedLaunch t(tm(), "welford", s());
          launch_welford(wb, mean_p, m2_p, nv, N, s());
          folded += wb.count;
          wb.count = 0;
        }
      }
    }
    if (overlap) HIP_OK(hipStreamWaitEvent(s(), theta_done.e, 0));
  }
};

// Several registrations iterated in lockstep (batched EM): every iteration
// runs the E-step of all active registrations in one launch, then their
// histogram reductions and theta steps in one launch each, then their
// smoothings (batched launches for small grids, a few pool streams for large
// ones), and one copy of all objective terms to the host.  A registration's
// kernels compute exactly what its own EM loop computes (same logical
// blocks, same kernels), so each keeps the result it would have alone.
// Slots can be refilled between runs, so pairs at different pyramid levels
// share the launches.
constexpr int kBatchPool = 4;

struct binder_batch {
  Stream stream;
  Stream pool[kBatchPool];
  Event theta_done[2], iter_done[2];
  Event pool_done[2][kBatchPool];
  int K = -1, L = -1;
  bool is2d = false;
  // one slot per registration: its handle, its EM call's state (as the
  // local variables of binder_registration_em keep it)
  struct Slot {
    binder_registration *r = nullptr;
    int cap = 0, k = 0;
    double mlp_old = 1e100;
    bool done = true;
    float *F[2] = {}, *DR[2] = {}, *TT[2] = {};
    double *TH[2] = {};
  };
  std::vector<Slot> slots;
  int update = 1;  // the update flag the slots' buffers were prepared for
  DBuf<double> stats;      // [2][P][8] objective terms per parity and slot
  double *host = nullptr;  // pinned copy of stats
  // per iteration parity of a run: the device argument set of the (slot,
  // buffer parity) list it was built for
  struct ArgSet {
    std::vector<long long> key;
    DBuf<unsigned char> dev;  // EStepArgs[n] | HistReduceArgs[n] | ThetaArgs[n] | work
    unsigned char *pin = nullptr;
    size_t pin_cap = 0;
    size_t off_h = 0, off_t = 0, off_w = 0;
    int nwork = 0, max_rows = 0;
    bool valid = false;
    // the smoothing of registrations whose passes are all small runs as
    // batched launches on the batch stream; the others' on the pool streams
    SmoothBatch smooth;
    bool any_small = false;
    std::vector<int> big;  // positions in the list
  } set[2];
  ~binder_batch() {
    if (host) (void)hipHostFree(host);
    for (auto &a : set)
      if (a.pin) (void)hipHostFree(a.pin);
  }
};

struct binder_nlt {
  Stream stream;
  int nx, ny, nz, N;
  long long nv;
  std::unique_ptr<Smoother> sm;
  DBuf<float> in, out;
  DBuf<double> scal;
};

extern "C" {

const char *binder_last_error(void) { return g_last_error.c_str(); }
const char *binder_backend(void) { return "hip"; }

int binder_philox4x32_10(const unsigned int *ctr4, const unsigned int *key2,
                         unsigned int *out4) {
  return guarded([&] {
    BINDER_REQUIRE(ctr4 && key2 && out4, "null argument");
    const u32x4 r =
        philox4x32_10(u32x4{ctr4[0], ctr4[1], ctr4[2], ctr4[3]}, key2[0], key2[1]);
    out4[0] = r.x;
    out4[1] = r.y;
    out4[2] = r.z;
    out4[3] = r.w;
  });
}

}  // extern "C"

namespace {

// Shape checks and buffers shared by the two constructors; the caller fills
// the current field r->d.
// voxel_pos == nullptr: the voxel positions are the grid indices (i, j[, k]).
std::unique_ptr<binder_registration> new_registration(const float *voxel_pos, int Nx_v,
                                                      int Ny_v, int Nz_v, int Nx_n, int Ny_n,
                                                      int Nz_n, int spline_order,
                                                      double non_zero_voxels) {
  BINDER_REQUIRE(Nx_v >= 2 && Ny_v >= 2 && Nz_v >= 1, "bad fixed-grid shape");
  BINDER_REQUIRE(Nx_n >= 1 && Ny_n >= 1 && Nz_n >= 1, "bad moving-grid shape");
  BINDER_REQUIRE(spline_order >= 1 && spline_order <= 3, "spline_order must be 1, 2 or 3");
  BINDER_REQUIRE((Nz_v == 1) == (Nz_n == 1),
                 "fixed and moving images must both be 2D or both 3D");
  auto r = std::make_unique<binder_registration>();
  r->order = spline_order;
  r->nxv = Nx_v;
  r->nyv = Ny_v;
  r->nzv = Nz_v;
  r->nxn = Nx_n;
  r->nyn = Ny_n;
  r->nzn = Nz_n;
  r->N = Nz_v == 1 ? 2 : 3;
  r->is2d = Nz_v == 1;
  r->nv = (long long)Nx_v * Ny_v * Nz_v;
  // the sampling E-step addresses the [N][nv] fields with 32-bit offsets
  BINDER_REQUIRE(r->nv * 3 < (1ll << 30), "fixed grid too large (N * Nv must be below 2^30)");
  r->nn = (long long)Nx_n * Ny_n * Nz_n;
  r->nnz = non_zero_voxels;
  const size_t nf = (size_t)r->N * r->nv;
  r->x.alloc(nf);
  if (voxel_pos) {
    upload_planes(voxel_pos, r->x, r->stage, r->nv, r->N, r->s());
    // the smoothing passes generate x when it is the index grid
    DBuf<int> flag;
    const int one = 1;
    flag.upload(&one, 1, r->s());
    hipLaunchKernelGGL(k_check_index_grid, dim3(grid_blocks(r->nv)), dim3(256), 0, r->s(),
                       r->x.get(), Nx_v, Ny_v, Nz_v, r->N, flag.get());
    HIP_LAUNCH_OK();
    int got = 0;
    flag.download(&got, 1, r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
    r->x_index_grid = got == 1;
  } else {
    hipLaunchKernelGGL(k_index_grid, dim3(grid_blocks(r->nv)), dim3(256), 0, r->s(),
                       r->x.get(), Nx_v, Ny_v, Nz_v, r->N);
    HIP_LAUNCH_OK();
    r->x_index_grid = true;
  }
  r->d.alloc(nf);
  r->dream.alloc(nf);
  r->dream.zero(r->s());
  r->stats.alloc(16);
  r->stats.zero(r->s());
  // The sampling E-step: blocks of 256 threads, up to 2048 (8 per CU on 256
  // CUs).  The EM E-step (512-thread blocks with 64 KB of private LDS
  // columns each): about 2048 voxels (4 per thread) per block, at most 2048
  // and at least min(64, nv / 256) blocks, so that mid-size pyramid levels do
  // not spend their time setting up and flushing blocks of a few hundred
  // voxels.  Each block keeps one partial histogram row.
  long long blocks = (r->nv + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  long long em = (r->nv + 2047) / 2048;
  em = std::max(em, std::min(blocks, 64ll));
  if (em > 2048) em = 2048;
  r->estep_blocks = (int)em;
  r->sample_blocks = (int)blocks;
  r->hist_rows = (int)em;
  r->mllpart.alloc(std::max(blocks, em));
  HIP_OK(hipStreamSynchronize(r->s()));
  return r;
}

template <class T>
void swap_buffers(DBuf<T> &a, DBuf<T> &b) {
  std::swap(a.p, b.p);
  std::swap(a.n, b.n);
}


// Philox key of a sampler seed.
void sampler_key(int seed, uint32_t &k0, uint32_t &k1) {
  k0 = (uint32_t)seed;
  k1 = 0x85EBCA6Bu ^ (uint32_t)((long long)seed >> 32);
}

}  // namespace

extern "C" {

int binder_registration_create(binder_registration **out,
                               const float *final_locations, int threads,
                               const float *voxel_pos, const float *node_pos,
                               int Nx_v, int Ny_v, int Nz_v, int Nx_n,
                               int Ny_n, int Nz_n, int spline_order,
                               double non_zero_voxels) {
  (void)threads;
  (void)node_pos;
  return guarded([&] {
    BINDER_REQUIRE(out != nullptr, "out handle pointer is null");
    *out = nullptr;
    BINDER_REQUIRE(final_locations && voxel_pos, "null input array");
    auto r = new_registration(voxel_pos, Nx_v, Ny_v, Nz_v, Nx_n, Ny_n, Nz_n, spline_order,
                              non_zero_voxels);
    upload_planes(final_locations, r->d, r->stage, r->nv, r->N, r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
    *out = r.release();
  });
}

int binder_registration_create_resampled(binder_registration **out,
                                         binder_registration *src, double scale,
                                         const float *voxel_pos, int Nx_v, int Ny_v,
                                         int Nz_v, int Nx_n, int Ny_n, int Nz_n,
                                         int spline_order, double non_zero_voxels) {
  return guarded([&] {
    BINDER_REQUIRE(out != nullptr, "out handle pointer is null");
    *out = nullptr;
    BINDER_REQUIRE(src != nullptr, "null source handle");
    auto r = new_registration(voxel_pos, Nx_v, Ny_v, Nz_v, Nx_n, Ny_n, Nz_n, spline_order,
                              non_zero_voxels);
    BINDER_REQUIRE(r->N == src->N, "source and new grid differ in dimensionality");
    HIP_OK(hipStreamSynchronize(src->s()));
    resize_planes_device(src->dcur(), src->nxv, src->nyv, src->nzv, src->N, r->d.get(),
                         r->nxv, r->nyv, r->nzv, scale, r->s());
    *out = r.release();
  });
}

void binder_registration_destroy(binder_registration *reg) {
  if (reg) {
    (void)hipStreamSynchronize(reg->s());
    delete reg;
  }
}

}  // extern "C"

namespace {

// The likelihood set-up shared by the MI and SSD models: the fixed bins
// (bytes, [0, L)) and the moving bins (bytes packed four z per word, [0, K))
// on the device and the histogram buffers for K x L bins.  Leaves the
// uploads queued on the handle's stream (the caller synchronizes).
void set_bins(binder_registration *r, int K, int L, const int *binned_nodes,
              const int *binned_voxels) {
  BINDER_REQUIRE(binned_nodes && binned_voxels, "null input array");
  BINDER_REQUIRE(K >= 1 && K <= 256 && L >= 1 && L <= 256 && K * L <= 4096,
                 "K and L must be in [1,256] with K*L <= 4096");
  // the fixed and (3D) moving bins go up as bytes, converted by host
  // threads straight into page-locked memory and copied by one DMA each
  const long long nn8 = r->is2d ? 0 : r->nn;
  std::lock_guard<std::mutex> lock(g_upload_mu);
  uint8_t *pin = pinned_upload((size_t)(r->nv + nn8));
  std::atomic<bool> bad_v{false}, bad_n{false};
  parallel_ranges(r->nv + nn8, [&](long long b0, long long b1) {
    bool bv_ = false, bn_ = false;
    for (long long i = b0; i < b1; ++i) {
      if (i < r->nv) {
        const int b = binned_voxels[i];
        bv_ |= b < 0 || b >= L;
        pin[i] = (uint8_t)b;
      } else {
        const int b = binned_nodes[i - r->nv];
        bn_ |= b < 0 || b >= K;
        pin[i] = (uint8_t)b;
      }
    }
    if (bv_) bad_v = true;
    if (bn_) bad_n = true;
  });
  BINDER_REQUIRE(!bad_v, "binned_voxels value outside [0, L)");
  BINDER_REQUIRE(!bad_n, "binned_nodes value outside [0, K)");
  r->K = K;
  r->L = L;
  r->bv.alloc((size_t)r->nv);
  HIP_OK(hipMemcpyAsync(r->bv.get(), pin, (size_t)r->nv, hipMemcpyHostToDevice, r->s()));
  const int Xp = r->nxn + 2, Yp = r->nyn + 2;
  // the E-step indexes the padded moving grid with 32-bit offsets and
  // 24-bit multiplies
  const long long slab = r->is2d ? (long long)Yp : (long long)Yp * (r->nzn + 5);
  BINDER_REQUIRE(slab < (1ll << 24) && (long long)Xp * slab < (1ll << 30),
                 "moving grid too large for the E-step's 32-bit indexing");
  auto node_bin = [K](int b) {
    BINDER_REQUIRE(b >= 0 && b < K, "binned_nodes value outside [0, K)");
    return (uint8_t)b;
  };
  if (r->is2d) {
    std::vector<uint8_t> p((size_t)Xp * Yp, 0);
    for (int x = 0; x < r->nxn; ++x)
      for (int y = 0; y < r->nyn; ++y)
        p[(size_t)(x + 1) * Yp + (y + 1)] = node_bin(binned_nodes[(long long)x * r->nyn + y]);
    r->bins2d.upload(p.data(), p.size(), r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
  } else {
    // word zq of a column packs the bins of z = zq - 4 .. zq - 1 (byte c =
    // z - zq + 4), zero outside the grid: the bins go up as bytes and
    // k_pack_quads lays the words out on the device
    const int Zq = r->nzn + 5;
    r->nb8.alloc((size_t)r->nn);
    HIP_OK(hipMemcpyAsync(r->nb8.get(), pin + r->nv, (size_t)r->nn, hipMemcpyHostToDevice,
                          r->s()));
    const long long nq = (long long)Xp * Yp * Zq;
    r->quads.alloc((size_t)nq);
    hipLaunchKernelGGL(k_pack_quads, dim3(grid_blocks(nq)), dim3(256), 0, r->s(),
                       r->nb8.get(), r->nxn, r->nyn, r->nzn, r->quads.get());
    HIP_LAUNCH_OK();
  }
  r->Hd.alloc((size_t)K * L);
  r->Hd.zero(r->s());
  r->rowbuf.alloc((size_t)K * L + K + 64);
  {
    const int rows = std::max(r->estep_blocks, r->sample_blocks);
    r->hpart.alloc((size_t)rows * K * L);
    r->hpart2.alloc((size_t)((rows + kHistReduceRows - 1) / kHistReduceRows) * K * L);
  }
  // the page-locked bins must have gone up before the lock is released
  HIP_OK(hipStreamSynchronize(r->s()));
}

}  // namespace

extern "C" {

int binder_registration_set_mi_likelihood(binder_registration *r, double alpha,
                                          const double *theta, int K, int L,
                                          int threads, const int *binned_nodes,
                                          const int *binned_voxels) {
  (void)threads;
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    BINDER_REQUIRE(theta, "null input array");
    set_bins(r, K, L, binned_nodes, binned_voxels);
    r->alpha = alpha;
    r->lik = 0;
    std::vector<float> tt((size_t)K * L);
    for (int k = 0; k < K; ++k)
      for (int l = 0; l < L; ++l) tt[(size_t)l * K + k] = (float)theta[(size_t)k * L + l];
    r->theta.upload(theta, (size_t)K * L, r->s());
    r->thT.upload(tt.data(), tt.size(), r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
    r->has_mi = true;
  });
}

int binder_registration_set_ssd_likelihood(binder_registration *r, int bins, int threads,
                                           double sigma_sq, int scale_shift, double scale,
                                           double shift, const int *binned_nodes,
                                           const int *binned_voxels, double alpha_0,
                                           double beta_0) {
  (void)threads;
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    BINDER_REQUIRE(bins >= 2 && bins <= kSsdMaxBins, "SSD bins must be in [2, 64]");
    BINDER_REQUIRE(sigma_sq > 0.0, "sigma_sq must be positive");
    BINDER_REQUIRE(scale_shift == 0 || scale_shift == 1, "scale_shift must be 0 or 1");
    set_bins(r, bins, bins, binned_nodes, binned_voxels);
    r->lik = 1;
    r->ssd_ss = scale_shift;
    r->ssd_a0 = alpha_0;
    r->ssd_b0 = beta_0;
    // theta[0..3]: sigma^2, shift, scale, draw flag, in a bins^2 buffer (>= 4
    // for bins >= 2) so the EM and sampler buffers of theta / thT keep their
    // sizes
    std::vector<double> par((size_t)bins * bins, 0.0);
    par[0] = sigma_sq;
    par[1] = scale_shift ? shift : 0.0;
    par[2] = scale_shift ? scale : 1.0;
    r->theta.upload(par.data(), par.size(), r->s());
    r->thT.alloc((size_t)bins * bins);
    // the scale and shift as given are what get_likelihood_parameters
    // returns until an update (SSDLikelihood.cpp:328-334)
    r->ssd_given[0] = scale;
    r->ssd_given[1] = shift;
    HIP_OK(hipStreamSynchronize(r->s()));
    r->has_mi = true;
  });
}

int binder_registration_get_ssd_parameters(binder_registration *r, double *out3) {
  return guarded([&] {
    BINDER_REQUIRE(r && out3, "null argument");
    if (r->lik != 1) throw Error(BINDER_ESTATE, "setSSDLikelihood was not called");
    double par[kSsdParams];
    r->theta.download(par, kSsdParams, r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
    out3[0] = r->ssd_ss ? par[2] : r->ssd_given[0];
    out3[1] = r->ssd_ss ? par[1] : r->ssd_given[1];
    out3[2] = par[0];
  });
}

int binder_registration_set_affine_transformation(binder_registration *r, const double *A,
                                                  const double *t) {
  return guarded([&] {
    BINDER_REQUIRE(r && A && t, "null argument");
    r->lin = std::make_unique<LinearModel>(kLinAffine, r->N, r->nv, r->x.get(), A, t, r->s());
    r->sm.reset();
  });
}

int binder_registration_set_rigid_transformation(binder_registration *r, const double *R,
                                                 const double *t) {
  return guarded([&] {
    BINDER_REQUIRE(r && R && t, "null argument");
    r->lin = std::make_unique<LinearModel>(kLinRigid, r->N, r->nv, r->x.get(), R, t, r->s());
    r->sm.reset();
  });
}

int binder_registration_get_linear_parameters(binder_registration *r, double *out,
                                              int *count) {
  return guarded([&] {
    BINDER_REQUIRE(r && out, "null argument");
    if (!r->lin) throw Error(BINDER_ESTATE, "no affine or rigid transformation was set");
    const int n = r->lin->get(out);
    if (count) *count = n;
  });
}

int binder_registration_set_nonlinear_transformation(
    binder_registration *r, double gamma_x, double gamma_y, double gamma_z,
    double voxels_size_x, double voxels_size_y, double voxels_size_z,
    double nodes_size_x, double nodes_size_y, double nodes_size_z,
    const float *D) {
  return guarded([&] {
    BINDER_REQUIRE(r && D, "null argument");
    const double g[3] = {gamma_x, gamma_y, gamma_z};
    const double rho[3] = {voxels_size_x, voxels_size_y, voxels_size_z};
    const double eta[3] = {nodes_size_x, nodes_size_y, nodes_size_z};
    r->sm = std::make_unique<Smoother>(r->nxv, r->nyv, r->nzv, g, rho, eta, D,
                                       r->sigma_spline(), r->s());
    r->sm->set_index_grid(r->x_index_grid);
    r->lin.reset();  // the DCT prior replaces an affine / rigid model
  });
}

int binder_registration_set_nonlinear_transformation_dct_prior(
    binder_registration *r, double gamma_x, double gamma_y, double gamma_z,
    double voxels_size_x, double voxels_size_y, double voxels_size_z,
    double nodes_size_x, double nodes_size_y, double nodes_size_z, int order) {
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    BINDER_REQUIRE(order == 1 || order == 2, "finite-difference order must be 1 or 2");
    const double g[3] = {gamma_x, gamma_y, gamma_z};
    const double rho[3] = {voxels_size_x, voxels_size_y, voxels_size_z};
    const double eta[3] = {nodes_size_x, nodes_size_y, nodes_size_z};
    // the per-axis terms of binder_dct_prior_eigenvalues, same expressions
    const int dims[3] = {r->nxv, r->nyv, r->nzv};
    const int nd = r->is2d ? 2 : 3;
    std::vector<double> t((size_t)dims[0] + dims[1] + dims[2], 0.0);
    size_t o = 0;
    for (int a = 0; a < 3; ++a) {
      const double v2 = rho[a] * rho[a];
      for (int k = 0; k < dims[a]; ++k, ++o)
        if (a < nd) t[o] = (2.0 - 2.0 * std::cos(M_PI * (double)k / (double)dims[a])) / v2;
    }
    r->sm = std::make_unique<Smoother>(r->nxv, r->nyv, r->nzv, g, rho, eta, t.data(), order,
                                       r->sigma_spline(), r->s());
    r->sm->set_index_grid(r->x_index_grid);
    r->lin.reset();  // the DCT prior replaces an affine / rigid model
  });
}

int binder_registration_em(binder_registration *r, int max_it, double th,
                           int update, int debug) {
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    r->require_ready();
    r->trace.clear();
    if (max_it <= 0) return;
    const size_t nf = (size_t)r->N * r->nv, KL = (size_t)r->K * r->L;
    if (r->dslot != 0) {  // the current field goes back to slot 0
      HIP_OK(hipMemcpyAsync(r->d.get(), r->dcur(), nf * sizeof(float), hipMemcpyDeviceToDevice,
                            r->s()));
      r->dslot = 0;
    }
    r->tune_estep_em();
    r->hist_rows = r->estep_blocks;
    r->em_d.alloc(nf);
    r->em_dream.alloc(nf);
    r->em_stats.alloc(32);
    if (update) {
      r->em_theta.alloc(KL);
      r->em_thT.alloc(KL);
    }
    float *F[2] = {r->d.get(), r->em_d.get()};
    float *DR[2] = {r->dream.get(), r->em_dream.get()};
    // theta only changes with update_likelihood_parameters (MAP step)
    double *TH[2] = {r->theta.get(), update ? r->em_theta.get() : r->theta.get()};
    float *TT[2] = {r->thT.get(), update ? r->em_thT.get() : r->thT.get()};
    double *ST[2] = {r->em_stats.get(), r->em_stats.get() + 16};
    double *host = r->host_stats.p;  // pinned: parity p at host + 8 p
    auto enqueue = [&](int k) {
      const int p = k & 1, q = p ^ 1;
      EStepArgs a = r->estep_args(false, 0, 0, 0);
      a.d = F[p];
      a.thT = TT[p];
      if (a.ssd_par) a.ssd_par = TH[p];
      a.dream = DR[p];
      {
        TimedLaunch t(r->tm(), "estep", r->s());
        launch_estep(a, r->is2d, false, r->estep_blocks, r->s());
      }
      r->hist_theta(r->s(), update ? kThetaMap : kThetaKeep, 0, 0, 0, nullptr,
                    TH[update ? q : p], TT[update ? q : p], ST[p], TH[p]);
      SmoothOptions opt;
      opt.prior_out = ST[p] + 2;
      // the prior cost's sum also writes the three objective terms to the
      // page-locked host buffer (no copy launch)
      opt.host_stats = host + 8 * p;
      opt.stats_in = ST[p];
      r->transform(DR[p], F[q], opt);
      HIP_OK(hipEventRecord(r->em_done[p].e, r->s()));
    };
    double mlp_old = 1e100;
    int k = 0;
    bool stop = false;
    enqueue(0);
    while (!stop) {
      // queue the next iteration before reading this one's objective; if
      // this one turns out to be the last, the next one's outputs (in the
      // other parity) are simply not used
      if (k + 1 < max_it) enqueue(k + 1);
      HIP_OK(hipEventSynchronize(r->em_done[k & 1].e));
      const double *hs = host + 8 * (k & 1);
      const double mll = hs[0], pt = hs[1], pd = hs[2];
      const double mlp = mll + pd + pt;
      r->trace.insert(r->trace.end(), {mll, pt, pd, mlp});
      if (debug)
        std::printf("EM %d: mll %.8f prior_theta %.8f prior_d %.8f mlp %.8f "
                    "per nonzero voxel %.8f\n",
                    k + 1, mll, pt, pd, mlp, mlp / r->nnz);
      if ((mlp_old - mlp) / r->nnz < th) stop = true;
      mlp_old = mlp;
      ++k;
      if (k >= max_it) stop = true;
    }
    HIP_OK(hipStreamSynchronize(r->s()));
    // k iterations counted: the state after them is parity k & 1, the dream
    // locations are those of iteration k - 1 (parity (k - 1) & 1)
    if (k & 1) {
      swap_buffers(r->d, r->em_d);
      if (update) {
        swap_buffers(r->theta, r->em_theta);
        swap_buffers(r->thT, r->em_thT);
      }
    }
    if ((k - 1) & 1) swap_buffers(r->dream, r->em_dream);
  });
}

namespace {

// Puts `r` (or nothing) into slot `p` of a batch with an EM call of `cap`
// iterations ahead of it, preparing its double buffers as
// binder_registration_em does.
void batch_fill(binder_batch *b, int p, binder_registration *r, int cap, int update) {
  binder_batch::Slot &q = b->slots[p];
  q = binder_batch::Slot{};
  for (auto &A : b->set) A.valid = false;
  if (!r) return;
  r->require_ready();
  BINDER_REQUIRE(estep_em32_supported(r->K), "batched EM needs K <= 32 moving bins");
  BINDER_REQUIRE(r->lik == 0 && !r->lin && r->sm,
                 "batched EM runs the MI likelihood with the DCT prior");
  if (b->K < 0) {
    b->K = r->K;
    b->L = r->L;
    b->is2d = r->is2d;
  }
  BINDER_REQUIRE(r->K == b->K && r->L == b->L,
                 "all registrations of a batch need the same K and L");
  BINDER_REQUIRE(r->is2d == b->is2d, "a batch is either all 2D or all 3D");
  for (int o = 0; o < (int)b->slots.size(); ++o)
    BINDER_REQUIRE(o == p || b->slots[o].r != r, "a registration appears twice in the batch");
  const size_t KL = (size_t)r->K * r->L;
  r->trace.clear();
  const size_t nf = (size_t)r->N * r->nv;
  if (r->dslot != 0) {
    HIP_OK(hipMemcpyAsync(r->d.get(), r->dcur(), nf * sizeof(float), hipMemcpyDeviceToDevice,
                          r->s()));
    r->dslot = 0;
  }
  r->tune_estep_em();
  r->hist_rows = r->estep_blocks;
  r->em_d.alloc(nf);
  r->em_dream.alloc(nf);
  if (update) {
    r->em_theta.alloc(KL);
    r->em_thT.alloc(KL);
  }
  // the handle's own stream has finished its set-up work before the batch
  // stream reads the handle's buffers
  HIP_OK(hipStreamSynchronize(r->s()));
  q.r = r;
  q.cap = cap;
  q.done = cap <= 0;
  q.F[0] = r->d.get();
  q.F[1] = r->em_d.get();
  q.DR[0] = r->dream.get();
  q.DR[1] = r->em_dream.get();
  q.TH[0] = r->theta.get();
  q.TH[1] = update ? r->em_theta.get() : r->theta.get();
  q.TT[0] = r->thT.get();
  q.TT[1] = update ? r->em_thT.get() : r->thT.get();
}

// The end of a slot's EM call: the state after q.k iterations is parity
// q.k & 1, the dream locations those of iteration q.k - 1.
void batch_finish(binder_batch::Slot &q, int update) {
  binder_registration *r = q.r;
  const int kp = q.k;
  if (kp == 0) return;
  if (kp & 1) {
    swap_buffers(r->d, r->em_d);
    if (update) {
      swap_buffers(r->theta, r->em_theta);
      swap_buffers(r->thT, r->em_thT);
    }
  }
  if ((kp - 1) & 1) swap_buffers(r->dream, r->em_dream);
}

// Runs the EM iterations of every unfinished slot in lockstep until at
// least `min_finished` slots finish in this call (or all have); finished[p]
// receives the iteration count of a slot that finished in this call, else
// -1.  Iteration k + 1 is queued before iteration k's objectives are read,
// and the last queued iteration is drained before returning, so a slot left
// unfinished resumes exactly where it stands.
void batch_run(binder_batch *b, double th, int min_finished, int *finished) {
  const int P = (int)b->slots.size();
  const int update = b->update;
  const size_t KL = (size_t)std::max(b->K, 0) * std::max(b->L, 0);
  for (int p = 0; p < P; ++p) finished[p] = -1;
  double *dstats = b->stats.get();
  // argument set of call-iteration parity `ip` for `list` (each slot at its
  // own buffer parity q.k & 1); rebuilt only when the list changes, once the
  // iteration that last used it has completed
  auto build = [&](int ip, const std::vector<int> &list) {
    binder_batch::ArgSet &A = b->set[ip];
    std::vector<long long> key;
    for (int p : list) key.push_back(((long long)p << 1) | (b->slots[p].k & 1));
    if (A.valid && A.key == key) return;
    const int n = (int)list.size();
    int nwork = 0, max_rows = 0;
    for (int p : list) {
      nwork += b->slots[p].r->estep_blocks;
      max_rows = std::max(max_rows, b->slots[p].r->estep_blocks);
    }
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    A.off_h = align(sizeof(EStepArgs) * n);
    A.off_t = A.off_h + align(sizeof(HistReduceArgs) * n);
    A.off_w = A.off_t + align(sizeof(ThetaArgs) * n);
    const size_t bytes = A.off_w + sizeof(uint32_t) * (size_t)nwork;
    if (A.pin_cap < bytes) {
      if (A.pin) HIP_OK(hipHostFree(A.pin));
      A.pin = nullptr;
      HIP_OK(hipHostMalloc((void **)&A.pin, bytes, 0));
      A.pin_cap = bytes;
    }
    if (A.dev.n < bytes) A.dev.alloc(bytes);
    auto *ea = reinterpret_cast<EStepArgs *>(A.pin);
    auto *ha = reinterpret_cast<HistReduceArgs *>(A.pin + A.off_h);
    auto *ta = reinterpret_cast<ThetaArgs *>(A.pin + A.off_t);
    auto *wk = reinterpret_cast<uint32_t *>(A.pin + A.off_w);
    std::vector<SmoothBatch::Job> jobs;
    A.big.clear();
    int w = 0;
    for (int j = 0; j < n; ++j) {
      const int p = list[j];
      const binder_batch::Slot &x = b->slots[p];
      binder_registration *r = x.r;
      const int par = x.k & 1, q = par ^ 1;
      EStepArgs e = r->estep_args(false, 0, 0, 0);
      e.d = x.F[par];
      e.thT = x.TT[par];
      e.dream = x.DR[par];
      e.nblocks = r->estep_blocks;
      ea[j] = e;
      ha[j] = HistReduceArgs{r->hpart.get(), r->hpart2.get(), r->Hd.get(), r->estep_blocks,
                             (int)KL};
      ta[j] = r->theta_args(update ? kThetaMap : kThetaKeep, 0, 0, 0, nullptr,
                            x.TH[update ? q : par], x.TT[update ? q : par],
                            dstats + ((size_t)ip * P + p) * 8);
      for (int lb = 0; lb < r->estep_blocks; ++lb) wk[w++] = ((uint32_t)j << 16) | (uint32_t)lb;
      if (r->sm->batchable_em())
        jobs.push_back(SmoothBatch::Job{r->sm.get(), x.DR[par], r->x.get(), x.F[q],
                                        dstats + ((size_t)ip * P + p) * 8 + 2});
      else
        A.big.push_back(j);
    }
    HIP_OK(hipMemcpyAsync(A.dev.get(), A.pin, bytes, hipMemcpyHostToDevice, b->stream.s));
    A.any_small = !jobs.empty();
    if (A.any_small) A.smooth.build(jobs, b->stream.s);
    A.key = key;
    A.nwork = nwork;
    A.max_rows = max_rows;
    A.valid = true;
  };
  auto enqueue = [&](int ip, const std::vector<int> &list) {
    build(ip, list);
    binder_batch::ArgSet &A = b->set[ip];
    const int n = (int)list.size();
    hipStream_t bs = b->stream.s;
    unsigned char *dev = A.dev.get();
    bool cubic = true;
    for (int p : list) cubic = cubic && b->slots[p].r->order == 3;
    launch_estep_em_batch(reinterpret_cast<const EStepArgs *>(dev),
                          reinterpret_cast<const uint32_t *>(dev + A.off_w), A.nwork, b->is2d,
                          cubic, b->K, b->L, bs);
    launch_hist_reduce_batch(reinterpret_cast<const HistReduceArgs *>(dev + A.off_h), n,
                             A.max_rows, (int)KL, bs);
    launch_theta_batch(reinterpret_cast<const ThetaArgs *>(dev + A.off_t), n, bs);
    const int nbig = (int)A.big.size();
    const int np = std::min(nbig, kBatchPool);
    if (np > 0) {
      HIP_OK(hipEventRecord(b->theta_done[ip].e, bs));
      for (int i = 0; i < np; ++i) HIP_OK(hipStreamWaitEvent(b->pool[i].s, b->theta_done[ip].e, 0));
    }
    if (A.any_small) A.smooth.launch(bs);
    for (int i = 0; i < nbig; ++i) {
      const int p = list[A.big[i]];
      const binder_batch::Slot &x = b->slots[p];
      const int par = x.k & 1;
      SmoothOptions opt;
      opt.prior_out = dstats + ((size_t)ip * P + p) * 8 + 2;
      opt.exec_stream = b->pool[i % kBatchPool].s;
      x.r->sm->run(x.DR[par], x.r->x.get(), x.F[par ^ 1], opt, nullptr);
    }
    for (int i = 0; i < np; ++i) {
      HIP_OK(hipEventRecord(b->pool_done[ip][i].e, b->pool[i].s));
      HIP_OK(hipStreamWaitEvent(bs, b->pool_done[ip][i].e, 0));
    }
    HIP_OK(hipMemcpyAsync(b->host + (size_t)ip * P * 8, dstats + (size_t)ip * P * 8,
                          sizeof(double) * P * 8, hipMemcpyDeviceToHost, bs));
    HIP_OK(hipEventRecord(b->iter_done[ip].e, bs));
  };
  // objective of list's iteration at call-iteration parity ip: each slot's
  // stop test; returns the slots that go on
  int nfin = 0;
  auto read = [&](int ip, const std::vector<int> &list) {
    std::vector<int> still;
    for (int p : list) {
      binder_batch::Slot &x = b->slots[p];
      binder_registration *r = x.r;
      const double *hs = b->host + ((size_t)ip * P + p) * 8;
      const double mll = hs[0], pt = hs[1], pd = hs[2];
      const double mlp = mll + pd + pt;
      r->trace.insert(r->trace.end(), {mll, pt, pd, mlp});
      bool stop = (x.mlp_old - mlp) / r->nnz < th;
      x.mlp_old = mlp;
      x.k += 1;
      if (x.k >= x.cap) stop = true;
      if (stop) {
        x.done = true;
        finished[p] = x.k;
        ++nfin;
        batch_finish(x, update);
      } else {
        still.push_back(p);
      }
    }
    return still;
  };
  std::vector<int> cur;
  for (int p = 0; p < P; ++p)
    if (b->slots[p].r && !b->slots[p].done) cur.push_back(p);
  if (cur.empty()) return;
  if (min_finished > (int)cur.size()) min_finished = (int)cur.size();
  int ip = 0;
  enqueue(ip, cur);
  while (true) {
    // queue the next iteration (each slot's k + 1) before reading this one;
    // a slot that stops now has that iteration computed into buffers that
    // are not used (as in binder_registration_em), so its state was already
    // fixed by batch_finish in read() -- the queued iteration writes only
    // the other parity
    std::vector<int> nxt;
    for (int p : cur)
      if (b->slots[p].k + 1 < b->slots[p].cap) nxt.push_back(p);
    const bool more = nfin < min_finished;  // decided before this read
    if (!nxt.empty() && more) {
      // the queued iteration's slots are at buffer parity k + 1: advance
      // their counters for the argument build, then restore
      for (int p : nxt) b->slots[p].k += 1;
      enqueue(ip ^ 1, nxt);
      for (int p : nxt) b->slots[p].k -= 1;
    }
    HIP_OK(hipEventSynchronize(b->iter_done[ip].e));
    std::vector<int> still = read(ip, cur);
    if (!more || nxt.empty()) {
      if (nxt.empty() || still.empty()) break;
      // drain: the queued iteration was not enqueued (more == false), so
      // every slot in `still` stands after its iteration k: return
      break;
    }
    // slots still running whose next iteration is queued
    std::vector<int> go;
    for (int p : still)
      if (std::find(nxt.begin(), nxt.end(), p) != nxt.end()) go.push_back(p);
    cur = go;
    ip ^= 1;
    if (cur.empty()) {
      HIP_OK(hipStreamSynchronize(b->stream.s));
      break;
    }
  }
  HIP_OK(hipStreamSynchronize(b->stream.s));
  for (int i = 0; i < kBatchPool; ++i) HIP_OK(hipStreamSynchronize(b->pool[i].s));
}

}  // namespace

int binder_batch_create(binder_batch **out, int npairs, binder_registration *const *regs) {
  return guarded([&] {
    BINDER_REQUIRE(out && npairs >= 1, "need at least one slot");
    BINDER_REQUIRE(npairs < 65536, "at most 65535 slots per batch");
    auto b = std::make_unique<binder_batch>();
    b->slots.resize(npairs);
    for (int p = 0; p < npairs; ++p)
      if (regs && regs[p]) batch_fill(b.get(), p, regs[p], 0, 1);
    b->stats.alloc((size_t)2 * npairs * 8);
    HIP_OK(hipHostMalloc((void **)&b->host, sizeof(double) * 2 * npairs * 8, 0));
    *out = b.release();
  });
}

void binder_batch_destroy(binder_batch *b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream.s);
  delete b;
}

int binder_batch_set(binder_batch *b, int slot, binder_registration *reg, int max_iterations,
                     int update) {
  return guarded([&] {
    BINDER_REQUIRE(b, "null batch");
    BINDER_REQUIRE(slot >= 0 && slot < (int)b->slots.size(), "slot out of range");
    bool others = false;
    for (int p = 0; p < (int)b->slots.size(); ++p)
      if (p != slot && b->slots[p].r && !b->slots[p].done) others = true;
    BINDER_REQUIRE(!others || (update != 0) == (b->update != 0),
                   "all running slots of a batch share update_likelihood_parameters");
    b->update = update ? 1 : 0;
    batch_fill(b, slot, reg, max_iterations, b->update);
  });
}

int binder_batch_run(binder_batch *b, double th, int min_finished, int *finished) {
  return guarded([&] {
    BINDER_REQUIRE(b && finished, "null argument");
    batch_run(b, th, min_finished, finished);
  });
}

int binder_batch_em(binder_batch *b, const int *max_iterations, double th, int update,
                    int *iterations_out) {
  return guarded([&] {
    BINDER_REQUIRE(b && max_iterations && iterations_out, "null argument");
    const int P = (int)b->slots.size();
    b->update = update ? 1 : 0;
    for (int p = 0; p < P; ++p) {
      BINDER_REQUIRE(b->slots[p].r, "binder_batch_em needs every slot filled");
      batch_fill(b, p, b->slots[p].r, max_iterations[p], b->update);
    }
    std::vector<int> fin(P);
    batch_run(b, th, P, fin.data());
    for (int p = 0; p < P; ++p) iterations_out[p] = b->slots[p].k;
  });
}

int binder_registration_e_step(binder_registration *r, double *mll,
                               double *hist_out) {
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    r->require_ready();
    r->e_and_theta(false, kThetaKeep, 0, 0, 0);
    HIP_OK(hipMemcpyAsync(r->host_stats.p, r->stats.get(), sizeof(double),
                          hipMemcpyDeviceToHost, r->s()));
    if (hist_out)
      HIP_OK(hipMemcpyAsync(hist_out, r->Hd.get(), sizeof(double) * r->K * r->L,
                            hipMemcpyDeviceToHost, r->s()));
    HIP_OK(hipStreamSynchronize(r->s()));
    if (mll) *mll = r->host_stats.p[0];
  });
}

int binder_registration_sampler_run(binder_registration *r, int seed, int N_b,
                                    int N_s, int sample_gamma, int every,
                                    int sp, int sl, int st) {
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    BINDER_REQUIRE(N_b >= 0 && N_s >= 0, "sweep counts must be non-negative");
    r->require_ready();
    const int N = r->N;
    const int npairs = N * (N + 1) / 2;
    r->mean.alloc((size_t)N * r->nv);
    r->m2.alloc((size_t)npairs * r->nv);
    if (N_s == 0) {
      // no Welford launch: the result is zero (otherwise the first launch
      // writes every voxel's accumulators without reading them)
      r->mean.zero(r->s());
      r->m2.zero(r->s());
    }
    r->samples = N_s;
    if (sp) r->tune_estep_sample();
    uint32_t k0, k1;
    sampler_key(seed, k0, k1);
    if (st) r->prepare_ring(N_b + N_s);
    r->enqueue_sweeps(N_b, N_s, sample_gamma, every, sp, sl, st, k0, k1, nullptr,
                      r->mean.get(), r->m2.get());
    HIP_OK(hipStreamSynchronize(r->s()));
  });
}

int binder_registration_sampler_chains(binder_registration *r, int nchains, const int *seeds,
                                       int N_b, int N_s, int sp, int sl, int st,
                                       int use_graph) {
  return guarded([&] {
    BINDER_REQUIRE(r && seeds, "null argument");
    BINDER_REQUIRE(nchains >= 1 && N_b >= 0 && N_s >= 1,
                   "need nchains >= 1, N_b >= 0 and N_s >= 1");
    r->require_ready();
    BINDER_REQUIRE(!(use_graph && r->timer.enabled()),
                   "kernel timing cannot be recorded inside a captured graph");
    const int N = r->N;
    const int npairs = N * (N + 1) / 2;
    const size_t nf = (size_t)N * r->nv, KL = (size_t)r->K * r->L;
    r->mean.alloc(nf);
    r->m2.alloc((size_t)npairs * r->nv);
    if (sp) r->tune_estep_sample();
    if (st) r->prepare_ring(N_b + N_s);  // the current field is now slot 0
    // slot 0 keeps the start field unless the ring wraps onto it (more than
    // kWelfordBatch - 1 sweeps): only then does a chain restore it
    const bool wraps = st && N_b + N_s + 1 > kWelfordBatch;
    DBuf<float> d0;
    DBuf<double> th0;
    DBuf<float> thT0;
    th0.alloc(KL);
    thT0.alloc(KL);
    HIP_OK(hipMemcpyAsync(th0.get(), r->theta.get(), KL * sizeof(double),
                          hipMemcpyDeviceToDevice, r->s()));
    HIP_OK(hipMemcpyAsync(thT0.get(), r->thT.get(), KL * sizeof(float),
                          hipMemcpyDeviceToDevice, r->s()));
    if (wraps) {
      d0.alloc(nf);
      HIP_OK(hipMemcpyAsync(d0.get(), r->d.get(), nf * sizeof(float), hipMemcpyDeviceToDevice,
                            r->s()));
    }
    // chain c's key is sampler_key(seeds[c]): the same draws as
    // sampler_run(seeds[c], ...).  The sweeps use key (0, 0) ^ slot[0..1]
    // and fold samples with Welford counts offset by slot[2..3] (c * N_s), so
    // the accumulators see all chains' samples in order and one graph serves
    // every chain.
    std::vector<uint32_t> tab((size_t)4 * nchains);
    for (int c = 0; c < nchains; ++c) {
      sampler_key(seeds[c], tab[4 * c], tab[4 * c + 1]);
      const unsigned long long base = (unsigned long long)c * (unsigned long long)N_s;
      tab[4 * c + 2] = (uint32_t)base;
      tab[4 * c + 3] = (uint32_t)(base >> 32);
    }
    DBuf<uint32_t> ktab, kslot;
    ktab.upload(tab.data(), tab.size(), r->s());
    kslot.alloc(4);
    const uint32_t *keyx = kslot.get();
    const long long *n0b = reinterpret_cast<const long long *>(kslot.get() + 2);
    // no clearing of mean / M2: chain 0's first Welford launch (sample
    // offset 0) writes them without reading (N_s >= 1)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto release = [&] {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    };
    auto restore = [&] {
      if (wraps)
        HIP_OK(hipMemcpyAsync(r->d.get(), d0.get(), nf * sizeof(float),
                              hipMemcpyDeviceToDevice, r->s()));
      HIP_OK(hipMemcpyAsync(r->theta.get(), th0.get(), KL * sizeof(double),
                            hipMemcpyDeviceToDevice, r->s()));
      HIP_OK(hipMemcpyAsync(r->thT.get(), thT0.get(), KL * sizeof(float),
                            hipMemcpyDeviceToDevice, r->s()));
      r->dslot = 0;
    };
    try {
      for (int c = 0; c < nchains; ++c) {
        if (c > 0) restore();
        HIP_OK(hipMemcpyAsync(kslot.get(), ktab.get() + 4 * c, 4 * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, r->s()));
        if (!use_graph || c == 0) {
          // chain 0 runs eagerly: it also builds every lazily made buffer and
          // pass matrix, so that the capture below records kernels only
          r->enqueue_sweeps(N_b, N_s, 0, 1, sp, sl, st, 0u, 0u, keyx, r->mean.get(),
                            r->m2.get(), n0b);
        } else {
          if (!exec) {
            HIP_OK(hipStreamBeginCapture(r->s(), hipStreamCaptureModeThreadLocal));
            r->enqueue_sweeps(N_b, N_s, 0, 1, sp, sl, st, 0u, 0u, keyx, r->mean.get(),
                              r->m2.get(), n0b);
            HIP_OK(hipStreamEndCapture(r->s(), &graph));
            HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
          }
          HIP_OK(hipGraphLaunch(exec, r->s()));
        }
      }
      restore();
      HIP_OK(hipStreamSynchronize(r->s()));
    } catch (...) {
      hipGraph_t dummy = nullptr;
      hipStreamCaptureStatus st_cap = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(r->s(), &st_cap) == hipSuccess &&
          st_cap == hipStreamCaptureStatusActive)
        (void)hipStreamEndCapture(r->s(), &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      (void)hipStreamSynchronize(r->s());
      release();
      throw;
    }
    release();
    r->samples = nchains * N_s;
  });
}

int binder_registration_sampler_result(binder_registration *r,
                                       double *mean_out, double *cov_out) {
  return guarded([&] {
    BINDER_REQUIRE(r && mean_out && cov_out, "null argument");
    if (r->samples < 0 || !r->mean.get())
      throw Error(BINDER_ESTATE, "sampler_run has not been called");
    const int N = r->N;
    const int npairs = N * (N + 1) / 2;
    std::vector<double> mh((size_t)N * r->nv), m2h((size_t)npairs * r->nv);
    r->mean.download(mh.data(), mh.size(), r->s());
    r->m2.download(m2h.data(), m2h.size(), r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
    soa_to_aos(mh.data(), mean_out, r->nv, N);
    const double inv = 1.0 / (r->samples - 1);  // M2 / (N_s - 1) as the reference
    for (long long i = 0; i < r->nv; ++i) {
      int p = 0;
      for (int a = 0; a < N; ++a)
        for (int b = a; b < N; ++b, ++p) {
          const double v = m2h[(size_t)p * r->nv + i] * inv;
          cov_out[(i * N + a) * N + b] = v;
          cov_out[(i * N + b) * N + a] = v;
        }
    }
  });
}

int binder_registration_sampler(binder_registration *r, int seed, int N_b,
                                int N_s, int sample_gamma, int every, int sp,
                                int sl, int st, double *mean_out,
                                double *cov_out) {
  if (!mean_out || !cov_out) {
    set_last_error("null argument");
    return BINDER_EINVAL;
  }
  const int rc = binder_registration_sampler_run(r, seed, N_b, N_s, sample_gamma,
                                                 every, sp, sl, st);
  if (rc) return rc;
  return binder_registration_sampler_result(r, mean_out, cov_out);
}

// Spline(order, is2D).spline(d) / .log_spline(d)
// (BINDER/src/cpu/Spline.cpp:44-145, BINDER/src/cpu/bindings.cpp:436-445):
// Eigenvalues of the finite-difference prior in the DCT domain,
// D = (sum_a (2 - 2 cos(pi k_a / N_a)) / vs_a^2) ^ order over the first nd
// axes of (nx, ny, nz), summed x + y (+ z) in FP64 and rounded to float32:
// the values of BINDER/utils.py:331-349 (initialize_fourier_domain_smoothing,
// DCT branch) as the handle stores them, from a host loop over the separable
// terms instead of a full-grid NumPy evaluation.
int binder_dct_prior_eigenvalues(int nx, int ny, int nz, int nd, int order,
                                 const double *voxel_size, float *out) {
  return guarded([&] {
    BINDER_REQUIRE(out && voxel_size, "null argument");
    BINDER_REQUIRE(nd >= 1 && nd <= 3, "nd must be 1, 2 or 3");
    BINDER_REQUIRE(order >= 1, "order must be positive");
    const int dims[3] = {nx, nd >= 2 ? ny : 1, nd >= 3 ? nz : 1};
    BINDER_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "sizes must be positive");
    std::vector<double> t[3];
    for (int a = 0; a < 3; ++a) {
      t[a].assign(dims[a], 0.0);
      if (a >= nd) continue;
      const double v2 = voxel_size[a] * voxel_size[a];
      for (int k = 0; k < dims[a]; ++k)
        t[a][k] = (2.0 - 2.0 * std::cos(M_PI * (double)k / (double)dims[a])) / v2;
    }
    auto power = [order](double s) {
      return order == 1 ? s : (order == 2 ? s * s : std::pow(s, (double)order));
    };
    size_t o = 0;
    for (int x = 0; x < dims[0]; ++x)
      for (int y = 0; y < dims[1]; ++y) {
        const double sxy = nd >= 2 ? t[0][x] + t[1][y] : t[0][x];
        for (int z = 0; z < dims[2]; ++z)
          out[o++] = (float)power(nd >= 3 ? sxy + t[2][z] : sxy);
      }
  });
}

// B-spline weights of order 1, 2 or 3 at the order's node offsets for the
// fractional distance d, and their logarithms with -1e100 for a weight that
// is exactly zero.  Unused trailing entries are 0 (weights) / -1e100 (logs).
int binder_spline_weights(int order, double d, int log_space, double *out4) {
  return guarded([&] {
    BINDER_REQUIRE(out4, "null argument");
    BINDER_REQUIRE(order >= 1 && order <= 3, "spline order must be 1, 2 or 3");
    const double kNeg = -1e100;
    const double e = 1.0 - d;
    if (!log_space) {
      if (order == 3) {
        out4[0] = e * e * e / 6.0;
        out4[1] = 2.0 / 3.0 - d * d + d * d * d * 0.5;
        out4[2] = 2.0 / 3.0 - e * e + e * e * e * 0.5;
        out4[3] = d * d * d / 6.0;
      } else if (order == 2) {
        const double a = 1.0 - 2.0 * d, b = -2.0 * d - 1.0;
        out4[0] = 0.125 * a * a;
        out4[1] = 0.75 - d * d;
        out4[2] = 0.125 * b * b;
        out4[3] = 0.0;
      } else {
        out4[0] = e;
        out4[1] = d;
        out4[2] = out4[3] = 0.0;
      }
      return;
    }
    out4[0] = out4[1] = out4[2] = out4[3] = kNeg;
    if (order == 3) {
      out4[0] = 3.0 * std::log(e) - std::log(6.0);
      out4[1] = std::log(2.0 / 3.0 - d * d + d * d * d / 2.0);
      out4[2] = std::log(2.0 / 3.0 - e * e + e * e * e / 2.0);
      out4[3] = d < 1e-300 ? kNeg : 3.0 * std::log(d) - std::log(6.0);
    } else if (order == 2) {
      if (d <= 0.5) {
        const double a = 1.0 - 2.0 * d, b = -2.0 * d - 1.0;
        out4[0] = std::log(0.125 * a * a);
        out4[1] = std::log(0.75 - d * d);
        out4[2] = std::log(0.125 * b * b);
      }
    } else {
      out4[0] = d >= 1.0 ? kNeg : std::log(e);
      out4[1] = d < 1e-300 ? kNeg : std::log(d);
    }
  });
}

int binder_registration_correct_topology(binder_registration *r, double e1,
                                         double e2, int max_outer_it, int max_it,
                                         const double *voxel_sizes, double *stats) {
  return guarded([&] {
    BINDER_REQUIRE(r && voxel_sizes, "null argument");
    topology_on_device(r->dcur(), r->N, r->nxv, r->nyv, r->nzv, voxel_sizes, e1, e2,
                       max_outer_it, max_it, r->s(), stats);
  });
}

int binder_registration_set_field_from_dct(binder_registration *r, const float *coef,
                                           int ncoef, int fx, int fy, int fz, int res) {
  return guarded([&] {
    BINDER_REQUIRE(r && coef, "null argument");
    BINDER_REQUIRE(!r->is2d, "the DCT warm start is defined for 3D registrations");
    BINDER_REQUIRE(res >= 1 && fx >= 1 && fy >= 1 && fz >= 1, "bad grid or resolution");
    BINDER_REQUIRE((fx + res - 1) / res == r->nxv && (fy + res - 1) / res == r->nyv &&
                       (fz + res - 1) / res == r->nzv,
                   "the handle's grid is not ceil(full grid / res)");
    BINDER_REQUIRE(ncoef >= 1 && ncoef <= kMaxWarmCoef, "ncoef must be in [1, 16]");
    DBuf<float> c;
    c.upload(coef, (size_t)3 * ncoef * ncoef * ncoef, r->s());
    launch_dct_field(c.get(), ncoef, fx, fy, fz, res, r->x.get(), r->dcur(), r->nxv, r->nyv,
                     r->nzv, r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
  });
}

int binder_registration_create_from_dct(binder_registration **out, const float *coef,
                                        int ncoef, int fx, int fy, int fz, int res,
                                        const float *voxel_pos,
                                        int Nx_v, int Ny_v, int Nz_v, int Nx_n, int Ny_n,
                                        int Nz_n, int spline_order, double non_zero_voxels) {
  return guarded([&] {
    BINDER_REQUIRE(out != nullptr, "out handle pointer is null");
    *out = nullptr;
    BINDER_REQUIRE(coef != nullptr, "null coefficients");
    auto r = new_registration(voxel_pos, Nx_v, Ny_v, Nz_v, Nx_n, Ny_n, Nz_n, spline_order,
                              non_zero_voxels);
    binder_registration *h = r.get();
    const int rc = binder_registration_set_field_from_dct(h, coef, ncoef, fx, fy, fz, res);
    if (rc != BINDER_OK) throw Error(rc, binder_last_error());
    *out = r.release();
  });
}

int binder_registration_get_final_locations(binder_registration *r,
                                            float *out) {
  return guarded([&] {
    BINDER_REQUIRE(r && out, "null argument");
    download_planes(r->dcur(), out, r->stage, r->nv, r->N, r->s());
  });
}

int binder_registration_get_dream_locations(binder_registration *r,
                                            float *out) {
  return guarded([&] {
    BINDER_REQUIRE(r && out, "null argument");
    download_planes(r->dream.get(), out, r->stage, r->nv, r->N, r->s());
  });
}

int binder_registration_set_dream_locations(binder_registration *r,
                                            const float *in) {
  return guarded([&] {
    BINDER_REQUIRE(r && in, "null argument");
    upload_planes(in, r->dream, r->stage, r->nv, r->N, r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
  });
}

int binder_registration_get_likelihood_parameters(binder_registration *r,
                                                  double *theta_out) {
  return guarded([&] {
    BINDER_REQUIRE(r && theta_out, "null argument");
    if (!r->has_mi) throw Error(BINDER_ESTATE, "setMILikelihood was not called");
    if (r->lik == 1) {
      // SSDLikelihood::get_parameters (:328-334): [scale, shift]
      double p3[3];
      const int rc = binder_registration_get_ssd_parameters(r, p3);
      if (rc != BINDER_OK) throw Error(rc, binder_last_error());
      theta_out[0] = p3[0];
      theta_out[1] = p3[1];
      return;
    }
    r->theta.download(theta_out, (size_t)r->K * r->L, r->s());
    HIP_OK(hipStreamSynchronize(r->s()));
  });
}

int binder_registration_get_gamma(binder_registration *r, double *gamma3) {
  return guarded([&] {
    BINDER_REQUIRE(r && gamma3, "null argument");
    if (!r->sm) throw Error(BINDER_ESTATE, "setNonLinearTransformation was not called");
    r->sm->get_gamma(gamma3);
  });
}

int binder_registration_get_shape(binder_registration *r, int *shape4) {
  return guarded([&] {
    BINDER_REQUIRE(r && shape4, "null argument");
    shape4[0] = r->nxv;
    shape4[1] = r->nyv;
    shape4[2] = r->nzv;
    shape4[3] = r->N;
  });
}

int binder_registration_get_objective_trace(binder_registration *r,
                                            double *out, int *rows) {
  return guarded([&] {
    BINDER_REQUIRE(r && rows, "null argument");
    const int avail = (int)(r->trace.size() / 4);
    const int n = *rows < avail ? *rows : avail;
    if (out && n > 0) std::memcpy(out, r->trace.data(), sizeof(double) * 4 * n);
    *rows = out ? n : avail;
  });
}

int binder_registration_set_profiling(binder_registration *r, int enable) {
  return guarded([&] {
    BINDER_REQUIRE(r, "null handle");
    HIP_OK(hipStreamSynchronize(r->s()));
    r->timer.enable(enable != 0);
  });
}

int binder_registration_get_kernel_times(binder_registration *r, char *names,
                                         int names_cap, double *total_ms,
                                         long long *launches, int *count) {
  return guarded([&] {
    BINDER_REQUIRE(r && count, "null argument");
    HIP_OK(hipStreamSynchronize(r->s()));
    std::vector<std::string> nm;
    std::vector<double> ms;
    std::vector<long long> c;
    r->timer.summary(nm, ms, c);
    const int n = (int)nm.size() < *count ? (int)nm.size() : *count;
    std::string joined;
    for (int i = 0; i < n; ++i) {
      if (total_ms) total_ms[i] = ms[i];
      if (launches) launches[i] = c[i];
      joined += nm[i];
      joined += '\n';
    }
    if (names && names_cap > 0) {
      std::strncpy(names, joined.c_str(), (size_t)names_cap - 1);
      names[names_cap - 1] = '\0';
    }
    *count = n;
  });
}

// ---------------------------------------------------------------------------
// NonLinearTransformation
// ---------------------------------------------------------------------------
int binder_nlt_create(binder_nlt **out, double gamma_x, double gamma_y,
                      double gamma_z, double vsx, double vsy, double vsz,
                      double nsx, double nsy, double nsz, const float *D,
                      int Nx_v, int Ny_v, int Nz_v, int threads,
                      double sigma_spline, double spline_offset) {
  (void)threads;
  (void)spline_offset;
  return guarded([&] {
    BINDER_REQUIRE(out && D, "null argument");
    *out = nullptr;
    auto n = std::make_unique<binder_nlt>();
    n->nx = Nx_v;
    n->ny = Ny_v;
    n->nz = Nz_v;
    n->N = Nz_v == 1 ? 2 : 3;
    n->nv = (long long)Nx_v * Ny_v * Nz_v;
    const double g[3] = {gamma_x, gamma_y, gamma_z};
    const double rho[3] = {vsx, vsy, vsz};
    const double eta[3] = {nsx, nsy, nsz};
    n->sm = std::make_unique<Smoother>(Nx_v, Ny_v, Nz_v, g, rho, eta, D,
                                       sigma_spline, n->stream.s);
    n->in.alloc((size_t)n->N * n->nv);
    n->out.alloc((size_t)n->N * n->nv);
    n->scal.alloc(2);
    *out = n.release();
  });
}

void binder_nlt_destroy(binder_nlt *n) {
  if (n) {
    (void)hipStreamSynchronize(n->stream.s);
    delete n;
  }
}

int binder_nlt_smooth_deformation(binder_nlt *n, const float *deformation,
                                  float *out, double *cost) {
  return guarded([&] {
    BINDER_REQUIRE(n && deformation && out, "null argument");
    std::vector<float> h((size_t)n->N * n->nv);
    aos_to_soa(deformation, h.data(), n->nv, n->N);
    n->in.upload(h.data(), h.size(), n->stream.s);
    SmoothOptions o;
    o.prior_out = n->scal.get();
    n->sm->run(n->in.get(), nullptr, n->out.get(), o, nullptr);
    n->out.download(h.data(), h.size(), n->stream.s);
    double c = 0.0;
    HIP_OK(hipMemcpyAsync(&c, n->scal.get(), sizeof(double),
                          hipMemcpyDeviceToHost, n->stream.s));
    HIP_OK(hipStreamSynchronize(n->stream.s));
    soa_to_aos(h.data(), out, n->nv, n->N);
    if (cost) *cost = c;
  });
}

int binder_nlt_sample_deformation_field(binder_nlt *n, const float *deformation,
                                        int S, unsigned long long seed,
                                        float *out) {
  return guarded([&] {
    BINDER_REQUIRE(n && deformation && out, "null argument");
    BINDER_REQUIRE(S >= 0, "number_of_samples must be non-negative");
    std::vector<float> h((size_t)n->N * n->nv);
    aos_to_soa(deformation, h.data(), n->nv, n->N);
    n->in.upload(h.data(), h.size(), n->stream.s);
    for (int s = 0; s < S; ++s) {
      SmoothOptions o;
      o.sample = true;
      o.key0 = (uint32_t)seed;
      o.key1 = (uint32_t)(seed >> 32) ^ 0x27D4EB2Fu;
      o.sweep = (uint32_t)s;
      o.stream = kStreamNltSample;
      o.prior_out = n->scal.get();
      n->sm->run(n->in.get(), nullptr, n->out.get(), o, nullptr);
      n->out.download(h.data(), h.size(), n->stream.s);
      HIP_OK(hipStreamSynchronize(n->stream.s));
      soa_to_aos(h.data(), out + (size_t)s * n->N * n->nv, n->nv, n->N);
    }
  });
}

int binder_nlt_up_sample(const double *field, int N_x, int N_y, int N_z,
                         int N_x_o, int N_y_o, int N_z_o, int threads,
                         double *out) {
  (void)threads;
  return guarded([&] {
    BINDER_REQUIRE(field && out, "null argument");
    BINDER_REQUIRE(N_x >= 2 && N_y >= 2 && N_z >= 1 && N_x_o >= 2 && N_y_o >= 2 &&
                       N_z_o >= 1 && (N_z == 1) == (N_z_o == 1),
                   "bad input/output grid shape");
    BINDER_REQUIRE(N_x <= 256 && N_y <= 256 && N_z <= 256 && N_x_o <= 256 &&
                       N_y_o <= 256 && N_z_o <= 256,
                   "each axis must be at most 256 voxels");
    const int N = N_z_o == 1 ? 2 : 3;
    const long long nvi = (long long)N_x * N_y * N_z;
    const long long nvo = (long long)N_x_o * N_y_o * N_z_o;
    Stream st;
    MatrixCache mats;
    std::vector<float> h((size_t)N * nvi);
    aos_to_soa(field, h.data(), nvi, N);
    DBuf<float> in, o;
    in.upload(h.data(), h.size(), st.s);
    o.alloc((size_t)N * nvo);
    up_sample_field(in.get(), N_x, N_y, N_z, N_x_o, N_y_o, N_z_o, o.get(), st.s, mats);
    std::vector<float> ho((size_t)N * nvo);
    o.download(ho.data(), ho.size(), st.s);
    HIP_OK(hipStreamSynchronize(st.s));
    soa_to_aos(ho.data(), out, nvo, N);
  });
}

}  // extern "C"
