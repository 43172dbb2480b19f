// Per-axis DCT contractions for the deformation M-step / d-sample.
//
// A component's displacement t = dream - x is transformed with the separable
// orthonormal DCT-II: one pass per axis, each a batched matrix product whose
// operands are a dense transform matrix and the data (read with the axis
// either as the contiguous dimension -- "data-left": rows x n times n x n --
// or as a strided one -- "matrix-left": n x n times n x columns).  The pass
// that completes the forward transform applies the prior filter
// p = c / (1 + gamma sigma^2 D) (+ sqrt(sigma^2/(1+gamma sigma^2 D)) * eps when
// sampling) and accumulates the prior cost 0.5 gamma D p^2 per block; the
// inverse passes apply the orthonormal DCT-III and the last one adds x back.
// Reference semantics: BINDER/src/cpu/NonLinearTransformation.cpp:58-126,
// :128-223, :344-364, :433-451; ortho convention BINDER/src/cpu/DCT.cpp:83-175.
#include <cmath>

#include "smoothing.hpp"

namespace binder {

namespace {

constexpr int BM = 64, BN = 64, BK = 16, kThreads = 256;

enum : int { EPI_STORE = 0, EPI_FILTER = 1, EPI_ADDX = 2 };

struct PassArgs {
  const float *A, *B;
  float *C;
  int M, N, K, lda, ldb, ldc;
  long long sAc, sBc, sCc;  // per-component strides
  long long sAb, sBb, sCb;  // per inner-batch strides
  int nb;                   // inner batches per component
  const float *X;           // prologue: data operand minus X (same layout)
  const float *Xo;          // ADDX epilogue: + Xo (same layout as C)
  // filter epilogue
  const float *D;
  const double *gamma;
  double sigma2;
  double etarho[3];
  int noise, raw;
  uint32_t k0, k1, sweep, stream;
  double *cost_part, *raw_part;
};

template <bool DATA_LEFT, bool PRO, int EPI>
__global__ __launch_bounds__(kThreads) void k_axis_pass(PassArgs a) {
  __shared__ float As[BK][BM + 4];
  __shared__ float Bs[BK][BN + 4];
  const int tid = threadIdx.x;
  const int tn = tid & 15, tm = tid >> 4;
  const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const int comp = blockIdx.z / a.nb, inner = blockIdx.z % a.nb;
  const float *A = a.A + comp * a.sAc + inner * a.sAb;
  const float *B = a.B + comp * a.sBc + inner * a.sBb;
  const float *XA = nullptr, *XB = nullptr;
  if (PRO) {
    if (DATA_LEFT)
      XA = a.X + comp * a.sAc + inner * a.sAb;
    else
      XB = a.X + comp * a.sBc + inner * a.sBb;
  }
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;

  for (int k0 = 0; k0 < a.K; k0 += BK) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = tid + kThreads * i;
      const int r = e >> 4, c = e & 15;
      const int gm = m0 + r, gk = k0 + c;
      float v = 0.f;
      if (gm < a.M && gk < a.K) {
        const long long off = (long long)gm * a.lda + gk;
        v = A[off];
        if (PRO && DATA_LEFT) v -= XA[off];
      }
      As[c][r] = v;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = tid + kThreads * i;
      const int r = e >> 6, c = e & 63;
      const int gk = k0 + r, gn = n0 + c;
      float v = 0.f;
      if (gk < a.K && gn < a.N) {
        const long long off = (long long)gk * a.ldb + gn;
        v = B[off];
        if (PRO && !DATA_LEFT) v -= XB[off];
      }
      Bs[r][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      const float4 av = *reinterpret_cast<const float4 *>(&As[kk][tm * 4]);
      const float4 bv = *reinterpret_cast<const float4 *>(&Bs[kk][tn * 4]);
      const float ar[4] = {av.x, av.y, av.z, av.w};
      const float br[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(ar[i], br[j], acc[i][j]);
    }
    __syncthreads();
  }

  float *C = a.C + comp * a.sCc + inner * a.sCb;
  double cost = 0.0, raw = 0.0;
  double g = 0.0;
  if (EPI == EPI_FILTER) g = a.gamma[comp];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + tm * 4 + i;
    if (gm >= a.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tn * 4 + j;
      if (gn >= a.N) continue;
      const long long off = (long long)gm * a.ldc + gn;
      float v = acc[i][j];
      if (EPI == EPI_FILTER) {
        const long long o = inner * a.sCb + off;  // voxel index in component
        const float Dv = a.D[o];
        const float gs2 = (float)(g * a.sigma2);
        const float den = 1.0f + gs2 * Dv;
        float p = v * (1.0f / den);
        if (a.noise) {
          const u32x4 r = philox4x32_10(
              u32x4{(uint32_t)o, a.sweep, (uint32_t)comp, a.stream}, a.k0, a.k1);
          double e0, e1;
          box_muller(r.x, r.y, e0, e1);
          p += (float)e0 * sqrtf((float)a.sigma2 / den);
        }
        const double pd = (double)p;
        cost += 0.5 * g * (double)Dv * pd * pd;
        if (a.raw) raw += a.etarho[comp] * (double)Dv * pd * pd;
        v = p;
      } else if (EPI == EPI_ADDX) {
        if (a.Xo) v += a.Xo[comp * a.sCc + inner * a.sCb + off];
      }
      C[off] = v;
    }
  }
  if (EPI == EPI_FILTER) {
    __shared__ double red[2][kThreads];
    red[0][tid] = cost;
    red[1][tid] = raw;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if (tid < s) {
        red[0][tid] += red[0][tid + s];
        red[1][tid] += red[1][tid + s];
      }
      __syncthreads();
    }
    if (tid == 0) {
      const long long b =
          ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
      a.cost_part[b] = red[0][0];
      if (a.raw) a.raw_part[b] = red[1][0];
    }
  }
}

// Deterministic sum of `count` partials (fixed assignment and tree order).
__global__ __launch_bounds__(256) void k_sum_partials(const double *part,
                                                      long long count,
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

// nu ~ Gamma(alpha0 + N nv / 2, scale 1 / (beta0 + nv raw / 2)), then
// gamma_k = nu rho_x rho_y rho_z eta_k^2 nv
// (NonLinearTransformation.cpp:116-124, :140-148).
__global__ void k_draw_nu(const double *raw, double *gamma, int N, double nv,
                          double rho_prod, double e0, double e1, double e2,
                          uint32_t k0, uint32_t k1, uint32_t sweep) {
  if (threadIdx.x != 0) return;
  const double a_post = 1.0 + 0.5 * (double)N * nv;
  const double b_post = 10000.0 + 0.5 * nv * (*raw);
  const double g = gamma_mt(a_post, u32x4{0u, sweep, 0u, kStreamNu}, k0, k1);
  const double nu = g / b_post;
  gamma[0] = nu * rho_prod * e0 * e0 * nv;
  gamma[1] = nu * rho_prod * e1 * e1 * nv;
  gamma[2] = nu * rho_prod * e2 * e2 * nv;
}

template <bool DL, bool PRO, int EPI>
void launch(dim3 grid, hipStream_t s, const PassArgs &a) {
  hipLaunchKernelGGL((k_axis_pass<DL, PRO, EPI>), grid, dim3(kThreads), 0, s, a);
}

}  // namespace

Smoother::Smoother(int nx, int ny, int nz, const double gamma[3],
                   const double rho[3], const double eta[3],
                   const float *D_host, double sigma, hipStream_t stream)
    : nx_(nx), ny_(ny), nz_(nz), N_(nz == 1 ? 2 : 3),
      nv_((long long)nx * ny * nz), sigma_(sigma), stream_(stream) {
  BINDER_REQUIRE(nx >= 2 && ny >= 2 && nz >= 1,
                 "deformation grid must be at least 2 x 2 (x 1)");
  for (int i = 0; i < 3; ++i) {
    rho_[i] = rho[i];
    eta_[i] = eta[i];
  }
  nu_ = gamma[0] / (rho[0] * rho[1] * rho[2] * eta[0] * eta[0] * (double)nv_);
  D_.upload(D_host, nv_, stream_);
  gamma_.upload(gamma, 3, stream_);
  bufA_.alloc((size_t)N_ * nv_);
  bufB_.alloc((size_t)N_ * nv_);
  // partial sums of the filter pass (axis 0, matrix-left over ny*nz columns)
  const long long nparts =
      (long long)((ny * nz + BN - 1) / BN) * ((nx + BM - 1) / BM) * N_;
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

const AxisMatrices &Smoother::mats(int n) {
  auto it = mats_.find(n);
  if (it != mats_.end()) return *it->second;
  std::vector<float> f((size_t)n * n), t((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    const double s = k == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
    for (int j = 0; j < n; ++j) {
      const double c = s * std::cos(M_PI * (2.0 * j + 1.0) * k / (2.0 * n));
      f[(size_t)k * n + j] = (float)c;
      t[(size_t)j * n + k] = (float)c;
    }
  }
  auto m = std::make_unique<AxisMatrices>();
  m->fwd.upload(f.data(), f.size(), stream_);
  m->inv.upload(t.data(), t.size(), stream_);
  HIP_OK(hipStreamSynchronize(stream_));
  auto &ref = *m;
  mats_[n] = std::move(m);
  return ref;
}

// flags: bit0 = prologue (data - xpro), bits1-2 = epilogue kind.
void Smoother::launch_pass(int axis, bool forward, const float *in, float *out,
                           int flags, const float *xpro, const float *xepi,
                           const SmoothOptions &opt, KernelTimer *timer) {
  const bool pro = flags & 1;
  const int epi = (flags >> 1) & 3;
  const int inner_axis = nz_ > 1 ? 2 : 1;
  PassArgs a{};
  a.sigma2 = sigma_ * sigma_;
  a.D = D_.get();
  a.gamma = gamma_.get();
  a.X = xpro;
  a.Xo = xepi;
  a.noise = opt.sample ? 1 : 0;
  a.raw = opt.sample_gamma ? 1 : 0;
  const double rp = rho_[0] * rho_[1] * rho_[2];
  for (int i = 0; i < 3; ++i) a.etarho[i] = eta_[i] * eta_[i] * rp;
  a.k0 = opt.key0;
  a.k1 = opt.key1;
  a.sweep = opt.sweep;
  a.stream = opt.stream;
  a.cost_part = part_cost_.get();
  a.raw_part = part_raw_.get();
  const int dims[3] = {nx_, ny_, nz_};
  const int n = dims[axis];
  const AxisMatrices &m = mats(n);
  const bool data_left = (axis == inner_axis);
  if (data_left) {
    a.A = in;
    a.B = forward ? m.inv.get() : m.fwd.get();
    a.M = (int)(nv_ / n);
    a.N = n;
    a.K = n;
    a.lda = n;
    a.ldb = n;
    a.ldc = n;
    a.sAc = nv_;
    a.sBc = 0;
    a.sCc = nv_;
    a.nb = 1;
  } else {
    a.A = forward ? m.fwd.get() : m.inv.get();
    a.B = in;
    a.M = n;
    a.K = n;
    a.lda = n;
    a.sAc = 0;
    a.sBc = nv_;
    a.sCc = nv_;
    if (axis == 1) {  // 3D middle axis, batched over x
      a.N = nz_;
      a.ldb = nz_;
      a.ldc = nz_;
      a.nb = nx_;
      a.sBb = (long long)ny_ * nz_;
      a.sCb = (long long)ny_ * nz_;
    } else {  // outermost axis
      a.N = ny_ * nz_;
      a.ldb = ny_ * nz_;
      a.ldc = ny_ * nz_;
      a.nb = 1;
    }
  }
  a.C = out;
  dim3 grid((a.N + BN - 1) / BN, (a.M + BM - 1) / BM, N_ * a.nb);
  BINDER_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "grid too large for DCT pass");
  const char *names[3][2] = {{"dct_x_inv", "dct_x_fwd"},
                             {"dct_y_inv", "dct_y_fwd"},
                             {"dct_z_inv", "dct_z_fwd"}};
  TimedLaunch tl(timer, names[axis][forward ? 1 : 0], stream_);
  if (data_left) {
    if (pro) {
      if (epi == EPI_FILTER) launch<true, true, EPI_FILTER>(grid, stream_, a);
      else launch<true, true, EPI_STORE>(grid, stream_, a);
    } else if (epi == EPI_ADDX) launch<true, false, EPI_ADDX>(grid, stream_, a);
    else if (epi == EPI_FILTER) launch<true, false, EPI_FILTER>(grid, stream_, a);
    else launch<true, false, EPI_STORE>(grid, stream_, a);
  } else {
    if (pro) {
      if (epi == EPI_FILTER) launch<false, true, EPI_FILTER>(grid, stream_, a);
      else launch<false, true, EPI_STORE>(grid, stream_, a);
    } else if (epi == EPI_ADDX) launch<false, false, EPI_ADDX>(grid, stream_, a);
    else if (epi == EPI_FILTER) launch<false, false, EPI_FILTER>(grid, stream_, a);
    else launch<false, false, EPI_STORE>(grid, stream_, a);
  }
  HIP_LAUNCH_OK();
}

void Smoother::run(const float *dream, const float *x, float *out,
                   const SmoothOptions &opt, KernelTimer *timer) {
  // forward axes innermost first, inverse in the opposite order
  int axes[3], na = 0;
  if (nz_ > 1) axes[na++] = 2;
  axes[na++] = 1;
  axes[na++] = 0;
  const float *src = dream;
  float *bufs[2] = {bufA_.get(), bufB_.get()};
  int bi = 0;
  // Forward passes; the first subtracts x, the last filters.
  for (int i = 0; i < na; ++i) {
    int flags = 0;
    if (i == 0 && x) flags |= 1;
    if (i == na - 1) flags |= (EPI_FILTER << 1);
    launch_pass(axes[i], true, src, bufs[bi], flags, x, nullptr, opt, timer);
    src = bufs[bi];
    bi ^= 1;
  }
  const long long nparts =
      (long long)((ny_ * nz_ + BN - 1) / BN) * ((nx_ + BM - 1) / BM) * N_;
  {
    TimedLaunch tl(timer, "sum_partials", stream_);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream_,
                       part_cost_.get(), nparts,
                       opt.prior_out ? opt.prior_out : scal_.get());
    HIP_LAUNCH_OK();
  }
  if (opt.sample && opt.sample_gamma) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream_,
                       part_raw_.get(), nparts, scal_.get() + 1);
    HIP_LAUNCH_OK();
  }
  // Inverse passes; the last adds x and writes the output.
  for (int i = na - 1; i >= 0; --i) {
    const bool last = (i == 0);
    float *dst = last ? out : bufs[bi];
    const int flags = last ? (EPI_ADDX << 1) : 0;
    launch_pass(axes[i], false, src, dst, flags, nullptr, last ? x : nullptr,
                opt, timer);
    src = dst;
    bi ^= 1;
  }
  if (opt.sample && opt.sample_gamma) {
    hipLaunchKernelGGL(k_draw_nu, dim3(1), dim3(64), 0, stream_,
                       scal_.get() + 1, gamma_.get(), N_, (double)nv_,
                       rho_[0] * rho_[1] * rho_[2], eta_[0], eta_[1], eta_[2],
                       opt.key0, opt.key1, opt.sweep);
    HIP_LAUNCH_OK();
  }
}

}  // namespace binder
