// SSD likelihood and affine / rigid transformations (see models.hpp).
//
// SSD (BINDER/src/cpu/SSDLikelihood.cpp): the likelihood of a voxel with
// fixed bin u at a node with moving bin k (bin 0 outside the grid, the value
// the reference uses there, :60-86) depends on (u, k) only; the E-step
// kernels evaluate it per voxel relative to the voxel's closest node (exact
// however small sigma^2 is), writing the same soft joint histogram H[k][u]
// and FP64 mll partials as for the MI likelihood.  The reference's
// accumulate_info sums (:117-146) are moments of H, so the parameter step
// (update_shift_and_scale, update_sigma, sample_parameters, :169-307) runs
// on those moments in one block.
//
// Affine / rigid (AffineTransformation.cpp:25-130,
// RigidTransformation.cpp:57-154): the least-squares fit needs sum y and
// sum x y^T over the voxels (FP64 block partials in a fixed order), an
// (N+1) x (N+1) or N x N solve (one thread, FP64), and y = M x + t.
#include <cmath>
#include <vector>

#include "estep.hpp"
#include "models.hpp"

namespace binder {
namespace {

constexpr uint32_t kStreamSsd = 6u;  // Philox stream of the SSD parameter draws

constexpr int kSsdThreads = 256;

__global__ __launch_bounds__(kSsdThreads) void k_ssd_theta(SsdThetaArgs a) {
  __shared__ double H[kSsdMaxBins * kSsdMaxBins];
  __shared__ double red[6][kSsdThreads];
  __shared__ long long lred[kSsdThreads];
  __shared__ int lodd[kSsdThreads];
  const int tid = threadIdx.x, B = a.bins, BB = B * B;
  // mll: the E-step's partials, summed as the MI step sums them
  long long acc = 0;
  int odd = 0;
  double dsum = 0.0;
  if (a.mll_fixed) {
    const long long *fx = reinterpret_cast<const long long *>(a.mllpart);
    for (int i = tid; i < a.nmll; i += kSsdThreads) {
      const long long v = fx[i];
      if (v == kMllFixNaN) odd |= 2;
      else if (v == kMllFixZero) odd |= 1;
      else acc += v;
    }
  } else {
    for (int i = tid; i < a.nmll; i += kSsdThreads) dsum += a.mllpart[i];
  }
  lred[tid] = acc;
  lodd[tid] = odd;
  red[0][tid] = dsum;
  // the histogram
  for (int i = tid; i < BB; i += kSsdThreads) {
    double h;
    if (a.hsum) {
      unsigned long long s = 0;
      for (int r = 0; r < a.hrows; ++r) s += a.hsum[(long long)r * BB + i];
      h = (double)s * (1.0 / kHistScale);
      a.Hd[i] = h;
    } else {
      h = a.Hd[i];
    }
    H[i] = h;
  }
  __syncthreads();
  for (int w = kSsdThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      lred[tid] += lred[tid + w];
      lodd[tid] |= lodd[tid + w];
      red[0][tid] += red[0][tid + w];
    }
    __syncthreads();
  }
  double mll = 0.0;
  if (tid == 0)
    mll = a.mll_fixed ? (lodd[0] & 2 ? (double)NAN
                                     : (lodd[0] & 1 ? (double)INFINITY
                                                    : -kLn2 * ((double)lred[0] / kMllFixScale)))
                      : red[0][0];
  __syncthreads();
  // moments of H[k][u]: sum h, h k, h k^2, h u, h u k, h u^2
  double m[6] = {0, 0, 0, 0, 0, 0};
  for (int i = tid; i < BB; i += kSsdThreads) {
    const double h = H[i], k = (double)(i / B), u = (double)(i % B);
    m[0] += h;
    m[1] += h * k;
    m[2] += h * k * k;
    m[3] += h * u;
    m[4] += h * u * k;
    m[5] += h * u * u;
  }
  for (int j = 0; j < 6; ++j) red[j][tid] = m[j];
  __syncthreads();
  for (int w = kSsdThreads / 2; w > 0; w >>= 1) {
    if (tid < w)
      for (int j = 0; j < 6; ++j) red[j][tid] += red[j][tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    const double g = red[0][0], gk = red[1][0], gkk = red[2][0], gu = red[3][0],
                 guk = red[4][0], guu = red[5][0];
    double s2 = a.par_in[0], sh = a.par_in[1], sc = a.par_in[2], flag = a.par_in[3];
    if (a.mode != kSsdKeep) {
      if (a.mode == kSsdDraw) flag = 1.0;
      // with shift and scale the MAP step draws once the flag is set
      // (SSDLikelihood.cpp:211, :273); without them only a draw step draws
      // (:148-167 vs :284-307)
      const bool draw = a.shift_scale ? flag != 0.0 : a.mode == kSsdDraw;
      const uint32_t key0 = a.keyx ? a.k0 ^ a.keyx[0] : a.k0;
      const uint32_t key1 = a.keyx ? a.k1 ^ a.keyx[1] : a.k1;
      if (a.shift_scale) {
        // A = [[g, gk], [gk, gkk]] (outside nodes count as k = 0), b = [gu, guk]
        const double det = g * gkk - gk * gk;
        const double i00 = gkk / det, i01 = -gk / det, i11 = g / det;
        double x0 = i00 * gu + i01 * guk, x1 = i01 * gu + i11 * guk;
        if (draw) {
          // x + chol(A^-1) z, z ~ N(0, I)
          const u32x4 r = philox4x32_10(u32x4{0u, a.sweep, 0u, kStreamSsd}, key0, key1);
          double z0, z1;
          box_muller(r.x, r.y, z0, z1);
          const double l00 = sqrt(i00), l10 = i01 / l00, l11 = sqrt(i11 - l10 * l10);
          x0 += l00 * z0;
          x1 += l10 * z0 + l11 * z1;
        }
        sh = x0;
        sc = x1;
      }
      // sum h (u - m0 - m1 k)^2 from the moments (m0 = 0, m1 = 1 without
      // shift and scale)
      const double m0 = a.shift_scale ? sh : 0.0, m1 = a.shift_scale ? sc : 1.0;
      const double sq = guu - 2.0 * m0 * gu - 2.0 * m1 * guk + m0 * m0 * g +
                        2.0 * m0 * m1 * gk + m1 * m1 * gkk;
      const double num = sq + a.a0 * a.b0, den = g + a.a0;
      if (!draw) {
        s2 = num / den;
      } else {
        // the reference's draw (:278, :300): 1 / Gamma(shape den / 2, scale 2 den / num)
        const double gd =
            gamma_mt(den * 0.5, u32x4{1u, a.sweep, 0u, kStreamSsd}, key0, key1) *
            (2.0 * den / num);
        s2 = 1.0 / gd;
      }
    }
    a.par_out[0] = s2;
    a.par_out[1] = sh;
    a.par_out[2] = sc;
    a.par_out[3] = flag;
    // -log InvGamma(sigma^2; a0, a0 b0) (:311-323)
    const double pa = a.a0, pb = a.a0 * a.b0;
    a.stats[0] = mll;
    a.stats[1] = -(pa * log(pb) - lgamma(pa) + (-pa - 1.0) * log(s2) - pb / s2);
  }
}

// ---------------------------------------------------------------------------
// Linear transformations
// ---------------------------------------------------------------------------
constexpr int kLinThreads = 256;
constexpr int kLinMoments = 12;  // sum y (N) and sum x_i y_n (N x N), N <= 3

// Per-block FP64 partial sums of y_n and x_i y_n (x, y as N planes of nv).
__global__ __launch_bounds__(kLinThreads) void k_lin_moments(const float *__restrict__ y,
                                                             const float *__restrict__ x,
                                                             long long nv, int N,
                                                             double *__restrict__ part) {
  double m[kLinMoments] = {0};
  for (long long v = (long long)blockIdx.x * kLinThreads + threadIdx.x; v < nv;
       v += (long long)gridDim.x * kLinThreads) {
    double xv[3] = {0, 0, 0}, yv[3] = {0, 0, 0};
    for (int n = 0; n < N; ++n) {
      xv[n] = x[n * nv + v];
      yv[n] = y[n * nv + v];
    }
    for (int n = 0; n < N; ++n) {
      m[n] += yv[n];
      for (int i = 0; i < N; ++i) m[3 + i * 3 + n] += xv[i] * yv[n];
    }
  }
  __shared__ double red[kLinMoments][kLinThreads];
  for (int j = 0; j < kLinMoments; ++j) red[j][threadIdx.x] = m[j];
  __syncthreads();
  for (int w = kLinThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int j = 0; j < kLinMoments; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kLinMoments) part[blockIdx.x * kLinMoments + threadIdx.x] =
      red[threadIdx.x][0];
}

__device__ double det_small(int n, const double *M) {
  if (n == 2) return M[0] * M[3] - M[1] * M[2];
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
         M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// H = U diag(s) V^T (n = 2 or 3, row-major), singular values decreasing: V
// from the cyclic Jacobi eigen-decomposition of H^T H, U = H V / s, and a
// zero singular value's left vector completed to an orthonormal basis.
__device__ void svd_small_dev(int n, const double *H, double *U, double *s, double *V) {
  double B[9] = {0}, W[9] = {0};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += H[k * n + i] * H[k * n + j];
      B[i * n + j] = acc;
      W[i * n + j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 60; ++sweep) {
    // stop once the off-diagonal part is below 1e-15 of the diagonal
    // (relative: the moments scale with the voxel count)
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < n; ++p) {
      dia += B[p * n + p] * B[p * n + p];
      for (int q = p + 1; q < n; ++q) off += B[p * n + q] * B[p * n + q];
    }
    if (off <= 1e-30 * dia || off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = B[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        const double th = (B[q * n + q] - B[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; ++k) {
          const double bkp = B[k * n + p], bkq = B[k * n + q];
          B[k * n + p] = c * bkp - sn * bkq;
          B[k * n + q] = sn * bkp + c * bkq;
        }
        for (int k = 0; k < n; ++k) {
          const double bpk = B[p * n + k], bqk = B[q * n + k];
          B[p * n + k] = c * bpk - sn * bqk;
          B[q * n + k] = sn * bpk + c * bqk;
        }
        for (int k = 0; k < n; ++k) {
          const double wkp = W[k * n + p], wkq = W[k * n + q];
          W[k * n + p] = c * wkp - sn * wkq;
          W[k * n + q] = sn * wkp + c * wkq;
        }
      }
  }
  int idx[3] = {0, 1, 2};
  for (int a = 0; a < n; ++a)  // order by decreasing eigenvalue
    for (int b = a + 1; b < n; ++b)
      if (B[idx[b] * n + idx[b]] > B[idx[a] * n + idx[a]]) {
        const int t = idx[a];
        idx[a] = idx[b];
        idx[b] = t;
      }
  for (int j = 0; j < n; ++j) {
    s[j] = sqrt(fmax(B[idx[j] * n + idx[j]], 0.0));
    for (int k = 0; k < n; ++k) V[k * n + j] = W[k * n + idx[j]];
  }
  for (int j = 0; j < n; ++j) {
    double col[3] = {0, 0, 0}, nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      for (int k = 0; k < n; ++k) col[i] += H[i * n + k] * V[k * n + j];
      nrm += col[i] * col[i];
    }
    nrm = sqrt(nrm);
    if (nrm > 1e-12 * (s[0] > 0 ? s[0] : 1.0)) {
      for (int i = 0; i < n; ++i) U[i * n + j] = col[i] / nrm;
    } else if (n == 3 && j == 2) {
      U[2] = U[3] * U[7] - U[6] * U[4];
      U[5] = U[6] * U[1] - U[0] * U[7];
      U[8] = U[0] * U[4] - U[3] * U[1];
    } else if (n == 2 && j == 1) {
      U[1] = -U[2];
      U[3] = U[0];
    } else {
      for (int i = 0; i < n; ++i) U[i * n + j] = i == j ? 1.0 : 0.0;
    }
  }
}

// Sums the block partials (each thread a fixed strided subset, then a
// fixed tree: the same order every call), fits the transform, writes the
// parameters, the zero prior cost and (optionally) the host statistics.
__global__ __launch_bounds__(kLinThreads) void k_lin_solve(const double *__restrict__ part,
                                                           int nblk, int kind, int N,
                                                           long long nv,
                                                           const double *__restrict__ xstat,
                                                           double *par, double *prior_out,
                                                           const double *stats_in,
                                                           double *host_stats) {
  __shared__ double red[kLinMoments][kLinThreads];
  {
    double acc[kLinMoments] = {0};
    for (int b = threadIdx.x; b < nblk; b += kLinThreads)
      for (int j = 0; j < kLinMoments; ++j) acc[j] += part[b * kLinMoments + j];
    for (int j = 0; j < kLinMoments; ++j) red[j][threadIdx.x] = acc[j];
  }
  __syncthreads();
  for (int w = kLinThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int j = 0; j < kLinMoments; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double S[kLinMoments];
  for (int j = 0; j < kLinMoments; ++j) S[j] = red[j][0];
  const int n1 = N + 1;
  if (kind == kLinAffine) {
    // T = (X^T X)^-1 X^T Y with X = [1, x]: rows 0 (sum y) and 1 + i (sum x_i y)
    double XY[12];
    for (int nn = 0; nn < N; ++nn) {
      XY[0 * N + nn] = S[nn];
      for (int i = 0; i < N; ++i) XY[(i + 1) * N + nn] = S[3 + i * 3 + nn];
    }
    double T[12];
    for (int i = 0; i < n1; ++i)
      for (int nn = 0; nn < N; ++nn) {
        double acc = 0.0;
        for (int j = 0; j < n1; ++j) acc += xstat[i * n1 + j] * XY[j * N + nn];
        T[i * N + nn] = acc;
      }
    for (int i = 0; i < n1 * n1; ++i) par[i] = 0.0;
    for (int nn = 0; nn < N; ++nn) {
      for (int n2 = 0; n2 < N; ++n2) par[nn * n1 + n2] = T[(n2 + 1) * N + nn];
      par[n1 * n1 + nn] = T[0 * N + nn];
    }
    par[n1 * n1 - 1] = 1.0;
  } else {
    // H(n, n2) = sum x~[n2] y~[n] = sum x[n2] y[n] - xbar[n2] sum y[n]
    double Hm[9], U[9], sv[3], V[9], VUt[9], R[9], ybar[3];
    for (int nn = 0; nn < N; ++nn) {
      ybar[nn] = S[nn] / (double)nv;
      for (int n2 = 0; n2 < N; ++n2) Hm[nn * N + n2] = S[3 + n2 * 3 + nn] - xstat[n2] * S[nn];
    }
    svd_small_dev(N, Hm, U, sv, V);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        double acc = 0.0;
        for (int k = 0; k < N; ++k) acc += V[i * N + k] * U[j * N + k];
        VUt[i * N + j] = acc;
      }
    const double dsign = det_small(N, VUt);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        double acc = 0.0;
        for (int k = 0; k < N; ++k) acc += V[i * N + k] * (k == N - 1 ? dsign : 1.0) * U[j * N + k];
        R[i * N + j] = acc;
      }
    for (int i = 0; i < N * N; ++i) par[i] = R[i];
    for (int nn = 0; nn < N; ++nn) {
      double acc = ybar[nn];
      for (int n2 = 0; n2 < N; ++n2) acc -= R[nn * N + n2] * xstat[n2];
      par[N * N + nn] = acc;
    }
  }
  if (prior_out) prior_out[0] = 0.0;  // constrain_transformation returns 0
  if (host_stats) {
    host_stats[0] = stats_in[0];
    host_stats[1] = stats_in[1];
    host_stats[2] = 0.0;
    __threadfence_system();
  }
}

// out = M x + t per voxel (N planes of nv floats)
__global__ __launch_bounds__(256) void k_lin_apply(const double *__restrict__ par, int kind,
                                                   int N, const float *__restrict__ x,
                                                   float *__restrict__ out, long long nv) {
  const int stride = kind == kLinAffine ? N + 1 : N;
  const double *t = par + (kind == kLinAffine ? (N + 1) * (N + 1) : N * N);
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (long long)gridDim.x * blockDim.x) {
    double xv[3] = {0, 0, 0};
    for (int n = 0; n < N; ++n) xv[n] = x[n * nv + v];
    for (int n = 0; n < N; ++n) {
      double acc = t[n];
      for (int n2 = 0; n2 < N; ++n2) acc += par[n * stride + n2] * xv[n2];
      out[n * nv + v] = (float)acc;
    }
  }
}

// Inverse of an n x n matrix (n <= 4) by Gauss-Jordan elimination with
// partial pivoting (host; the affine normal matrix, once per model).
void invert_host(int n, const double *M, double *Minv) {
  double a[4][8];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) a[i][j] = j < n ? M[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (p != c)
      for (int j = 0; j < 2 * n; ++j) std::swap(a[p][j], a[c][j]);
    BINDER_REQUIRE(a[c][c] != 0.0, "the voxel positions do not span the space (X^T X singular)");
    const double inv = 1.0 / a[c][c];
    for (int j = 0; j < 2 * n; ++j) a[c][j] *= inv;
    for (int r = 0; r < n; ++r)
      if (r != c) {
        const double f = a[r][c];
        for (int j = 0; j < 2 * n; ++j) a[r][j] -= f * a[c][j];
      }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Minv[i * n + j] = a[i][j + n];
}

int lin_blocks(long long nv) {
  long long b = (nv + kLinThreads * 8 - 1) / (kLinThreads * 8);
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (int)b;
}

}  // namespace

void launch_ssd_theta(const SsdThetaArgs &a, hipStream_t s) {
  BINDER_REQUIRE(a.bins >= 1 && a.bins <= kSsdMaxBins, "SSD bins must be in [1, 64]");
  hipLaunchKernelGGL(k_ssd_theta, dim3(1), dim3(kSsdThreads), 0, s, a);
  HIP_LAUNCH_OK();
}

LinearModel::LinearModel(int kind, int N, long long nv, const float *x, const double *M0,
                         const double *t0, hipStream_t stream)
    : kind_(kind), N_(N), nv_(nv), stream_(stream), blocks_(lin_blocks(nv)) {
  BINDER_REQUIRE(kind == kLinAffine || kind == kLinRigid, "unknown linear transformation");
  BINDER_REQUIRE(N == 2 || N == 3, "linear transformations need 2 or 3 dimensions");
  const int nm = kind == kLinAffine ? (N + 1) * (N + 1) : N * N;
  std::vector<double> p0((size_t)nm + N);
  for (int i = 0; i < nm; ++i) p0[i] = M0[i];
  for (int n = 0; n < N; ++n) p0[nm + n] = t0[n];
  par_.upload(p0.data(), p0.size(), stream_);
  part_.alloc((size_t)blocks_ * kLinMoments);
  // moments of the positions themselves: sum x and sum x x^T
  hipLaunchKernelGGL(k_lin_moments, dim3(blocks_), dim3(kLinThreads), 0, stream_, x, x, nv,
                     N, part_.get());
  HIP_LAUNCH_OK();
  std::vector<double> part((size_t)blocks_ * kLinMoments), S(kLinMoments, 0.0);
  part_.download(part.data(), part.size(), stream_);
  HIP_OK(hipStreamSynchronize(stream_));
  for (int b = 0; b < blocks_; ++b)
    for (int j = 0; j < kLinMoments; ++j) S[j] += part[(size_t)b * kLinMoments + j];
  if (kind == kLinAffine) {
    // X^T X with X = [1, x] (AffineTransformation.cpp:32-49), inverted once
    const int n1 = N + 1;
    std::vector<double> M((size_t)n1 * n1), Minv((size_t)n1 * n1);
    M[0] = (double)nv;
    for (int i = 0; i < N; ++i) {
      M[0 * n1 + i + 1] = M[(i + 1) * n1 + 0] = S[i];
      for (int j = 0; j < N; ++j) M[(i + 1) * n1 + j + 1] = S[3 + i * 3 + j];
    }
    invert_host(n1, M.data(), Minv.data());
    xstat_.upload(Minv.data(), Minv.size(), stream_);
  } else {
    // x means (RigidTransformation.cpp:22-37)
    std::vector<double> xm(N);
    for (int n = 0; n < N; ++n) xm[n] = S[n] / (double)nv;
    xstat_.upload(xm.data(), xm.size(), stream_);
  }
  HIP_OK(hipStreamSynchronize(stream_));
}

void LinearModel::run(const float *dream, const float *x, float *out, const SmoothOptions &opt) {
  hipStream_t st = opt.exec_stream ? opt.exec_stream : stream_;
  BINDER_REQUIRE(!opt.host_stats || opt.stats_in, "host statistics need the device statistics");
  hipLaunchKernelGGL(k_lin_moments, dim3(blocks_), dim3(kLinThreads), 0, st, dream, x, nv_, N_,
                     part_.get());
  HIP_LAUNCH_OK();
  hipLaunchKernelGGL(k_lin_solve, dim3(1), dim3(kLinThreads), 0, st, part_.get(), blocks_, kind_, N_,
                     nv_, xstat_.get(), par_.get(), opt.want_prior ? opt.prior_out : nullptr,
                     opt.stats_in, opt.host_stats);
  HIP_LAUNCH_OK();
  const long long b = std::min<long long>((nv_ + 255) / 256, 4096);
  hipLaunchKernelGGL(k_lin_apply, dim3((unsigned)b), dim3(256), 0, st, par_.get(), kind_, N_, x,
                     out, nv_);
  HIP_LAUNCH_OK();
}

int LinearModel::get(double *out) {
  const int n = (kind_ == kLinAffine ? (N_ + 1) * (N_ + 1) : N_ * N_) + N_;
  par_.download(out, (size_t)n, stream_);
  HIP_OK(hipStreamSynchronize(stream_));
  return n;
}

}  // namespace binder
