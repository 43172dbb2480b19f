// ORACLE -- test infrastructure, not product code.
//
// FP64 CPU restatement of BINDER's registration-and-sampling path, written
// from the reference's behaviour (file:line cited per function).  It is the
// checker for the HIP backend in tests/, the reference for
// __graft_entry__.smoke(), and the CPU baseline timed by bench.py.  The
// product library (binder_amd) never loads it.
//
// Each function is pinned by tests/test_oracle.py: golden vectors and
// fresh-input comparisons against the reference's own FFTW/Eigen-free sources
// compiled into oracle/_ref (Jacobians, spline, MI histogram/theta update),
// closed forms (DCT, smoothing and cost, up_sample, topology integration),
// brute force on tiny grids (E-step), and distribution tests (draws).
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <vector>

#include "fft.h"
#include "oracle.h"

namespace orc {

// ---------------------------------------------------------------------------
// Cubic B-spline weights at offsets {-1,0,1,2} for 0 <= d < 1
// (BINDER/src/cpu/Spline.cpp:44-73; order 2 :75-93, order 1 :94-108).
// ---------------------------------------------------------------------------
static inline void spline_weights(int order, double d, double w[4]) {
  if (order == 3) {
    const double d2 = d * d;
    const double d3 = d2 * d;
    const double omd = 1.0 - d;
    const double omd2 = omd * omd;
    const double omd3 = omd2 * omd;
    w[0] = omd3 / 6.0;
    w[1] = 2.0 / 3.0 - d2 + d3 * 0.5;
    w[2] = 2.0 / 3.0 - omd2 + omd3 * 0.5;
    w[3] = d3 / 6.0;
  } else if (order == 2) {
    const double d2 = d * d;
    const double a = -2.0 * d + 1.0;
    const double b = -2.0 * d - 1.0;
    w[0] = 0.125 * (a * a);
    w[1] = 0.75 - d2;
    w[2] = 0.125 * (b * b);
    w[3] = 0.0;
  } else {
    w[0] = 1 - d;
    w[1] = d;
    w[2] = 0.0;
    w[3] = 0.0;
  }
}

// log of the spline weights with the reference's sentinels for zero weights
// (BINDER/src/cpu/Spline.cpp:111-146)
static inline void log_spline_weights(int order, double d, double o[4]) {
  o[3] = -1e100;
  if (order == 3) {
    o[3] = d < 1e-300 ? -1e100 : 3 * std::log(d) - std::log(6.0);
    o[0] = 3 * std::log(1.0 - d) - std::log(6.0);
    o[1] = std::log(2.0 / 3.0 - d * d + std::pow(d, 3) / 2.0);
    o[2] = std::log(2.0 / 3.0 - std::pow(1.0 - d, 2) + std::pow(1.0 - d, 3) / 2.0);
  } else if (order == 2) {
    if (d > 0.5) {
      o[0] = o[1] = o[2] = -1e100;
    } else {
      o[0] = std::log(0.125 * std::pow(-2.0 * d + 1.0, 2));
      o[1] = std::log(0.75 - d * d);
      o[2] = std::log(0.125 * std::pow(-2.0 * d - 1.0, 2));
    }
  } else {
    o[0] = d >= 1 ? -1e100 : std::log(1 - d);
    o[1] = d < 1e-300 ? -1e100 : std::log(d);
    o[2] = -1e100;
  }
}

// Inverse of an n x n matrix (n <= 4) by Gauss-Jordan elimination with
// partial pivoting (row-major); the reference uses Eigen's inverse().
static void invert(int n, const double *M, double *Minv) {
  double a[4][8];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) a[i][j] = j < n ? M[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (p != c)
      for (int j = 0; j < 2 * n; ++j) std::swap(a[p][j], a[c][j]);
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

static double det_n(int n, const double *M) {
  if (n == 2) return M[0] * M[3] - M[1] * M[2];
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
         M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// SVD H = U diag(s) V^T of an n x n matrix (n = 2 or 3, row-major), singular
// values in decreasing order: V from the cyclic Jacobi eigen-decomposition of
// H^T H, U = H V / s, and a zero singular value's left vector completed to
// an orthonormal basis (Eigen's JacobiSVD in the reference).
static void svd_small(int n, const double *H, double *U, double *s, double *V) {
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
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < n; ++p) {
      dia += B[p * n + p] * B[p * n + p];
      for (int q = p + 1; q < n; ++q) off += B[p * n + q] * B[p * n + q];
    }
    if (off <= 1e-30 * dia || off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = B[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (B[q * n + q] - B[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; ++k) {  // B <- J^T B J
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
  std::sort(idx, idx + n, [&](int a, int b) { return B[a * n + a] > B[b * n + b]; });
  for (int j = 0; j < n; ++j) {
    s[j] = std::sqrt(std::max(B[idx[j] * n + idx[j]], 0.0));
    for (int k = 0; k < n; ++k) V[k * n + j] = W[k * n + idx[j]];
  }
  for (int j = 0; j < n; ++j) {
    double col[3] = {0, 0, 0}, nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      for (int k = 0; k < n; ++k) col[i] += H[i * n + k] * V[k * n + j];
      nrm += col[i] * col[i];
    }
    nrm = std::sqrt(nrm);
    if (nrm > 1e-12 * (s[0] > 0 ? s[0] : 1.0)) {
      for (int i = 0; i < n; ++i) U[i * n + j] = col[i] / nrm;
    } else if (n == 3 && j == 2) {  // u3 = u1 x u2
      U[0 * 3 + 2] = U[1 * 3 + 0] * U[2 * 3 + 1] - U[2 * 3 + 0] * U[1 * 3 + 1];
      U[1 * 3 + 2] = U[2 * 3 + 0] * U[0 * 3 + 1] - U[0 * 3 + 0] * U[2 * 3 + 1];
      U[2 * 3 + 2] = U[0 * 3 + 0] * U[1 * 3 + 1] - U[1 * 3 + 0] * U[0 * 3 + 1];
    } else if (n == 2 && j == 1) {
      U[0 * 2 + 1] = -U[1 * 2 + 0];
      U[1 * 2 + 1] = U[0 * 2 + 0];
    } else {
      for (int i = 0; i < n; ++i) U[i * n + j] = i == j ? 1.0 : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Per-axis orthonormal DCT of a [nx,ny,nz] C-order volume, OpenMP over lines.
// ---------------------------------------------------------------------------
static void dct3(double *data, int nx, int ny, int nz, bool inverse,
                 int threads) {
  const int dims[3] = {nx, ny, nz};
  const long long strides[3] = {(long long)ny * nz, nz, 1};
  for (int axis = 2; axis >= 0; --axis) {
    const int n = dims[axis];
    if (n == 1) continue;
    const long long st = strides[axis];
    const long long lines = (long long)nx * ny * nz / n;
#pragma omp parallel num_threads(threads)
    {
      DCT1D t(n);
      std::vector<double> buf(n);
#pragma omp for schedule(static)
      for (long long l = 0; l < lines; ++l) {
        // enumerate the lines orthogonal to `axis`
        long long base;
        if (axis == 2) {
          base = l * nz;
        } else if (axis == 1) {
          const long long x = l / nz, z = l % nz;
          base = x * ny * nz + z;
        } else {
          base = l;  // (y,z) flattened
        }
        for (int j = 0; j < n; ++j) buf[j] = data[base + j * st];
        if (inverse)
          t.inverse(buf.data());
        else
          t.forward(buf.data());
        for (int j = 0; j < n; ++j) data[base + j * st] = buf[j];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// DCT-domain smoothing and sampling of a deformation
// (BINDER/src/cpu/NonLinearTransformation.cpp:15-223, :344-364, :433-451).
// ---------------------------------------------------------------------------
struct Smoother {
  int nx, ny, nz, N, threads;
  long long nv;
  double sigma, offset;
  double gamma[3], rho[3], eta[3];
  double alpha0 = 1.0, beta0 = 10000.0, nu;
  std::vector<double> D;
  std::mt19937 gen;  // default seeded, replaced by set_generator()
  bool sample = false;
  std::vector<double> work, coeff_saved, tmp_sigma, noise;
  bool recompute_sigma = true;

  Smoother(double g[3], double r[3], double e[3], const double *Din, int nx_,
           int ny_, int nz_, int threads_, double sigma_, double offset_)
      : nx(nx_), ny(ny_), nz(nz_), N(nz_ == 1 ? 2 : 3), threads(threads_),
        nv((long long)nx_ * ny_ * nz_), sigma(sigma_), offset(offset_) {
    for (int i = 0; i < 3; ++i) {
      gamma[i] = g[i];
      rho[i] = r[i];
      eta[i] = e[i];
    }
    D.assign(Din, Din + nv);
    // NonLinearTransformation.cpp:43
    nu = gamma[0] / (rho[0] * rho[1] * rho[2] * eta[0] * eta[0] * (double)nv);
    work.resize(nv);
  }

  // compute_fourier_tmp_sigma_DCT (:344-364)
  void compute_sigma() {
    tmp_sigma.resize((size_t)N * nv);
    for (int n = 0; n < N; ++n)
      for (long long i = 0; i < nv; ++i)
        tmp_sigma[n * nv + i] = std::sqrt(
            sigma * sigma / (1.0 + gamma[n] * std::pow(sigma, 2) * D[i]));
  }

  // precompute_sample_from_gaussian_distribution (:433-451)
  void precompute_noise() {
    if (recompute_sigma) {
      compute_sigma();
      recompute_sigma = false;
    }
    std::normal_distribution<double> dist(0.0, 1.0);
    noise.resize((size_t)N * nv);
    for (long long it = 0; it < (long long)N * nv; ++it)
      noise[it] = dist(gen) * tmp_sigma[it];
  }

  // smooth(n) (:190-223): DCT, filter (+ noise), cost, IDCT.  `buf` holds the
  // displacement on entry and the smoothed displacement on exit; `coef`
  // receives the (noisy) coefficients when sample && sample_gamma.
  double smooth(int n, double *buf, bool sample_gamma, double *coef) {
    dct3(buf, nx, ny, nz, false, threads);
    double cost = 0.0;
    for (long long i = 0; i < nv; ++i) {
      double p = buf[i] * (1.0 / (1.0 + gamma[n] * sigma * sigma * D[i]));
      if (sample) p += noise[n * nv + i];
      if (sample && sample_gamma) coef[i] = p;
      cost += 0.5 * gamma[n] * std::pow(std::sqrt(D[i]) * std::abs(p), 2);
      buf[i] = p;
    }
    dct3(buf, nx, ny, nz, true, threads);
    return cost;
  }

  // constrain_transformation (:58-76) / sample_transformation (:79-126):
  // dream, voxel_pos, final_locations are [nv, N] interleaved.
  double constrain(const double *dream, const double *xpos, double *final_loc,
                   bool do_sample, bool sample_gamma) {
    double prior = 0.0, total_raw = 0.0;
    if (do_sample) {
      sample = true;
      recompute_sigma = true;
      precompute_noise();
      coeff_saved.resize(nv);
    }
    for (int n = 0; n < N; ++n) {
      for (long long i = 0; i < nv; ++i)
        work[i] = dream[i * N + n] - xpos[i * N + n];  // fill_in_data (:128-137)
      prior += smooth(n, work.data(), do_sample && sample_gamma,
                      coeff_saved.empty() ? nullptr : coeff_saved.data());
      if (do_sample && sample_gamma) {  // compute_raw_regularization (:151-171)
        const double rp = rho[0] * rho[1] * rho[2];
        double raw = 0.0;
        for (long long i = 0; i < nv; ++i)
          raw += eta[n] * eta[n] * rp * D[i] * coeff_saved[i] * coeff_saved[i];
        total_raw += raw;
      }
      for (long long i = 0; i < nv; ++i)  // update_final_locations (:173-188)
        final_loc[i * N + n] = work[i] + xpos[i * N + n];
    }
    if (do_sample && sample_gamma) {  // :116-124, updateGammasFromNu :140-148
      const double a_post = alpha0 + 0.5 * (double)(N * nv);
      const double b_post = beta0 + 0.5 * (double)nv * total_raw;
      std::gamma_distribution<double> gd(a_post, 1.0 / b_post);
      nu = gd(gen);
      const double rp = rho[0] * rho[1] * rho[2];
      for (int k = 0; k < 3; ++k) gamma[k] = nu * rp * eta[k] * eta[k] * (double)nv;
      recompute_sigma = true;
    }
    return prior;
  }
};

// ---------------------------------------------------------------------------
// Registration: E-step, theta update, EM loop and Gibbs sampler
// (BINDER/src/cpu/Registration.cpp, BINDER/src/cpu/MILikelihood.cpp).
// ---------------------------------------------------------------------------
struct Registration {
  int nxv, nyv, nzv, nxn, nyn, nzn, N, threads, order;
  long long nv, nn;
  double nnz;
  bool is2D;
  std::vector<double> final_loc, xpos, dream;
  // likelihood (MILikelihood.cpp:8-30)
  int K = 0, L = 0;
  double alpha = 0;
  std::vector<double> theta, H;
  std::vector<int> bn, bv;
  std::mt19937 lk_gen;
  std::unique_ptr<Smoother> sm;
  std::vector<double> trace;  // rows of (mll, prior_theta, prior_d, mlp)
  // SSD likelihood (SSDLikelihood.h:13-88): lik 1 replaces the MI table
  int lik = 0;
  int ssd_shift_scale = 0;
  double sigma2 = 1.0, shift = 0.0, scale = 1.0, a0 = 0.0, b0 = 0.0, log_constant = 0.0;
  int ssd_sample = 0;  // set by sample_parameters and never cleared (SSDLikelihood.cpp:285)
  // affine / rigid transformation (AffineTransformation.cpp,
  // RigidTransformation.cpp): tmodel 1 or 2 replaces the DCT smoother
  int tmodel = 0;
  std::vector<double> A_aff, R_rig, t_lin, x_means;

  Registration(const double *fl, int threads_, const double *vp, int nxv_,
               int nyv_, int nzv_, int nxn_, int nyn_, int nzn_, int order_,
               double nnz_)
      : nxv(nxv_), nyv(nyv_), nzv(nzv_), nxn(nxn_), nyn(nyn_), nzn(nzn_),
        N(nzv_ == 1 ? 2 : 3), threads(threads_ < 1 ? 1 : threads_),
        order(order_), nv((long long)nxv_ * nyv_ * nzv_),
        nn((long long)nxn_ * nyn_ * nzn_), nnz(nnz_), is2D(nzv_ == 1) {
    final_loc.assign(fl, fl + nv * N);
    xpos.assign(vp, vp + nv * N);
    dream.assign(nv * N, 0.0);
  }

  double offset() const { return order == 2 ? 0.5 : 0.0; }
  double sigma_spline() const {
    if (order == 3) return std::sqrt(9.0 / (8.0 * M_PI));
    if (order == 2) return std::sqrt(8.0 / (9.0 * M_PI));
    return 0.4;
  }

  // MILikelihood::compute_likelihood / accumulate_info bin lookup
  // (MILikelihood.cpp:42-55, :89-98): out-of-range nodes use moving bin 0.
  inline int node_bin(int x, int y, int z) const {
    if (x < 0 || y < 0 || z < 0 || x >= nxn || y >= nyn || z >= nzn) return 0;
    return bn[((long long)x * nyn + y) * nzn + z];
  }

  // estimate_posterior_and_dream_location (Registration.cpp:219-330) for one
  // voxel; `sample` selects the categorical draw (:278-307).
  // Returns -log(normalizer).
  double voxel(long long vi, bool sample, std::mt19937 *rng, double *Hacc) {
    const int nb = order + 1;
    const int ZZ = is2D ? 1 : nb;
    static const int off3[4] = {-1, 0, 1, 2}, off2[4] = {-1, 0, 1, 0},
                     off1[4] = {0, 1, 0, 0};
    const int *off = order == 3 ? off3 : (order == 2 ? off2 : off1);
    double wx[4], wy[4], wz[4] = {1.0, 1.0, 1.0, 1.0};
    int ix, iy, iz = 0;
    {
      const double o = offset();
      double t = final_loc[vi * N + 0];
      ix = (int)std::floor(t + o);
      spline_weights(order, t - std::floor(t + o), wx);
      t = final_loc[vi * N + 1];
      iy = (int)std::floor(t + o);
      spline_weights(order, t - std::floor(t + o), wy);
      if (!is2D) {
        t = final_loc[vi * N + 2];
        iz = (int)std::floor(t + o);
        spline_weights(order, t - std::floor(t + o), wz);
      }
    }
    const int u = bv[vi];
    double post[64];
    int bins[64];
    double Z = 0.0;
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b) {
        const double wxy = wx[a] * wy[b];
        for (int c = 0; c < ZZ; ++c) {
          const int j = (a * nb + b) * ZZ + c;
          const int tz = is2D ? 0 : iz + off[c];
          const int k = node_bin(ix + off[a], iy + off[b], tz);
          bins[j] = k;
          post[j] = wxy * wz[c] * theta[(long long)k * L + u];
          Z += post[j];
        }
      }
    if (sample) {
      std::uniform_real_distribution<double> dis(0.0, 1.0);
      const double r = dis(*rng);
      double cum = 0.0;
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
          for (int c = 0; c < ZZ; ++c) {
            const int j = (a * nb + b) * ZZ + c;
            post[j] /= Z;
            cum += post[j];
            if (r <= cum) {
              Hacc[(long long)bins[j] * L + u] += 1.0;
              dream[vi * N + 0] = (double)(ix + off[a]);
              dream[vi * N + 1] = (double)(iy + off[b]);
              if (!is2D) dream[vi * N + 2] = (double)(iz + off[c]);
              return -std::log(Z);
            }
          }
    } else {
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
          for (int c = 0; c < ZZ; ++c) {
            const int j = (a * nb + b) * ZZ + c;
            post[j] /= Z;
            Hacc[(long long)bins[j] * L + u] += post[j];
            dream[vi * N + 0] += post[j] * (double)(ix + off[a]);
            dream[vi * N + 1] += post[j] * (double)(iy + off[b]);
            if (!is2D) dream[vi * N + 2] += post[j] * (double)(iz + off[c]);
          }
    }
    return -std::log(Z);
  }

  // compute_log_likelihood (SSDLikelihood.cpp:60-86) for fixed bin u and
  // the moving bin k of the node (0 outside the grid: the value the
  // reference uses for an outside node, voxel - 0 or voxel - shift)
  double ssd_loglik(int u, int k) const {
    const double d = ssd_shift_scale == 0 ? (double)u - (double)k
                                           : (double)u - shift - scale * (double)k;
    return (d * d) / (-2.0 * sigma2) + log_constant;
  }

  // estimate_posterior_and_dream_location_log (Registration.cpp:332-440),
  // the E-step of the SSD likelihood (log_space, Registration.cpp:478)
  double voxel_log(long long vi, bool sample, std::mt19937 *rng, double *Hacc) {
    const int nb = order + 1;
    const int ZZ = is2D ? 1 : nb;
    static const int off3[4] = {-1, 0, 1, 2}, off2[4] = {-1, 0, 1, 0},
                     off1[4] = {0, 1, 0, 0};
    const int *off = order == 3 ? off3 : (order == 2 ? off2 : off1);
    double lx[4], ly[4], lz[4] = {0.0, 0.0, 0.0, 0.0};
    int ix, iy, iz = 0;
    {
      const double o = offset();
      double t = final_loc[vi * N + 0];
      ix = (int)std::floor(t + o);
      log_spline_weights(order, t - std::floor(t + o), lx);
      t = final_loc[vi * N + 1];
      iy = (int)std::floor(t + o);
      log_spline_weights(order, t - std::floor(t + o), ly);
      if (!is2D) {
        t = final_loc[vi * N + 2];
        iz = (int)std::floor(t + o);
        log_spline_weights(order, t - std::floor(t + o), lz);
      }
    }
    const int u = bv[vi];
    double lp[64];
    int bins[64];
    double mx = -std::numeric_limits<double>::infinity();
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b)
        for (int c = 0; c < ZZ; ++c) {
          const int j = (a * nb + b) * ZZ + c;
          const int tz = is2D ? 0 : iz + off[c];
          const int k = node_bin(ix + off[a], iy + off[b], tz);
          bins[j] = k;
          lp[j] = lx[a] + ly[b] + lz[c] + ssd_loglik(u, k);
          mx = std::max(mx, lp[j]);
        }
    double nrm = 0.0;
    for (int j = 0; j < nb * nb * ZZ; ++j) nrm += std::exp(lp[j] - mx);
    const double lnz = std::log(nrm) + mx;
    if (sample) {
      std::uniform_real_distribution<double> dis(0.0, 1.0);
      const double r = dis(*rng);
      double cum = 0.0;
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
          for (int c = 0; c < ZZ; ++c) {
            const int j = (a * nb + b) * ZZ + c;
            cum += std::exp(lp[j] - lnz);
            if (r <= cum) {
              Hacc[(long long)bins[j] * L + u] += 1.0;
              dream[vi * N + 0] = (double)(ix + off[a]);
              dream[vi * N + 1] = (double)(iy + off[b]);
              if (!is2D) dream[vi * N + 2] = (double)(iz + off[c]);
              return -lnz;
            }
          }
    } else {
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
          for (int c = 0; c < ZZ; ++c) {
            const int j = (a * nb + b) * ZZ + c;
            const double p = std::exp(lp[j] - lnz);
            Hacc[(long long)bins[j] * L + u] += p;
            dream[vi * N + 0] += p * (double)(ix + off[a]);
            dream[vi * N + 1] += p * (double)(iy + off[b]);
            if (!is2D) dream[vi * N + 2] += p * (double)(iz + off[c]);
          }
    }
    return -lnz;
  }

  // The SSD statistics as sums over the soft joint histogram H[k][u] (the
  // posterior mass of moving bin k at voxels of fixed bin u): the per-node
  // sums of SSDLikelihood::accumulate_info (:117-146) grouped by bin.
  //   g = sum post, gk = sum post k, gkk = sum post k^2, gu = sum post u,
  //   guk = sum post u k, and sq(m0, m1) = sum post (u - m0 - m1 k)^2.
  void ssd_moments(double &g, double &gk, double &gkk, double &gu, double &guk) const {
    g = gk = gkk = gu = guk = 0.0;
    for (int k = 0; k < K; ++k)
      for (int u = 0; u < L; ++u) {
        const double h = H[(size_t)k * L + u];
        g += h;
        gk += h * k;
        gkk += h * k * (double)k;
        gu += h * u;
        guk += h * u * (double)k;
      }
  }
  double ssd_sq(double m0, double m1) const {
    double acc = 0.0;
    for (int k = 0; k < K; ++k)
      for (int u = 0; u < L; ++u) {
        const double d = (double)u - m0 - m1 * (double)k;
        acc += H[(size_t)k * L + u] * d * d;
      }
    return acc;
  }

  // update_parameters / update_shift_and_scale / update_sigma /
  // sample_parameters (SSDLikelihood.cpp:148-307)
  void ssd_update(bool draw) {
    if (draw) ssd_sample = 1;
    // with shift and scale, update_shift_and_scale / update_sigma draw once
    // the flag is set (:211, :273); without them update_parameters is the
    // MAP step whatever the flag (:148-167) and only sample_parameters draws
    const bool smp = ssd_shift_scale == 1 ? ssd_sample != 0 : draw;
    double g, gk, gkk, gu, guk;
    ssd_moments(g, gk, gkk, gu, guk);
    if (ssd_shift_scale == 1) {
      // A = [[g, gk], [gk, gkk]], b = [gu, guk] (outside nodes count as k = 0)
      const double M[4] = {g, gk, gk, gkk};
      double Minv[4];
      invert(2, M, Minv);
      double x0 = Minv[0] * gu + Minv[1] * guk, x1 = Minv[2] * gu + Minv[3] * guk;
      if (smp) {
        // (shift, scale) = x + chol(A^-1) z, z ~ N(0, I)
        std::normal_distribution<double> normal(0.0, 1.0);
        const double z0 = normal(lk_gen), z1 = normal(lk_gen);
        const double l00 = std::sqrt(Minv[0]);
        const double l10 = Minv[2] / l00;
        const double l11 = std::sqrt(Minv[3] - l10 * l10);
        x0 += l00 * z0;
        x1 += l10 * z0 + l11 * z1;
      }
      shift = x0;
      scale = x1;
    }
    const double num = (ssd_shift_scale == 1 ? ssd_sq(shift, scale) : ssd_sq(0.0, 1.0)) +
                       a0 * b0;
    const double den = g + a0;
    if (!smp) {
      sigma2 = num / den;
    } else {
      // the reference's draw: 1 / Gamma(shape den / 2, scale 2 den / num)
      std::gamma_distribution<double> gd(den * 0.5, 2.0 * den / num);
      sigma2 = 1.0 / gd(lk_gen);
    }
  }

  // compute_cost (SSDLikelihood.cpp:311-323): -log InvGamma(sigma^2; a0, a0 b0)
  double ssd_cost() const {
    const double a = a0, b = a0 * b0, x = sigma2;
    return -(a * std::log(b) - std::lgamma(a) + (-a - 1.0) * std::log(x) - b / x);
  }

  // AffineTransformation::constrain_transformation (:25-130) and
  // RigidTransformation::constrain_transformation (:57-154): the fitted
  // transform of the voxel positions, written to final_loc; prior cost 0.
  double lin_constrain() {
    const int n1 = N + 1;
    if (tmodel == 1) {
      double M[16] = {0}, XY[12] = {0};
      for (long long v = 0; v < nv; ++v) {
        double xr[4] = {1.0, 0, 0, 0};
        for (int n = 0; n < N; ++n) xr[n + 1] = xpos[v * N + n];
        for (int i = 0; i < n1; ++i) {
          for (int j = 0; j < n1; ++j) M[i * n1 + j] += xr[i] * xr[j];
          for (int n = 0; n < N; ++n) XY[i * N + n] += xr[i] * dream[v * N + n];
        }
      }
      double Minv[16];
      invert(n1, M, Minv);
      double T[12] = {0};  // (X^T X)^-1 X^T Y, rows 0..N
      for (int i = 0; i < n1; ++i)
        for (int n = 0; n < N; ++n)
          for (int j = 0; j < n1; ++j) T[i * N + n] += Minv[i * n1 + j] * XY[j * N + n];
      A_aff.assign((size_t)n1 * n1, 0.0);
      for (int n = 0; n < N; ++n) {
        t_lin[n] = T[0 * N + n];
        for (int n2 = 0; n2 < N; ++n2) A_aff[n * n1 + n2] = T[(n2 + 1) * N + n];
      }
      A_aff[(size_t)n1 * n1 - 1] = 1.0;
      for (long long v = 0; v < nv; ++v)
        for (int n = 0; n < N; ++n) {
          double acc = t_lin[n];
          for (int n2 = 0; n2 < N; ++n2) acc += A_aff[n * n1 + n2] * xpos[v * N + n2];
          final_loc[v * N + n] = acc;
        }
    } else {
      double ym[3] = {0, 0, 0};
      for (long long v = 0; v < nv; ++v)
        for (int n = 0; n < N; ++n) ym[n] += dream[v * N + n];
      for (int n = 0; n < N; ++n) ym[n] /= (double)nv;
      double Hm[9] = {0};  // H(n, n2) = sum x~[n2] y~[n]
      for (long long v = 0; v < nv; ++v)
        for (int n = 0; n < N; ++n) {
          const double yt = dream[v * N + n] - ym[n];
          for (int n2 = 0; n2 < N; ++n2) Hm[n * N + n2] += (xpos[v * N + n2] - x_means[n2]) * yt;
        }
      double U[9], sv[3], V[9], VUt[9];
      svd_small(N, Hm, U, sv, V);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          double acc = 0.0;
          for (int k = 0; k < N; ++k) acc += V[i * N + k] * U[j * N + k];
          VUt[i * N + j] = acc;
        }
      const double dsign = det_n(N, VUt);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          double acc = 0.0;
          for (int k = 0; k < N; ++k) acc += V[i * N + k] * (k == N - 1 ? dsign : 1.0) * U[j * N + k];
          R_rig[i * N + j] = acc;
        }
      for (int n = 0; n < N; ++n) {
        double acc = ym[n];
        for (int n2 = 0; n2 < N; ++n2) acc -= R_rig[n * N + n2] * x_means[n2];
        t_lin[n] = acc;
      }
      for (long long v = 0; v < nv; ++v)
        for (int n = 0; n < N; ++n) {
          double acc = t_lin[n];
          for (int n2 = 0; n2 < N; ++n2) acc += R_rig[n * N + n2] * xpos[v * N + n2];
          final_loc[v * N + n] = acc;
        }
    }
    return 0.0;
  }

  double constrain(bool sample, bool sample_gamma) {
    if (tmodel != 0) return lin_constrain();  // sample_transformation = constrain
    return sm->constrain(dream.data(), xpos.data(), final_loc.data(), sample, sample_gamma);
  }

  // The E-step voxel loop (Registration.cpp:144-181 and :551-571).
  // Histogram and mll are accumulated per x-slab and summed in slab order so
  // the oracle's result does not depend on the thread count.
  double estep(bool sample, std::vector<std::mt19937> *rngs) {
    std::fill(dream.begin(), dream.end(), 0.0);  // initialize_dream_locations
    // SSDLikelihood::initialize_parameters (:39-58)
    if (lik == 1) log_constant = -std::log(std::sqrt(2.0 * M_PI * sigma2));
    std::vector<double> slabH((size_t)nxv * K * L, 0.0);
    std::vector<double> slab_mll(nxv, 0.0);
#pragma omp parallel for schedule(static, 1) num_threads(threads)
    for (int x = 0; x < nxv; ++x) {
      std::mt19937 *rng = rngs ? &(*rngs)[omp_get_thread_num()] : nullptr;
      double *Hs = &slabH[(size_t)x * K * L];
      double m = 0.0;
      for (int y = 0; y < nyv; ++y)
        for (int z = 0; z < nzv; ++z) {
          const long long vi = ((long long)x * nyv + y) * nzv + z;
          m += lik == 1 ? voxel_log(vi, sample, rng, Hs) : voxel(vi, sample, rng, Hs);
        }
      slab_mll[x] = m;
    }
    H.assign((size_t)K * L, 0.0);
    double mll = 0.0;
    for (int x = 0; x < nxv; ++x) {
      mll += slab_mll[x];
      for (int i = 0; i < K * L; ++i) H[i] += slabH[(size_t)x * K * L + i];
    }
    return mll;
  }

  // update_parameters (MILikelihood.cpp:71-86)
  void theta_update() {
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int l = 0; l < L; ++l) s += (H[k * L + l] + alpha - 1.0);
      for (int l = 0; l < L; ++l) theta[k * L + l] = (H[k * L + l] + alpha - 1.0) / s;
    }
  }

  // compute_cost / dirichlet_logpdf (MILikelihood.cpp:33-39, :183-206)
  double theta_cost() const {
    double cost = 0.0;
    for (int k = 0; k < K; ++k) {
      double sa = 0.0;
      for (int l = 0; l < L; ++l) sa += alpha;
      double lp = std::lgamma(sa);
      for (int l = 0; l < L; ++l) lp -= std::lgamma(alpha);
      for (int l = 0; l < L; ++l) lp += (alpha - 1.0) * std::log(theta[k * L + l]);
      cost += -lp;
    }
    return cost;
  }

  // sample_parameters (MILikelihood.cpp:132-169): theta_k ~ Dir(H_k + alpha)
  void theta_sample() {
    std::vector<double> tmp((size_t)K * L);
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int l = 0; l < L; ++l) {
        std::gamma_distribution<double> dist(H[k * L + l] + alpha, 1.0);
        tmp[k * L + l] = dist(lk_gen);
        s += tmp[k * L + l];
      }
      for (int l = 0; l < L; ++l) tmp[k * L + l] /= s;
    }
    theta = tmp;
  }

  // Registration::EM (Registration.cpp:122-217)
  int em(int max_it, double th, bool update, bool debug) {
    trace.clear();
    double mlp_old = 1e100;
    int it = 0;
    bool converged = false;
    while (it < max_it && !converged) {
      const double mll = estep(false, nullptr);
      if (update) {
        if (lik == 1) ssd_update(false);
        else theta_update();
      }
      const double pt = lik == 1 ? ssd_cost() : theta_cost();
      const double pd = constrain(false, false);
      const double mlp = mll + pd + pt;
      trace.insert(trace.end(), {mll, pt, pd, mlp});
      if (debug)
        std::printf("oracle EM %d: mll %.8f prior_theta %.8f prior_d %.8f mlp %.8f\n",
                    it + 1, mll, pt, pd, mlp);
      if ((mlp_old - mlp) / nnz < th) converged = true;
      mlp_old = mlp;
      ++it;
    }
    return it;
  }

  // Registration::sampler (Registration.cpp:515-642)
  void sampler(int seed, int Nb, int Ns, int sample_gamma, int every, int sp,
               int sl, int st, double *mean_out, double *cov_out) {
    std::vector<std::mt19937> rngs(threads);
    for (int i = 0; i < threads; ++i) rngs[i].seed(seed + i);
    std::mt19937 rng;
    rng.seed(seed);
    lk_gen = rng;   // lkObj->set_generator(rng) copies the engine
    if (sm) sm->gen = rng;  // transformationObj->set_generator(rng)
    std::vector<double> avg((size_t)nv * N, 0.0), m2((size_t)nv * N * N, 0.0);
    for (int sweep = 0; sweep < Nb + Ns; ++sweep) {
      if (sp) estep(true, &rngs);
      if (sl) {
        if (lik == 1) ssd_update(true);
        else theta_sample();
      }
      if (st) {
        const bool sg = sample_gamma && (sweep % every == 0);
        constrain(true, sg);
      }
      if (sweep >= Nb) {
        const double inv_n = 1.0 / ((double)(sweep - Nb) + 1.0);
        for (long long v = 0; v < nv; ++v) {
          double d1[3], d2[3];
          for (int n = 0; n < N; ++n) {
            const double val = final_loc[v * N + n];
            const double mo = avg[v * N + n];
            d1[n] = val - mo;
            const double mn = mo + d1[n] * inv_n;
            avg[v * N + n] = mn;
            d2[n] = val - mn;
          }
          for (int r = 0; r < N; ++r)
            for (int c = 0; c < N; ++c) m2[(v * N + r) * N + c] += d1[r] * d2[c];
        }
      }
    }
    const double inv = 1.0 / (Ns - 1);
    for (size_t i = 0; i < avg.size(); ++i) mean_out[i] = avg[i];
    for (size_t i = 0; i < m2.size(); ++i) cov_out[i] = m2[i] * inv;
  }
};

}  // namespace orc

using namespace orc;

// ===========================================================================
// C API (ctypes) -- mirrors include/binder_c.h with FP64 arrays.
// ===========================================================================
extern "C" {

void orc_spline(int order, double d, double *out4) { spline_weights(order, d, out4); }

void orc_dct3(double *data, int nx, int ny, int nz, int inverse, int threads) {
  dct3(data, nx, ny, nz, inverse != 0, threads < 1 ? 1 : threads);
}

void *orc_reg_create(const double *final_locations, int threads,
                     const double *voxel_pos, int Nx_v, int Ny_v, int Nz_v,
                     int Nx_n, int Ny_n, int Nz_n, int spline_order,
                     double non_zero_voxels) {
  return new Registration(final_locations, threads, voxel_pos, Nx_v, Ny_v,
                          Nz_v, Nx_n, Ny_n, Nz_n, spline_order,
                          non_zero_voxels);
}
void orc_reg_destroy(void *h) { delete (Registration *)h; }

void orc_reg_set_mi(void *h, double alpha, const double *theta, int K, int L,
                    const int *binned_nodes, const int *binned_voxels) {
  auto *r = (Registration *)h;
  r->alpha = alpha;
  r->K = K;
  r->L = L;
  r->theta.assign(theta, theta + (size_t)K * L);
  r->H.assign((size_t)K * L, 0.0);
  r->bn.assign(binned_nodes, binned_nodes + r->nn);
  r->bv.assign(binned_voxels, binned_voxels + r->nv);
}

// setSSDLikelihood (Registration.cpp:472-479; bindings.cpp:120-152): the
// fixed and moving bins both in [0, bins)
void orc_reg_set_ssd(void *h, int bins, double sigma_sq, int scale_shift, double scale,
                     double shift, const int *binned_nodes, const int *binned_voxels,
                     double alpha_0, double beta_0) {
  auto *r = (Registration *)h;
  r->lik = 1;
  r->K = r->L = bins;
  r->sigma2 = sigma_sq;
  r->ssd_shift_scale = scale_shift;
  r->scale = scale;
  r->shift = shift;
  r->a0 = alpha_0;
  r->b0 = beta_0;
  r->ssd_sample = 0;
  r->H.assign((size_t)bins * bins, 0.0);
  r->bn.assign(binned_nodes, binned_nodes + r->nn);
  r->bv.assign(binned_voxels, binned_voxels + r->nv);
}
// SSDLikelihood::get_parameters (:328-334): [scale, shift]; sigma^2 third
void orc_reg_get_ssd(void *h, double *out3) {
  auto *r = (Registration *)h;
  out3[0] = r->scale;
  out3[1] = r->shift;
  out3[2] = r->sigma2;
}

// setAffineTransformation / setRigidTransformation (Registration.cpp:507-518;
// bindings.cpp:223-247): A is (N+1) x (N+1), R is N x N, t has N entries
void orc_reg_set_affine(void *h, const double *A, const double *t) {
  auto *r = (Registration *)h;
  const int n1 = r->N + 1;
  r->tmodel = 1;
  r->A_aff.assign(A, A + (size_t)n1 * n1);
  r->t_lin.assign(t, t + r->N);
}
void orc_reg_set_rigid(void *h, const double *R, const double *t) {
  auto *r = (Registration *)h;
  r->tmodel = 2;
  r->R_rig.assign(R, R + (size_t)r->N * r->N);
  r->t_lin.assign(t, t + r->N);
  // x means of the voxel positions (RigidTransformation.cpp:22-37)
  r->x_means.assign(r->N, 0.0);
  for (long long v = 0; v < r->nv; ++v)
    for (int n = 0; n < r->N; ++n) r->x_means[n] += r->xpos[v * r->N + n];
  for (int n = 0; n < r->N; ++n) r->x_means[n] /= (double)r->nv;
}
// one transformation step from the current dream locations (test hook)
double orc_reg_constrain(void *h) { return ((Registration *)h)->constrain(false, false); }
// one SSD parameter update / draw from the last E-step's histogram and its
// prior cost (test hooks)
void orc_reg_ssd_update(void *h, int draw) { ((Registration *)h)->ssd_update(draw != 0); }
double orc_reg_ssd_cost(void *h) { return ((Registration *)h)->ssd_cost(); }
void orc_reg_ssd_seed(void *h, int seed) { ((Registration *)h)->lk_gen.seed(seed); }

// the fitted transform: the matrix (affine (N+1)^2, rigid N^2) then t
int orc_reg_get_linear(void *h, double *out) {
  auto *r = (Registration *)h;
  const std::vector<double> &M = r->tmodel == 1 ? r->A_aff : r->R_rig;
  std::memcpy(out, M.data(), sizeof(double) * M.size());
  std::memcpy(out + M.size(), r->t_lin.data(), sizeof(double) * r->t_lin.size());
  return (int)(M.size() + r->t_lin.size());
}

void orc_reg_set_nlt(void *h, double gx, double gy, double gz, double rx,
                     double ry, double rz, double ex, double ey, double ez,
                     const double *D) {
  auto *r = (Registration *)h;
  double g[3] = {gx, gy, gz}, rr[3] = {rx, ry, rz}, e[3] = {ex, ey, ez};
  r->sm.reset(new Smoother(g, rr, e, D, r->nxv, r->nyv, r->nzv, r->threads,
                           r->sigma_spline(), r->offset()));
}

int orc_reg_em(void *h, int max_it, double th, int update, int debug) {
  return ((Registration *)h)->em(max_it, th, update != 0, debug != 0);
}

double orc_reg_estep(void *h, double *H_out) {
  auto *r = (Registration *)h;
  const double mll = r->estep(false, nullptr);
  if (H_out) std::memcpy(H_out, r->H.data(), sizeof(double) * r->K * r->L);
  return mll;
}

// Likelihood lookups / histogram accumulation / MAP update / prior cost over
// a list of (voxel, node) queries, the same protocol as oracle/ref_shim.cpp's
// ref_mi_step around the reference's MILikelihood object.
double orc_mi_step(double alpha, double *theta_io, int K, int L,
                   const int *binned_nodes, const int *binned_voxels, int Nxn,
                   int Nyn, int Nzn, int Nxv, int Nyv, int Nzv, int nq,
                   const int *q_voxel, const int *q_node_xyz,
                   const double *q_post, double *q_lik, int update) {
  std::vector<double> zero((size_t)Nxv * Nyv * Nzv * 3, 0.0);
  Registration r(zero.data(), 1, zero.data(), Nxv, Nyv, Nzv, Nxn, Nyn, Nzn, 3, 1.0);
  r.K = K;
  r.L = L;
  r.alpha = alpha;
  r.theta.assign(theta_io, theta_io + (size_t)K * L);
  r.bn.assign(binned_nodes, binned_nodes + (size_t)Nxn * Nyn * Nzn);
  r.bv.assign(binned_voxels, binned_voxels + (size_t)Nxv * Nyv * Nzv);
  r.H.assign((size_t)K * L, 0.0);
  for (int i = 0; i < nq; ++i) {
    const int *p = q_node_xyz + 3 * i;
    const int k = r.node_bin(p[0], p[1], p[2]);
    const int u = r.bv[q_voxel[i]];
    q_lik[i] = r.theta[(size_t)k * L + u];
    r.H[(size_t)k * L + u] += q_post[i];
  }
  if (update) r.theta_update();
  const double c = r.theta_cost();
  std::memcpy(theta_io, r.theta.data(), sizeof(double) * K * L);
  return c;
}

void orc_reg_theta_update(void *h) { ((Registration *)h)->theta_update(); }
double orc_reg_theta_cost(void *h) { return ((Registration *)h)->theta_cost(); }

void orc_reg_sampler(void *h, int seed, int Nb, int Ns, int sg, int every,
                     int sp, int sl, int st, double *mean, double *cov) {
  ((Registration *)h)->sampler(seed, Nb, Ns, sg, every, sp, sl, st, mean, cov);
}

void orc_reg_get_final(void *h, double *out) {
  auto *r = (Registration *)h;
  std::memcpy(out, r->final_loc.data(), sizeof(double) * r->final_loc.size());
}
void orc_reg_get_dream(void *h, double *out) {
  auto *r = (Registration *)h;
  std::memcpy(out, r->dream.data(), sizeof(double) * r->dream.size());
}
void orc_reg_set_dream(void *h, const double *in) {
  auto *r = (Registration *)h;
  std::memcpy(r->dream.data(), in, sizeof(double) * r->dream.size());
}
void orc_reg_get_theta(void *h, double *out) {
  auto *r = (Registration *)h;
  std::memcpy(out, r->theta.data(), sizeof(double) * r->theta.size());
}
// gamma of the transformation model: as set, or as the last nu draw left it
// (NonLinearTransformation.cpp:140-148)
void orc_reg_get_gamma(void *h, double *out3) {
  auto *r = (Registration *)h;
  for (int k = 0; k < 3; ++k) out3[k] = r->sm->gamma[k];
}
int orc_reg_get_trace(void *h, double *out, int max_rows) {
  auto *r = (Registration *)h;
  const int rows = (int)(r->trace.size() / 4);
  const int n = rows < max_rows ? rows : max_rows;
  if (out) std::memcpy(out, r->trace.data(), sizeof(double) * 4 * n);
  return rows;
}

// NonLinearTransformation standalone object (NonLinearTransformation.cpp:15-56)
void *orc_nlt_create(double gx, double gy, double gz, double rx, double ry,
                     double rz, double ex, double ey, double ez,
                     const double *D, int nx, int ny, int nz, int threads,
                     double sigma, double offset) {
  double g[3] = {gx, gy, gz}, r[3] = {rx, ry, rz}, e[3] = {ex, ey, ez};
  return new Smoother(g, r, e, D, nx, ny, nz, threads < 1 ? 1 : threads,
                      sigma, offset);
}
void orc_nlt_destroy(void *h) { delete (Smoother *)h; }

// smooth_deformation (NonLinearTransformation.cpp:226-278)
double orc_nlt_smooth(void *h, const double *in, double *out) {
  auto *s = (Smoother *)h;
  std::vector<double> zero((size_t)s->nv * s->N, 0.0);
  return s->constrain(in, zero.data(), out, false, false);
}

// sampleDeformationField (NonLinearTransformation.cpp:280-341); seed < 0 keeps
// the default-constructed engine like the reference's standalone object.
void orc_nlt_sample(void *h, const double *in, int S, long long seed,
                    double *out) {
  auto *s = (Smoother *)h;
  if (seed >= 0) s->gen.seed((unsigned)seed);
  std::vector<double> zero((size_t)s->nv * s->N, 0.0), fl((size_t)s->nv * s->N);
  for (int i = 0; i < S; ++i) {
    s->constrain(in, zero.data(), fl.data(), true, false);
    std::memcpy(out + (size_t)i * s->nv * s->N, fl.data(), sizeof(double) * fl.size());
  }
}

// up_sample (NonLinearTransformation.cpp:366-430)
void orc_up_sample(const double *field, int nx, int ny, int nz, int nxo,
                   int nyo, int nzo, int threads, double *out) {
  const int N = nzo == 1 ? 2 : 3;
  const long long nvi = (long long)nx * ny * nz, nvo = (long long)nxo * nyo * nzo;
  std::vector<double> a(nvi), b(nvo);
  for (int n = 0; n < N; ++n) {
    for (long long i = 0; i < nvi; ++i) a[i] = field[i * N + n];
    dct3(a.data(), nx, ny, nz, false, threads < 1 ? 1 : threads);
    std::fill(b.begin(), b.end(), 0.0);
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z)
          b[((long long)x * nyo + y) * nzo + z] = a[((long long)x * ny + y) * nz + z];
    dct3(b.data(), nxo, nyo, nzo, true, threads < 1 ? 1 : threads);
    for (long long i = 0; i < nvo; ++i) out[i * N + n] = b[i];
  }
}

}  // extern "C"
