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

  // The E-step voxel loop (Registration.cpp:144-181 and :551-571).
  // Histogram and mll are accumulated per x-slab and summed in slab order so
  // the oracle's result does not depend on the thread count.
  double estep(bool sample, std::vector<std::mt19937> *rngs) {
    std::fill(dream.begin(), dream.end(), 0.0);  // initialize_dream_locations
    std::vector<double> slabH((size_t)nxv * K * L, 0.0);
    std::vector<double> slab_mll(nxv, 0.0);
#pragma omp parallel for schedule(static, 1) num_threads(threads)
    for (int x = 0; x < nxv; ++x) {
      std::mt19937 *rng = rngs ? &(*rngs)[omp_get_thread_num()] : nullptr;
      double *Hs = &slabH[(size_t)x * K * L];
      double m = 0.0;
      for (int y = 0; y < nyv; ++y)
        for (int z = 0; z < nzv; ++z)
          m += voxel(((long long)x * nyv + y) * nzv + z, sample, rng, Hs);
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
      if (update) theta_update();
      const double pt = theta_cost();
      const double pd =
          sm->constrain(dream.data(), xpos.data(), final_loc.data(), false, false);
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
    sm->gen = rng;  // transformationObj->set_generator(rng)
    std::vector<double> avg((size_t)nv * N, 0.0), m2((size_t)nv * N * N, 0.0);
    for (int sweep = 0; sweep < Nb + Ns; ++sweep) {
      if (sp) estep(true, &rngs);
      if (sl) theta_sample();
      if (st) {
        const bool sg = sample_gamma && (sweep % every == 0);
        sm->constrain(dream.data(), xpos.data(), final_loc.data(), true, sg);
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
