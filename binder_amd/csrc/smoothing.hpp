// DCT-domain smoothing of a deformation on the GPU: the separable orthonormal
// DCT-II / DCT-III of each displacement component computed as per-axis
// matrix contractions, with the prior filter, the sampling noise and the
// prior cost fused into the pass that completes the forward transform.
// Replaces NonLinearTransformation::constrain_transformation /
// sample_transformation (BINDER/src/cpu/NonLinearTransformation.cpp:58-126,
// :190-223) and their cuFFT-based GPU counterparts.
#pragma once
#include <map>
#include <memory>

#include "common.hpp"

namespace binder {

// Dense per-axis transform matrices, FP32 on the device.
struct AxisMatrices {
  DBuf<float> fwd;  // [n][n]: fwd[k*n + j] = s_k cos(pi (2j+1) k / (2n))
  DBuf<float> inv;  // [n][n]: inv = fwd^T
};

struct SmoothOptions {
  bool sample = false;         // add DCT-domain Gaussian noise
  bool sample_gamma = false;   // accumulate the raw regulariser and draw nu
  uint32_t key0 = 0, key1 = 0; // Philox key
  uint32_t sweep = 0;          // Philox counter word
  uint32_t stream = 0;         // Philox stream word
  double *prior_out = nullptr; // device double: receives sum_n 0.5 g_n D p^2
};

class Smoother {
 public:
  // gamma/rho/eta as in NonLinearTransformation.cpp:15-43; D is [nx,ny,nz].
  Smoother(int nx, int ny, int nz, const double gamma[3], const double rho[3],
           const double eta[3], const float *D_host, double sigma,
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

  // Flops of one run (forward + inverse contractions over all components).
  double flops() const;

 private:
  const AxisMatrices &mats(int n);
  void launch_pass(int axis, bool forward, const float *in, float *out,
                   int flags, const float *xpro, const float *xepi,
                   const SmoothOptions &opt, KernelTimer *timer);

  int nx_, ny_, nz_, N_;
  long long nv_;
  double sigma_, rho_[3], eta_[3], nu_;
  hipStream_t stream_;
  DBuf<float> D_, bufA_, bufB_;
  DBuf<double> gamma_, part_cost_, part_raw_, scal_;
  std::map<int, std::unique_ptr<AxisMatrices>> mats_;
};

}  // namespace binder
