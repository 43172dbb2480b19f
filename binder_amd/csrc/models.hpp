// The reference's other likelihood and transformation models on the GPU:
// the SSD likelihood's parameter step (BINDER/src/cpu/SSDLikelihood.cpp) from
// the soft joint histogram (the E-step kernels evaluate the likelihood
// itself), and the affine / rigid transformations
// (BINDER/src/cpu/AffineTransformation.cpp, RigidTransformation.cpp) as FP64
// voxel reductions, a one-thread solve and an apply pass.
#pragma once
#include "common.hpp"
#include "smoothing.hpp"

namespace binder {

// ---------------------------------------------------------------------------
// SSD likelihood.  Parameters on the device: par[0] sigma^2, par[1] shift,
// par[2] scale, par[3] the reference's `sample` flag (SSDLikelihood.h:46:
// set by the first parameter draw, after which every update draws).
// ---------------------------------------------------------------------------
constexpr int kSsdParams = 4;
constexpr int kSsdMaxBins = 64;

enum SsdMode { kSsdKeep = 0, kSsdUpdate = 1, kSsdDraw = 2 };

struct SsdThetaArgs {
  // histogram: hsum (hrows partial rows, fixed point) or Hd as it stands
  const unsigned long long *hsum;
  int hrows;
  double *Hd;           // [bins][bins], written when hsum is set
  const double *mllpart;  // E-step mll partials (int64 fixed point when mll_fixed)
  int nmll, mll_fixed;
  const double *par_in;   // parameters the E-step used
  double *par_out;        // parameters after the step (may equal par_in)
  int bins, shift_scale, mode;
  double a0, b0;
  double *stats;  // stats[0] = mll, stats[1] = -log prior of sigma^2
  uint32_t k0, k1, sweep;
  const uint32_t *keyx;
};

// mll, the parameter step of `mode` and the prior cost, in one block (the
// E-step kernels evaluate the likelihood from the parameters themselves,
// EStepArgs::ssd_par).
void launch_ssd_theta(const SsdThetaArgs &a, hipStream_t s);

// ---------------------------------------------------------------------------
// Affine (kind 1) and rigid (kind 2) transformations.  Parameters on the
// device: the matrix (affine (N+1) x (N+1) with the homogeneous last row,
// rigid N x N, row-major) followed by t (N).
// ---------------------------------------------------------------------------
enum LinearKind { kLinAffine = 1, kLinRigid = 2 };

class LinearModel {
 public:
  // x: the voxel positions as N planes of nv floats (device); M0, t0 the
  // initial matrix and translation (host).
  LinearModel(int kind, int N, long long nv, const float *x, const double *M0,
              const double *t0, hipStream_t stream);
  // The fitted transform of `dream` applied to x, written to out (N planes);
  // prior cost 0 to opt.prior_out, and opt.host_stats as Smoother::run
  // writes them.  Enqueued on opt.exec_stream (or the model's stream).
  void run(const float *dream, const float *x, float *out, const SmoothOptions &opt);
  int kind() const { return kind_; }
  // matrix then t, as the parameters above; returns the count
  int get(double *out);

 private:
  int kind_, N_;
  long long nv_;
  hipStream_t stream_;
  int blocks_;
  DBuf<double> part_, par_, xstat_;
};

}  // namespace binder
