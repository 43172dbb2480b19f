// ORACLE -- test infrastructure.  C wrapper around the reference's own
// FFTW/Eigen-free translation units, compiled from /root/reference into
// oracle/_ref/ (see Makefile).  Used only in this container to pin the oracle
// against the reference's actual arithmetic: the Jacobian determinant
// (BINDER/src/cpu/utils.cpp:10-345), the B-spline weights
// (BINDER/src/cpu/Spline.cpp:44-145) and the MI likelihood object
// (BINDER/src/cpu/MILikelihood.cpp:8-206).
#include <cstring>
#include <vector>

#include "MILikelihood.h"
#include "Spline.h"
#include "utils.h"

extern "C" {

void ref_jacobian_3d(const double *f, int Nx, int Ny, int Nz, double vx,
                     double vy, double vz, double *out) {
  std::vector<double> J = compute_jacobian_determinant_3D(f, Nx, Ny, Nz, vx, vy, vz);
  std::memcpy(out, J.data(), sizeof(double) * J.size());
}

void ref_jacobian_2d(const double *f, int Nx, int Ny, double vx, double vy,
                     double *out) {
  std::vector<double> J = compute_jacobian_determinant_2D(f, Nx, Ny, vx, vy);
  std::memcpy(out, J.data(), sizeof(double) * J.size());
}

void ref_spline(int order, int is2D, double d, double *out4, double *log4) {
  Spline s(order, is2D != 0);
  auto a = s.spline(d);
  auto b = s.log_spline(d);
  for (int i = 0; i < 4; ++i) {
    out4[i] = a[i];
    log4[i] = b[i];
  }
}

// Drives MILikelihood over a list of (voxel, node) lookups, the way
// Registration::estimate_posterior_and_dream_location uses it: likelihood
// values, accumulate_info with given posteriors, then update_parameters and
// compute_cost.  theta_io: [K*L] in, updated theta out.
double ref_mi_step(double alpha, double *theta_io, int K, int L,
                   const int *binned_nodes, const int *binned_voxels, int Nxn,
                   int Nyn, int Nzn, int Nxv, int Nyv, int Nzv, int nq,
                   const int *q_voxel, const int *q_node_xyz,
                   const double *q_post, double *q_lik, int update) {
  MILikelihood lk(alpha, theta_io, 1, binned_nodes, binned_voxels, K, L, Nxn,
                  Nyn, Nzn, Nxv, Nyv, Nzv);
  lk.initialize_parameters();
  for (int i = 0; i < nq; ++i) {
    const int *p = q_node_xyz + 3 * i;
    q_lik[i] = lk.compute_likelihood(q_voxel[i], p[0], p[1], p[2]);
    lk.accumulate_info(0, q_voxel[i], p[0], p[1], p[2], q_post[i]);
  }
  if (update) lk.update_parameters();
  const double c = lk.compute_cost();
  std::vector<double> t = lk.get_parameters();
  std::memcpy(theta_io, t.data(), sizeof(double) * t.size());
  return c;
}

}  // extern "C"
