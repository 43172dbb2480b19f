// Learned warm start of the registration: device helpers shared with the
// registration handle (declarations).
#pragma once
#include "common.hpp"

namespace binder {

// Largest number of low-frequency DCT coefficients per axis and component
// that the warm-start network predicts (a network's own count, ncoef, is at
// most this).
constexpr int kMaxWarmCoef = 16;

// Writes the deformation planes d[c][nx][ny][nz] (c < 3) of a pyramid level
// at resolution `res` from normalised low-frequency coefficients
// coef[c][a][b][k] (a, b, k < ncoef) of the full-resolution displacement
// (grid fx, fy, fz):
//   u_c(i, j, k) = sqrt(fx fy fz) sum_{a,b,k'} coef[c][a][b][k'] C_fx[a][i] C_fy[b][j] C_fz[k'][k]
// with C_n the orthonormal DCT-II basis of length n, and
//   d_c(p) = x_c(p) + u_c(res * p) / res
// for level-grid voxel p (its full-resolution index is res * p), where x
// are the voxel positions [c][nv] (x == nullptr: the grid indices p).
void launch_dct_field(const float *coef_dev, int ncoef, int fx, int fy, int fz, int res,
                      const float *x, float *d, int nx, int ny, int nz, hipStream_t s);

}  // namespace binder
