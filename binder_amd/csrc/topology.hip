// Jacobian determinant, fold detection and topology correction, FP64.
//
// This file is compiled with -ffp-contract=off: the determinant expressions
// are evaluated with the operand order of the reference, so on the same
// inputs the Jacobians and therefore the fold masks and counts are bit-exact
// with the FP64 CPU implementation (BINDER/src/cpu/utils.cpp:10-345).
//
// Topology correction (BINDER/src/cpu/topology_correction_3D.cpp:17-607):
//  - corner Jacobians of every cell, kept as an 8-bit out-of-range mask
//    (the reference's J_corner values are only ever compared with e1/e2);
//  - periodic forward-difference gradients, 9 planes [n*3+a][voxel];
//  - gradient limiting: the reference updates gradients sequentially in
//    x -> y -> z -> corner order and later cells read what earlier cells
//    wrote.  Two cells interact only if they share an edge, i.e. their offset
//    lies in {-1,0,1}^3; with t = 4x + 2y + z every lexicographically earlier
//    neighbour has a smaller t and two cells with the same t never share an
//    edge.  Processing flagged cells hyperplane by hyperplane (t ascending,
//    cells of one hyperplane in parallel, corners of a cell in order) is
//    therefore exactly the sequential result;
//  - periodic least-squares integration: the reference's FFT formula
//    conj(e^{iw}-1) g_hat / sum(2-2cos w) equals L^+ s with
//    s = sum_a (g_a(j-e_a) - g_a(j)) and L the periodic grid Laplacian.  L is
//    diagonalised by the separable orthonormal discrete Hartley transform
//    (cas(2 pi j k / n) / sqrt(n), symmetric and self-inverse), so the solve
//    is three forward contractions, a 1/lambda scaling (DC -> 0), and three
//    inverse contractions -- the same machinery as the DCT smoothing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "topology.hpp"

#pragma clang fp contract(off)

namespace binder {
namespace {

__device__ __forceinline__ double det3(double a, double b, double c, double d,
                                       double e, double f, double g, double h,
                                       double i) {
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

struct Field {
  const double *x, *y, *z;  // SoA planes
};

// Per-voxel Jacobian (utils.cpp:100-345): corners fff, bff, fbf, ffb, bbf,
// bfb, fbb, bbb; forward 'f' = w(p+e)-w(p), backward 'b' = w(p)-w(p-e).
__device__ double voxel_jacobian(const Field w, int x, int y, int z, int Nx,
                                 int Ny, int Nz, double vol_scale) {
  const long long syz = (long long)Ny * Nz;
  const long long p = x * syz + (long long)y * Nz + z;
  const double px = w.x[p], py = w.y[p], pz = w.z[p];
  double total = 0.0;
  int valid = 0;
  const int cb[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                        {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int bx = cb[c][0], by = cb[c][1], bz = cb[c][2];
    if (bx ? x == 0 : x >= Nx - 1) continue;
    if (by ? y == 0 : y >= Ny - 1) continue;
    if (bz ? z == 0 : z >= Nz - 1) continue;
    const long long qx = bx ? p - syz : p + syz;
    const long long qy = by ? p - Nz : p + Nz;
    const long long qz = bz ? p - 1 : p + 1;
    double dx0, dx1, dx2, dy0, dy1, dy2, dz0, dz1, dz2;
    if (bx) { dx0 = px - w.x[qx]; dx1 = py - w.y[qx]; dx2 = pz - w.z[qx]; }
    else    { dx0 = w.x[qx] - px; dx1 = w.y[qx] - py; dx2 = w.z[qx] - pz; }
    if (by) { dy0 = px - w.x[qy]; dy1 = py - w.y[qy]; dy2 = pz - w.z[qy]; }
    else    { dy0 = w.x[qy] - px; dy1 = w.y[qy] - py; dy2 = w.z[qy] - pz; }
    if (bz) { dz0 = px - w.x[qz]; dz1 = py - w.y[qz]; dz2 = pz - w.z[qz]; }
    else    { dz0 = w.x[qz] - px; dz1 = w.y[qz] - py; dz2 = w.z[qz] - pz; }
    const double Jc = det3(dx0, dy0, dz0, dx1, dy1, dz1, dx2, dy2, dz2) * vol_scale;
    total += Jc;
    valid++;
  }
  return valid > 0 ? total / valid : 1.0;
}

// Block partials: [0] min, [1] max, [2] #(J<e1), [3] #(J>e2).
__global__ __launch_bounds__(256) void k_jacobian3d(Field w, int Nx, int Ny,
                                                    int Nz, double vol_scale,
                                                    double e1, double e2,
                                                    double *J, uint8_t *mask,
                                                    double *part) {
  const long long nv = (long long)Nx * Ny * Nz;
  double mn = INFINITY, mx = -INFINITY, lo = 0, hi = 0;
  // contiguous voxel range per logical block, logical blocks contiguous per
  // XCD: the 7-point stencil re-reads stay in one L2
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const long long per = (nv + gridDim.x - 1) / gridDim.x;
  const long long vend = ((long long)lb + 1) * per < nv ? ((long long)lb + 1) * per : nv;
  for (long long i = (long long)lb * per + threadIdx.x; i < vend; i += blockDim.x) {
    const int z = (int)(i % Nz);
    const long long r = i / Nz;
    const int y = (int)(r % Ny), x = (int)(r / Ny);
    const double v = voxel_jacobian(w, x, y, z, Nx, Ny, Nz, vol_scale);
    if (J) J[i] = v;
    if (mask) mask[i] = v < e1 ? 1 : 0;
    mn = fmin(mn, v);
    mx = fmax(mx, v);
    lo += v < e1 ? 1.0 : 0.0;
    hi += v > e2 ? 1.0 : 0.0;
  }
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  red[0][t] = mn;
  red[1][t] = mx;
  red[2][t] = lo;
  red[3][t] = hi;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = fmin(red[0][t], red[0][t + s]);
      red[1][t] = fmax(red[1][t], red[1][t + s]);
      red[2][t] += red[2][t + s];
      red[3][t] += red[3][t + s];
    }
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 4; ++k) part[lb * 4 + k] = red[k][0];
}

// 2D Jacobian (utils.cpp:10-98): corners ff, bf, fb, bb.
__device__ double voxel_jacobian2d(const double *wx, const double *wy, long long i,
                                   int Nx, int Ny, double vol_scale) {
  const int y = (int)(i % Ny), x = (int)(i / Ny);
  const double px = wx[i], py = wy[i];
  double total = 0.0;
  int valid = 0;
  const int cb[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int bx = cb[c][0], by = cb[c][1];
    if (bx ? x == 0 : x >= Nx - 1) continue;
    if (by ? y == 0 : y >= Ny - 1) continue;
    const long long qx = bx ? i - Ny : i + Ny;
    const long long qy = by ? i - 1 : i + 1;
    const double X1 = bx ? px - wx[qx] : wx[qx] - px;
    const double Y1 = bx ? py - wy[qx] : wy[qx] - py;
    const double X2 = by ? px - wx[qy] : wx[qy] - px;
    const double Y2 = by ? py - wy[qy] : wy[qy] - py;
    total += (X1 * Y2 - X2 * Y1) * vol_scale;
    valid++;
  }
  return valid > 0 ? total / valid : 1.0;
}

__global__ __launch_bounds__(256) void k_jacobian2d(const double *wx,
                                                    const double *wy, int Nx,
                                                    int Ny, double vol_scale,
                                                    double *J) {
  const long long nv = (long long)Nx * Ny;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  J[i] = voxel_jacobian2d(wx, wy, i, Nx, Ny, vol_scale);
}

// Block partials of the 2D Jacobian: [0] min, [1] max, [2] #(J<e1), [3] #(J>e2).
__global__ __launch_bounds__(256) void k_jacobian2d_stats(const double *wx,
                                                          const double *wy, int Nx,
                                                          int Ny, double vol_scale,
                                                          double e1, double e2,
                                                          double *part) {
  const long long nv = (long long)Nx * Ny;
  double mn = INFINITY, mx = -INFINITY, lo = 0, hi = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    const double v = voxel_jacobian2d(wx, wy, i, Nx, Ny, vol_scale);
    mn = fmin(mn, v);
    mx = fmax(mx, v);
    lo += v < e1 ? 1.0 : 0.0;
    hi += v > e2 ? 1.0 : 0.0;
  }
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  red[0][t] = mn;
  red[1][t] = mx;
  red[2][t] = lo;
  red[3][t] = hi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      red[0][t] = fmin(red[0][t], red[0][t + st]);
      red[1][t] = fmax(red[1][t], red[1][t + st]);
      red[2][t] += red[2][t + st];
      red[3][t] += red[3][t + st];
    }
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 4; ++k) part[blockIdx.x * 4 + k] = red[k][0];
}

// Vertex of each corner Jacobian of a cell in J_corner order fff, bff, fbf,
// ffb, fbb, bfb, bbf, bbb (topology_correction_3D.cpp:573-580, :452-486).
__constant__ int kVert[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

// Corner Jacobians of each cell -> out-of-range mask; block partial counts.
__global__ __launch_bounds__(256) void k_corner_mask(Field w, int Nx, int Ny,
                                                     int Nz, double vol_scale,
                                                     double e1, double e2,
                                                     uint8_t *cmask,
                                                     double *Jc_out,
                                                     unsigned long long *part) {
  const int Cy = Ny - 1, Cz = Nz - 1;
  const long long ncell = (long long)(Nx - 1) * Cy * Cz;
  unsigned long long lo = 0, hi = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       cell < ncell; cell += stride) {
    const int z = (int)(cell % Cz);
    const long long r = cell / Cz;
    const int y = (int)(r % Cy), x = (int)(r / Cy);
    uint8_t m = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int cx = kVert[c][0], cy = kVert[c][1], cz = kVert[c][2];
      auto at = [&](int a, int b, int d) {
        return ((long long)(x + a) * Ny + (y + b)) * Nz + (z + d);
      };
      const long long ex1 = at(1, cy, cz), ex0 = at(0, cy, cz);
      const long long ey1 = at(cx, 1, cz), ey0 = at(cx, 0, cz);
      const long long ez1 = at(cx, cy, 1), ez0 = at(cx, cy, 0);
      double J = det3(w.x[ex1] - w.x[ex0], w.x[ey1] - w.x[ey0], w.x[ez1] - w.x[ez0],
                      w.y[ex1] - w.y[ex0], w.y[ey1] - w.y[ey0], w.y[ez1] - w.y[ez0],
                      w.z[ex1] - w.z[ex0], w.z[ey1] - w.z[ey0], w.z[ez1] - w.z[ez0]);
      J *= vol_scale;
      if (Jc_out) Jc_out[cell * 8 + c] = J;
      if (J < e1) lo++;
      if (J > e2) hi++;
      if (J < e1 || J > e2) m |= (uint8_t)(1u << c);
    }
    cmask[cell] = m;
  }
  __shared__ unsigned long long red[2][256];
  const int t = threadIdx.x;
  red[0][t] = lo;
  red[1][t] = hi;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] += red[0][t + s];
      red[1][t] += red[1][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x * 2 + 0] = red[0][0];
    part[blockIdx.x * 2 + 1] = red[1][0];
  }
}

// compute_gradients_3D (:172-245): g[(n*3+a)][p] = (w_n(p+e_a) - w_n(p))/vs_n
// with periodic wrap.
__global__ __launch_bounds__(256) void k_gradients(Field w, int Nx, int Ny,
                                                   int Nz, double vs0,
                                                   double vs1, double vs2,
                                                   double *g) {
  const long long nv = (long long)Nx * Ny * Nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    const int z = (int)(i % Nz);
    const long long r = i / Nz;
    const int y = (int)(r % Ny), x = (int)(r / Ny);
    const int x2 = x + 1 > Nx - 1 ? x + 1 - Nx : x + 1;
    const int y2 = y + 1 > Ny - 1 ? y + 1 - Ny : y + 1;
    const int z2 = z + 1 > Nz - 1 ? z + 1 - Nz : z + 1;
    const long long nb[3] = {((long long)x2 * Ny + y) * Nz + z,
                             ((long long)x * Ny + y2) * Nz + z,
                             ((long long)x * Ny + y) * Nz + z2};
    const double *pl[3] = {w.x, w.y, w.z};
    const double vs[3] = {vs0, vs1, vs2};
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double p = pl[n][i];
#pragma unroll
      for (int a = 0; a < 3; ++a) g[(n * 3 + a) * nv + i] = (pl[n][nb[a]] - p) / vs[n];
    }
  }
}

// Flagged cells -> per-hyperplane buckets (t = 4x + 2y + z).
__global__ void k_bucket_count(const uint8_t *cmask, int Nx, int Ny, int Nz,
                               int *cnt) {
  const int Cy = Ny - 1, Cz = Nz - 1;
  const long long ncell = (long long)(Nx - 1) * Cy * Cz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += stride) {
    if (!cmask[c]) continue;
    const int z = (int)(c % Cz);
    const long long r = c / Cz;
    const int y = (int)(r % Cy), x = (int)(r / Cy);
    atomicAdd(&cnt[4 * x + 2 * y + z], 1);
  }
}

// Exclusive scan of the bucket counts (single block); start has T+1 entries.
__global__ __launch_bounds__(1024) void k_bucket_scan(const int *cnt, int T,
                                                      int *start, int *cursor) {
  __shared__ int s_tot[1024];
  const int t = threadIdx.x;
  const int per = (T + 1023) / 1024;
  const int b0 = t * per, b1 = min(T, b0 + per);
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += cnt[b];
  s_tot[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? s_tot[t - off] : 0;
    __syncthreads();
    s_tot[t] += v;
    __syncthreads();
  }
  int run = s_tot[t] - sum;
  for (int b = b0; b < b1; ++b) {
    start[b] = run;
    cursor[b] = run;
    run += cnt[b];
  }
  if (t == 1023) start[T] = s_tot[1023];
}

__global__ void k_bucket_scatter(const uint8_t *cmask, int Nx, int Ny, int Nz,
                                 int *cursor, int *list) {
  const int Cy = Ny - 1, Cz = Nz - 1;
  const long long ncell = (long long)(Nx - 1) * Cy * Cz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += stride) {
    if (!cmask[c]) continue;
    const int z = (int)(c % Cz);
    const long long r = c / Cz;
    const int y = (int)(r % Cy), x = (int)(r / Cy);
    const int pos = atomicAdd(&cursor[4 * x + 2 * y + z], 1);
    list[pos] = (int)c;
  }
}

// limit_gradients_3D (:247-332) in hyperplane order, one workgroup.
__global__ __launch_bounds__(1024) void k_limit(const uint8_t *cmask,
                                                const int *start, int T,
                                                const int *list, int Nx, int Ny,
                                                int Nz, double e1, double e2,
                                                double *g) {
  const long long nv = (long long)Nx * Ny * Nz;
  const int Cy = Ny - 1, Cz = Nz - 1;
  for (int t = 0; t < T; ++t) {
    const int b0 = start[t], b1 = start[t + 1];
    if (b0 == b1) continue;
    for (int k = b0 + (int)threadIdx.x; k < b1; k += blockDim.x) {
      const int cell = list[k];
      const int z = cell % Cz;
      const int r = cell / Cz;
      const int y = r % Cy, x = r / Cy;
      const uint8_t m = cmask[cell];
      for (int c = 0; c < 8; ++c) {
        if (!((m >> c) & 1)) continue;
        long long loc[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          int px = x + kVert[c][0], py = y + kVert[c][1], pz = z + kVert[c][2];
          if (a == 0) px = x;
          if (a == 1) py = y;
          if (a == 2) pz = z;
          loc[a] = ((long long)px * Ny + py) * Nz + pz;
        }
        double J[9];
#pragma unroll
        for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
          for (int n2 = 0; n2 < 3; ++n2)
            J[n1 * 3 + n2] = g[(n1 * 3 + n2) * nv + loc[n2]];
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double alpha = 0.0;
        double det = det3(J[0], J[1], J[2], J[3], J[4], J[5], J[6], J[7], J[8]);
        int guard = 0;
        while ((det > e2 || det < e1) && guard < 64) {
          alpha += 0.1;
          if (alpha > 1.0) alpha = 1.0;
          double Jn[9];
#pragma unroll
          for (int q = 0; q < 9; ++q) Jn[q] = (1 - alpha) * J[q] + alpha * I[q];
          det = det3(Jn[0], Jn[1], Jn[2], Jn[3], Jn[4], Jn[5], Jn[6], Jn[7], Jn[8]);
          if (alpha >= 1.0) ++guard;
        }
        alpha += 0.1;
        if (alpha > 1.0) alpha = 1.0;
#pragma unroll
        for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
          for (int n2 = 0; n2 < 3; ++n2)
            g[(n1 * 3 + n2) * nv + loc[n2]] =
                (1 - alpha) * J[n1 * 3 + n2] + alpha * I[n1 * 3 + n2];
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

// Block partial sums of the three coordinate planes.
__global__ __launch_bounds__(256) void k_plane_sums(Field w, long long nv,
                                                    double *part) {
  double s[3] = {0, 0, 0};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    s[0] += w.x[i];
    s[1] += w.y[i];
    s[2] += w.z[i];
  }
  __shared__ double red[3][256];
  const int t = threadIdx.x;
  for (int k = 0; k < 3; ++k) red[k][t] = s[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st)
      for (int k = 0; k < 3; ++k) red[k][t] += red[k][t + st];
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 3; ++k) part[blockIdx.x * 3 + k] = red[k][0];
}

// s(p) = sum_a g_{n,a}(p - e_a) - g_{n,a}(p), periodic.
__global__ __launch_bounds__(256) void k_divergence(const double *g, int n,
                                                    int Nx, int Ny, int Nz,
                                                    double *s) {
  const long long nv = (long long)Nx * Ny * Nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const double *gx = g + (n * 3 + 0) * nv, *gy = g + (n * 3 + 1) * nv,
               *gz = g + (n * 3 + 2) * nv;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    const int z = (int)(i % Nz);
    const long long r = i / Nz;
    const int y = (int)(r % Ny), x = (int)(r / Ny);
    const int xm = x == 0 ? Nx - 1 : x - 1;
    const int ym = y == 0 ? Ny - 1 : y - 1;
    const int zm = z == 0 ? Nz - 1 : z - 1;
    const double v = (gx[((long long)xm * Ny + y) * Nz + z] - gx[i]) +
                     (gy[((long long)x * Ny + ym) * Nz + z] - gy[i]) +
                     (gz[((long long)x * Ny + y) * Nz + zm] - gz[i]);
    s[i] = v;
  }
}

// ---- 2D topology correction (topology_correction_2D.cpp:17-430) -------------
// Corner Jacobians Jff, Jbf, Jfb, Jbb of each cell (:156-200) -> 4-bit
// out-of-range mask; block partial counts of cells with any corner below e1
// / above e2.
__global__ __launch_bounds__(256) void k_corner_mask2d(const double *wx,
                                                       const double *wy, int Nx,
                                                       int Ny, double vol_scale,
                                                       double e1, double e2,
                                                       uint8_t *cmask,
                                                       unsigned long long *part) {
  const int Cy = Ny - 1;
  const long long ncell = (long long)(Nx - 1) * Cy;
  unsigned long long lo = 0, hi = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       cell < ncell; cell += stride) {
    const int y = (int)(cell % Cy), x = (int)(cell / Cy);
    const long long i00 = (long long)x * Ny + y, i01 = i00 + 1, i10 = i00 + Ny,
                    i11 = i10 + 1;
    const double wx00 = wx[i00], wx01 = wx[i01], wx10 = wx[i10], wx11 = wx[i11];
    const double wy00 = wy[i00], wy01 = wy[i01], wy10 = wy[i10], wy11 = wy[i11];
    const double J[4] = {
        ((wx10 - wx00) * (wy01 - wy00) - (wx01 - wx00) * (wy10 - wy00)) * vol_scale,
        ((wx10 - wx00) * (wy11 - wy10) - (wx11 - wx10) * (wy10 - wy00)) * vol_scale,
        ((wx11 - wx01) * (wy01 - wy00) - (wx01 - wx00) * (wy11 - wy01)) * vol_scale,
        ((wx11 - wx01) * (wy11 - wy10) - (wx11 - wx10) * (wy11 - wy01)) * vol_scale};
    uint8_t m = 0;
    bool l = false, h = false;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      l = l || J[c] < e1;
      h = h || J[c] > e2;
      if (J[c] < e1 || J[c] > e2) m |= (uint8_t)(1u << c);
    }
    lo += l ? 1 : 0;
    hi += h ? 1 : 0;
    cmask[cell] = m;
  }
  __shared__ unsigned long long red[2][256];
  const int t = threadIdx.x;
  red[0][t] = lo;
  red[1][t] = hi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      red[0][t] += red[0][t + st];
      red[1][t] += red[1][t + st];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x * 2 + 0] = red[0][0];
    part[blockIdx.x * 2 + 1] = red[1][0];
  }
}

// compute_gradients_2D (:202-245): g[(n*2+a)][p] = (w_n(p+e_a) - w_n(p))/vs_n,
// periodic.
__global__ __launch_bounds__(256) void k_gradients2d(const double *wx,
                                                     const double *wy, int Nx,
                                                     int Ny, double vs0,
                                                     double vs1, double *g) {
  const long long nv = (long long)Nx * Ny;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    const int y = (int)(i % Ny), x = (int)(i / Ny);
    const int x2 = x + 1 > Nx - 1 ? x + 1 - Nx : x + 1;
    const int y2 = y + 1 > Ny - 1 ? y + 1 - Ny : y + 1;
    const long long ix = (long long)x2 * Ny + y, iy = (long long)x * Ny + y2;
    g[0 * nv + i] = (wx[ix] - wx[i]) / vs0;
    g[1 * nv + i] = (wx[iy] - wx[i]) / vs0;
    g[2 * nv + i] = (wy[ix] - wy[i]) / vs1;
    g[3 * nv + i] = (wy[iy] - wy[i]) / vs1;
  }
}

// Flagged cells -> hyperplane buckets t = 2x + y: every lexicographically
// earlier cell within {-1,0,1}^2 has a smaller t and two cells of one
// hyperplane never share a gradient entry.
__global__ void k_bucket2d(const uint8_t *cmask, int Nx, int Ny, int *cnt,
                           int *cursor, int *list, int scatter) {
  const int Cy = Ny - 1;
  const long long ncell = (long long)(Nx - 1) * Cy;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += stride) {
    if (!cmask[c]) continue;
    const int y = (int)(c % Cy), x = (int)(c / Cy);
    if (scatter) list[atomicAdd(&cursor[2 * x + y], 1)] = (int)c;
    else atomicAdd(&cnt[2 * x + y], 1);
  }
}

// Column offsets of each 2D corner (get_jac_off_2D, :394-410).
__constant__ int kOff2[4][2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {1, 0}},
                                   {{0, 1}, {0, 0}}, {{0, 1}, {1, 0}}};

// limit_gradients_2D (:247-318) in hyperplane order, one workgroup.
__global__ __launch_bounds__(1024) void k_limit2d(const uint8_t *cmask,
                                                  const int *start, int T,
                                                  const int *list, int Nx,
                                                  int Ny, double e1, double e2,
                                                  double *g) {
  const long long nv = (long long)Nx * Ny;
  const int Cy = Ny - 1;
  for (int t = 0; t < T; ++t) {
    const int b0 = start[t], b1 = start[t + 1];
    if (b0 == b1) continue;
    for (int k = b0 + (int)threadIdx.x; k < b1; k += blockDim.x) {
      const int cell = list[k];
      const int y = cell % Cy, x = cell / Cy;
      const uint8_t m = cmask[cell];
      for (int c = 0; c < 4; ++c) {
        if (!((m >> c) & 1)) continue;
        long long loc[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
          loc[a] = (long long)(x + kOff2[c][a][0]) * Ny + (y + kOff2[c][a][1]);
        double J[4];
#pragma unroll
        for (int n1 = 0; n1 < 2; ++n1)
#pragma unroll
          for (int n2 = 0; n2 < 2; ++n2) J[n1 * 2 + n2] = g[(n1 * 2 + n2) * nv + loc[n2]];
        const double I[4] = {1, 0, 0, 1};
        double alpha = 0.0;
        double det = J[0] * J[3] - J[1] * J[2];
        int guard = 0;
        while ((det > e2 || det < e1) && guard < 64) {
          alpha += 0.1;
          if (alpha > 1.0) alpha = 1.0;
          double Jn[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) Jn[q] = (1 - alpha) * J[q] + alpha * I[q];
          det = Jn[0] * Jn[3] - Jn[1] * Jn[2];
          if (alpha >= 1.0) ++guard;
        }
        alpha += 0.1;
        if (alpha > 1.0) alpha = 1.0;
#pragma unroll
        for (int n1 = 0; n1 < 2; ++n1)
#pragma unroll
          for (int n2 = 0; n2 < 2; ++n2)
            g[(n1 * 2 + n2) * nv + loc[n2]] =
                (1 - alpha) * J[n1 * 2 + n2] + alpha * I[n1 * 2 + n2];
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

// s(p) = sum_a g_{n,a}(p - e_a) - g_{n,a}(p) over a = x, y, periodic.
__global__ __launch_bounds__(256) void k_divergence2d(const double *g, int n,
                                                      int Nx, int Ny, double *s) {
  const long long nv = (long long)Nx * Ny;
  const double *gx = g + (n * 2 + 0) * nv, *gy = g + (n * 2 + 1) * nv;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    const int y = (int)(i % Ny), x = (int)(i / Ny);
    const int xm = x == 0 ? Nx - 1 : x - 1;
    const int ym = y == 0 ? Ny - 1 : y - 1;
    s[i] = (gx[(long long)xm * Ny + y] - gx[i]) + (gy[(long long)x * Ny + ym] - gy[i]);
  }
}

// FP64 per-axis contraction (same scheme as the FP32 DCT passes).
constexpr int TBM = 64, TBN = 64, TBK = 8;
enum : int { TEPI_STORE = 0, TEPI_INVLAMBDA = 1, TEPI_FINAL = 2 };

struct DPass {
  const double *A, *B;
  double *C;
  int M, N, K, lda, ldb, ldc;
  long long sAb, sBb, sCb;
  int nb;
  int Nx, Ny, Nz;     // for the 1/lambda epilogue
  double scale, add;  // final epilogue: C = acc * scale + add
};

template <bool DATA_LEFT, int EPI>
__global__ __launch_bounds__(256) void k_dpass(DPass a) {
  __shared__ double As[TBK][TBM + 2];
  __shared__ double Bs[TBK][TBN + 2];
  const int tid = threadIdx.x;
  const int tn = tid & 15, tm = tid >> 4;
  const int n0 = blockIdx.x * TBN, m0 = blockIdx.y * TBM;
  const int inner = blockIdx.z;
  const double *A = a.A + inner * a.sAb;
  const double *B = a.B + inner * a.sBb;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < a.K; k0 += TBK) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int e = tid + 256 * i;
      const int r = e >> 3, c = e & 7;
      const int gm = m0 + r, gk = k0 + c;
      As[c][r] = (gm < a.M && gk < a.K) ? A[(long long)gm * a.lda + gk] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int e = tid + 256 * i;
      const int r = e >> 6, c = e & 63;
      const int gk = k0 + r, gn = n0 + c;
      Bs[r][c] = (gk < a.K && gn < a.N) ? B[(long long)gk * a.ldb + gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TBK; ++kk) {
      double ar[4], br[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) ar[i] = As[kk][tm * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) br[j] = Bs[kk][tn * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(ar[i], br[j], acc[i][j]);
    }
    __syncthreads();
  }
  double *C = a.C + inner * a.sCb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + tm * 4 + i;
    if (gm >= a.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tn * 4 + j;
      if (gn >= a.N) continue;
      const long long off = (long long)gm * a.ldc + gn;
      double v = acc[i][j];
      if (EPI == TEPI_INVLAMBDA) {
        const long long o = inner * a.sCb + off;
        const int kz = (int)(o % a.Nz);
        const long long rr = o / a.Nz;
        const int ky = (int)(rr % a.Ny), kx = (int)(rr / a.Ny);
        const double lam = (2.0 - 2.0 * cos(2.0 * M_PI * kx / a.Nx)) +
                           (2.0 - 2.0 * cos(2.0 * M_PI * ky / a.Ny)) +
                           (2.0 - 2.0 * cos(2.0 * M_PI * kz / a.Nz));
        v = (kx == 0 && ky == 0 && kz == 0) ? 0.0 : v / lam;
      } else if (EPI == TEPI_FINAL) {
        v = v * a.scale + a.add;
      }
      C[off] = v;
    }
  }
}

template <class T>
struct HostVec {
  std::vector<T> v;
};

struct TopoCtx {
  hipStream_t s = nullptr;
  int Nx, Ny, Nz;
  long long nv, ncell;
  double vs[3], e1, e2, vol_scale;
  DBuf<double> w, g, part, sbuf, tbuf, ubuf, Jmat[3];
  DBuf<uint8_t> cmask;
  DBuf<unsigned long long> cpart;
  DBuf<int> cnt, start, cursor, list;
  int T = 0;
  int blocks = 0;
  std::vector<double> hpart;
  std::vector<unsigned long long> hcpart;

  Field field() const {
    return Field{w.get(), w.get() + nv, w.get() + 2 * nv};
  }

  void jstats(double &jmin, double &jmax, long long &nb, long long &na) {
    hipLaunchKernelGGL(k_jacobian3d, dim3(blocks), dim3(256), 0, s, field(), Nx,
                       Ny, Nz, vol_scale, e1, e2, (double *)nullptr,
                       (uint8_t *)nullptr, part.get());
    HIP_LAUNCH_OK();
    hpart.resize((size_t)blocks * 4);
    part.download(hpart.data(), hpart.size(), s);
    HIP_OK(hipStreamSynchronize(s));
    jmin = INFINITY;
    jmax = -INFINITY;
    double lo = 0, hi = 0;
    for (int b = 0; b < blocks; ++b) {
      jmin = std::min(jmin, hpart[b * 4 + 0]);
      jmax = std::max(jmax, hpart[b * 4 + 1]);
      lo += hpart[b * 4 + 2];
      hi += hpart[b * 4 + 3];
    }
    nb = (long long)lo;
    na = (long long)hi;
  }

  void corner(long long &lo, long long &hi) {
    hipLaunchKernelGGL(k_corner_mask, dim3(blocks), dim3(256), 0, s, field(),
                       Nx, Ny, Nz, vol_scale, e1, e2, cmask.get(),
                       (double *)nullptr, cpart.get());
    HIP_LAUNCH_OK();
    hcpart.resize((size_t)blocks * 2);
    cpart.download(hcpart.data(), hcpart.size(), s);
    HIP_OK(hipStreamSynchronize(s));
    lo = hi = 0;
    for (int b = 0; b < blocks; ++b) {
      lo += (long long)hcpart[b * 2];
      hi += (long long)hcpart[b * 2 + 1];
    }
  }

  void dht_matrix(int n, DBuf<double> &m) {
    std::vector<double> h((size_t)n * n);
    const double sc = 1.0 / std::sqrt((double)n);
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        const long long kj = ((long long)k * j) % n;
        const double a = 2.0 * M_PI * (double)kj / n;
        h[(size_t)k * n + j] = (std::cos(a) + std::sin(a)) * sc;
      }
    m.upload(h.data(), h.size(), s);
    HIP_OK(hipStreamSynchronize(s));
  }

  // One separable transform pass (the DHT matrix is symmetric and its own
  // inverse, so forward and inverse passes use the same matrix).
  void dpass(int axis, const double *in, double *out, int epi, double scale,
             double add) {
    DPass a{};
    a.Nx = Nx;
    a.Ny = Ny;
    a.Nz = Nz;
    a.scale = scale;
    a.add = add;
    const int dims[3] = {Nx, Ny, Nz};
    const int n = dims[axis];
    a.C = out;
    if (axis == 2) {
      a.A = in;
      a.B = Jmat[2].get();
      a.M = (int)(nv / n);
      a.N = n;
      a.K = n;
      a.lda = n;
      a.ldb = n;
      a.ldc = n;
      a.nb = 1;
    } else {
      a.A = Jmat[axis].get();
      a.B = in;
      a.M = n;
      a.K = n;
      a.lda = n;
      if (axis == 1) {
        a.N = Nz;
        a.ldb = Nz;
        a.ldc = Nz;
        a.nb = Nx;
        a.sBb = (long long)Ny * Nz;
        a.sCb = (long long)Ny * Nz;
      } else {
        a.N = Ny * Nz;
        a.ldb = Ny * Nz;
        a.ldc = Ny * Nz;
        a.nb = 1;
      }
    }
    dim3 grid((a.N + TBN - 1) / TBN, (a.M + TBM - 1) / TBM, a.nb);
    const bool dl = axis == 2;
    if (dl) {
      if (epi == TEPI_FINAL) hipLaunchKernelGGL((k_dpass<true, TEPI_FINAL>), grid, dim3(256), 0, s, a);
      else if (epi == TEPI_INVLAMBDA) hipLaunchKernelGGL((k_dpass<true, TEPI_INVLAMBDA>), grid, dim3(256), 0, s, a);
      else hipLaunchKernelGGL((k_dpass<true, TEPI_STORE>), grid, dim3(256), 0, s, a);
    } else {
      if (epi == TEPI_FINAL) hipLaunchKernelGGL((k_dpass<false, TEPI_FINAL>), grid, dim3(256), 0, s, a);
      else if (epi == TEPI_INVLAMBDA) hipLaunchKernelGGL((k_dpass<false, TEPI_INVLAMBDA>), grid, dim3(256), 0, s, a);
      else hipLaunchKernelGGL((k_dpass<false, TEPI_STORE>), grid, dim3(256), 0, s, a);
    }
    HIP_LAUNCH_OK();
  }

  void limit() {
    cnt.zero(s);
    const int gb = (int)std::min<long long>((ncell + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bucket_count, dim3(gb), dim3(256), 0, s, cmask.get(),
                       Nx, Ny, Nz, cnt.get());
    HIP_LAUNCH_OK();
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, cnt.get(), T,
                       start.get(), cursor.get());
    HIP_LAUNCH_OK();
    hipLaunchKernelGGL(k_bucket_scatter, dim3(gb), dim3(256), 0, s, cmask.get(),
                       Nx, Ny, Nz, cursor.get(), list.get());
    HIP_LAUNCH_OK();
    hipLaunchKernelGGL(k_limit, dim3(1), dim3(1024), 0, s, cmask.get(),
                       start.get(), T, list.get(), Nx, Ny, Nz, e1, e2, g.get());
    HIP_LAUNCH_OK();
  }

  void integrate() {
    // mean of the current positions (compute_mean_coordinates :334-358)
    hipLaunchKernelGGL(k_plane_sums, dim3(blocks), dim3(256), 0, s, field(), nv,
                       part.get());
    HIP_LAUNCH_OK();
    hpart.resize((size_t)blocks * 3);
    part.download(hpart.data(), hpart.size(), s);
    HIP_OK(hipStreamSynchronize(s));
    double mean[3] = {0, 0, 0};
    for (int b = 0; b < blocks; ++b)
      for (int k = 0; k < 3; ++k) mean[k] += hpart[b * 3 + k];
    for (int k = 0; k < 3; ++k) mean[k] /= (double)nv;
    for (int n = 0; n < 3; ++n) {
      hipLaunchKernelGGL(k_divergence, dim3(blocks), dim3(256), 0, s, g.get(),
                         n, Nx, Ny, Nz, sbuf.get());
      HIP_LAUNCH_OK();
      dpass(2, sbuf.get(), tbuf.get(), TEPI_STORE, 1, 0);
      dpass(1, tbuf.get(), ubuf.get(), TEPI_STORE, 1, 0);
      dpass(0, ubuf.get(), tbuf.get(), TEPI_INVLAMBDA, 1, 0);
      dpass(0, tbuf.get(), ubuf.get(), TEPI_STORE, 1, 0);
      dpass(1, ubuf.get(), tbuf.get(), TEPI_STORE, 1, 0);
      dpass(2, tbuf.get(), w.get() + n * nv, TEPI_FINAL, vs[n], mean[n]);
    }
  }
};

__global__ void k_scale_planes(double *w, long long n, double a, double b,
                               double c, int divide) {
  const long long nv = n;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += stride) {
    if (divide) {
      w[i] /= a;
      w[nv + i] /= b;
      w[2 * nv + i] /= c;
    } else {
      w[i] *= a;
      w[nv + i] *= b;
      w[2 * nv + i] *= c;
    }
  }
}

struct OwnedStream {
  hipStream_t s = nullptr;
  OwnedStream() { HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~OwnedStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

void upload_soa3(const double *aos, long long nv, DBuf<double> &dst,
                 hipStream_t s) {
  std::vector<double> h((size_t)3 * nv);
  for (long long i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) h[k * nv + i] = aos[i * 3 + k];
  dst.upload(h.data(), h.size(), s);
  HIP_OK(hipStreamSynchronize(s));
}

void topo_init(TopoCtx &c, hipStream_t s, int Nx, int Ny, int Nz, const double vs[3],
               double e1, double e2) {
  c.s = s;
  c.Nx = Nx;
  c.Ny = Ny;
  c.Nz = Nz;
  c.nv = (long long)Nx * Ny * Nz;
  c.ncell = (long long)(Nx - 1) * (Ny - 1) * (Nz - 1);
  for (int k = 0; k < 3; ++k) c.vs[k] = vs[k];
  c.e1 = e1;
  c.e2 = e2;
  c.vol_scale = 1.0 / (c.vs[0] * c.vs[1] * c.vs[2]);
  c.blocks = (int)std::min<long long>((c.nv + 255) / 256, 4096);
  c.part.alloc((size_t)c.blocks * 4);
  c.cpart.alloc((size_t)c.blocks * 2);
}

// 3D correction of the grid-unit field c.w (three f64 planes, on the
// device): to physical units, the outer/inner loop of correct_topology
// (topology_correction_3D.cpp:17-143), back to grid units.
void topo3_run(TopoCtx &c, int max_outer, int max_inner, double *stats) {
  hipLaunchKernelGGL(k_scale_planes, dim3(c.blocks), dim3(256), 0, c.s,
                     c.w.get(), c.nv, c.vs[0], c.vs[1], c.vs[2], 0);
  HIP_LAUNCH_OK();
  double jmin, jmax;
  long long nb, na;
  c.jstats(jmin, jmax, nb, na);
  int outer = 0, inner_total = 0;
  if (!(jmin >= c.e1 && jmax <= c.e2)) {
    c.cmask.alloc(c.ncell);
    c.g.alloc((size_t)9 * c.nv);
    c.sbuf.alloc(c.nv);
    c.tbuf.alloc(c.nv);
    c.ubuf.alloc(c.nv);
    c.T = 4 * (c.Nx - 2) + 2 * (c.Ny - 2) + (c.Nz - 2) + 1;
    c.cnt.alloc(c.T);
    c.start.alloc(c.T + 1);
    c.cursor.alloc(c.T);
    c.list.alloc(c.ncell);
    c.dht_matrix(c.Nx, c.Jmat[0]);
    c.dht_matrix(c.Ny, c.Jmat[1]);
    c.dht_matrix(c.Nz, c.Jmat[2]);
    while ((jmin < c.e1 || jmax > c.e2) && outer < max_outer) {
      long long lo, hi;
      c.corner(lo, hi);
      int inner = 0;
      while ((lo > 0 || hi > 0) && inner < max_inner) {
        hipLaunchKernelGGL(k_gradients, dim3(c.blocks), dim3(256), 0, c.s,
                           c.field(), c.Nx, c.Ny, c.Nz, c.vs[0], c.vs[1],
                           c.vs[2], c.g.get());
        HIP_LAUNCH_OK();
        c.limit();
        c.integrate();
        c.corner(lo, hi);
        inner++;
      }
      inner_total += inner;
      c.jstats(jmin, jmax, nb, na);
      outer++;
    }
  }
  hipLaunchKernelGGL(k_scale_planes, dim3(c.blocks), dim3(256), 0, c.s,
                     c.w.get(), c.nv, c.vs[0], c.vs[1], c.vs[2], 1);
  HIP_LAUNCH_OK();
  if (stats) {
    stats[0] = outer;
    stats[1] = inner_total;
    stats[2] = jmin;
    stats[3] = jmax;
    stats[4] = (double)nb;
    stats[5] = (double)na;
  }
}

// 2D correction of the grid-unit field c.w (planes x, y, and a zero third
// plane, on the device), c set up on the grid (Nx, 1, Ny)
// (topology_correction_2D.cpp:17-140).
void topo2_run(TopoCtx &c, int Nx_v, int Ny_v, double vs0, double vs1, double e1,
             double e2, int max_outer, int max_inner, double *stats) {
  const long long nv = c.nv;
  const long long ncell = (long long)(Nx_v - 1) * (Ny_v - 1);
  const double vol_scale = 1.0 / (vs0 * vs1);
  double *wx = c.w.get(), *wy = c.w.get() + nv;
  hipLaunchKernelGGL(k_scale_planes, dim3(c.blocks), dim3(256), 0, c.s, c.w.get(),
                     nv, vs0, vs1, 1.0, 0);
  HIP_LAUNCH_OK();
  auto jstats = [&](double &jmin, double &jmax, long long &nb, long long &na) {
    hipLaunchKernelGGL(k_jacobian2d_stats, dim3(c.blocks), dim3(256), 0, c.s, wx, wy,
                       Nx_v, Ny_v, vol_scale, e1, e2, c.part.get());
    HIP_LAUNCH_OK();
    c.hpart.resize((size_t)c.blocks * 4);
    c.part.download(c.hpart.data(), c.hpart.size(), c.s);
    HIP_OK(hipStreamSynchronize(c.s));
    jmin = INFINITY;
    jmax = -INFINITY;
    double lo = 0, hi = 0;
    for (int b = 0; b < c.blocks; ++b) {
      jmin = std::min(jmin, c.hpart[b * 4 + 0]);
      jmax = std::max(jmax, c.hpart[b * 4 + 1]);
      lo += c.hpart[b * 4 + 2];
      hi += c.hpart[b * 4 + 3];
    }
    nb = (long long)lo;
    na = (long long)hi;
  };
  auto corner = [&](long long &lo, long long &hi) {
    hipLaunchKernelGGL(k_corner_mask2d, dim3(c.blocks), dim3(256), 0, c.s, wx, wy,
                       Nx_v, Ny_v, vol_scale, e1, e2, c.cmask.get(), c.cpart.get());
    HIP_LAUNCH_OK();
    c.hcpart.resize((size_t)c.blocks * 2);
    c.cpart.download(c.hcpart.data(), c.hcpart.size(), c.s);
    HIP_OK(hipStreamSynchronize(c.s));
    lo = hi = 0;
    for (int b = 0; b < c.blocks; ++b) {
      lo += (long long)c.hcpart[b * 2];
      hi += (long long)c.hcpart[b * 2 + 1];
    }
  };
  double jmin, jmax;
  long long nb, na;
  jstats(jmin, jmax, nb, na);
  int outer = 0, inner_total = 0;
  if (!(jmin >= e1 && jmax <= e2)) {
    c.cmask.alloc(ncell);
    c.g.alloc((size_t)4 * nv);
    c.sbuf.alloc(nv);
    c.tbuf.alloc(nv);
    c.ubuf.alloc(nv);
    c.T = 2 * (Nx_v - 2) + (Ny_v - 2) + 1;
    c.cnt.alloc(c.T);
    c.start.alloc(c.T + 1);
    c.cursor.alloc(c.T);
    c.list.alloc(ncell);
    c.dht_matrix(Nx_v, c.Jmat[0]);
    c.dht_matrix(Ny_v, c.Jmat[2]);
    const int gb = (int)std::min<long long>((ncell + 255) / 256, 4096);
    while ((jmin < e1 || jmax > e2) && outer < max_outer) {
      long long lo, hi;
      corner(lo, hi);
      int inner = 0;
      while ((lo > 0 || hi > 0) && inner < max_inner) {
        hipLaunchKernelGGL(k_gradients2d, dim3(c.blocks), dim3(256), 0, c.s, wx, wy,
                           Nx_v, Ny_v, vs0, vs1, c.g.get());
        HIP_LAUNCH_OK();
        // limit_gradients_2D in hyperplane order
        c.cnt.zero(c.s);
        hipLaunchKernelGGL(k_bucket2d, dim3(gb), dim3(256), 0, c.s, c.cmask.get(), Nx_v,
                           Ny_v, c.cnt.get(), (int *)nullptr, (int *)nullptr, 0);
        HIP_LAUNCH_OK();
        hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, c.s, c.cnt.get(), c.T,
                           c.start.get(), c.cursor.get());
        HIP_LAUNCH_OK();
        hipLaunchKernelGGL(k_bucket2d, dim3(gb), dim3(256), 0, c.s, c.cmask.get(), Nx_v,
                           Ny_v, (int *)nullptr, c.cursor.get(), c.list.get(), 1);
        HIP_LAUNCH_OK();
        hipLaunchKernelGGL(k_limit2d, dim3(1), dim3(1024), 0, c.s, c.cmask.get(),
                           c.start.get(), c.T, c.list.get(), Nx_v, Ny_v, e1, e2,
                           c.g.get());
        HIP_LAUNCH_OK();
        // compute_mean_coordinates, then integrate_gradient_field_2D
        hipLaunchKernelGGL(k_plane_sums, dim3(c.blocks), dim3(256), 0, c.s,
                           Field{wx, wy, wy}, nv, c.part.get());
        HIP_LAUNCH_OK();
        c.hpart.resize((size_t)c.blocks * 3);
        c.part.download(c.hpart.data(), c.hpart.size(), c.s);
        HIP_OK(hipStreamSynchronize(c.s));
        double mean[2] = {0, 0};
        for (int b = 0; b < c.blocks; ++b)
          for (int k = 0; k < 2; ++k) mean[k] += c.hpart[b * 3 + k];
        for (int k = 0; k < 2; ++k) mean[k] /= (double)nv;
        const double vsn[2] = {vs0, vs1};
        for (int n = 0; n < 2; ++n) {
          hipLaunchKernelGGL(k_divergence2d, dim3(c.blocks), dim3(256), 0, c.s,
                             c.g.get(), n, Nx_v, Ny_v, c.sbuf.get());
          HIP_LAUNCH_OK();
          c.dpass(2, c.sbuf.get(), c.tbuf.get(), TEPI_STORE, 1, 0);
          c.dpass(0, c.tbuf.get(), c.ubuf.get(), TEPI_INVLAMBDA, 1, 0);
          c.dpass(0, c.ubuf.get(), c.tbuf.get(), TEPI_STORE, 1, 0);
          c.dpass(2, c.tbuf.get(), c.w.get() + n * nv, TEPI_FINAL, vsn[n], mean[n]);
        }
        corner(lo, hi);
        inner++;
      }
      inner_total += inner;
      jstats(jmin, jmax, nb, na);
      outer++;
    }
  }
  hipLaunchKernelGGL(k_scale_planes, dim3(c.blocks), dim3(256), 0, c.s, c.w.get(), nv,
                     vs0, vs1, 1.0, 1);
  HIP_LAUNCH_OK();
  if (stats) {
    stats[0] = outer;
    stats[1] = inner_total;
    stats[2] = jmin;
    stats[3] = jmax;
    stats[4] = (double)nb;
    stats[5] = (double)na;
  }
}

__global__ void k_planes_f32_to_f64(const float *in, double *out, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (double)in[i];
}

__global__ void k_planes_f64_to_f32(const double *in, float *out, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (float)in[i];
}

}  // namespace

void topology_on_device(float *d_soa, int N, int Nx, int Ny, int Nz, const double *vs,
                        double e1, double e2, int max_outer, int max_inner, hipStream_t s,
                        double *stats) {
  TopoCtx c;
  if (N == 3) {
    BINDER_REQUIRE(Nx >= 2 && Ny >= 2 && Nz >= 2,
                   "topology correction needs at least 2 voxels per axis");
    topo_init(c, s, Nx, Ny, Nz, vs, e1, e2);
  } else {
    BINDER_REQUIRE(Nx >= 2 && Ny >= 2, "topology correction needs at least 2 voxels per axis");
    const double vs3[3] = {vs[0], 1.0, vs[1]};
    topo_init(c, s, Nx, 1, Ny, vs3, e1, e2);
  }
  const long long nv = c.nv;
  c.w.alloc((size_t)3 * nv);
  if (N == 2) HIP_OK(hipMemsetAsync(c.w.get() + 2 * nv, 0, sizeof(double) * nv, s));
  const int gb = (int)std::min<long long>((N * nv + 255) / 256, 8192);
  hipLaunchKernelGGL(k_planes_f32_to_f64, dim3(gb), dim3(256), 0, s, d_soa, c.w.get(),
                     (long long)N * nv);
  HIP_LAUNCH_OK();
  if (N == 3) topo3_run(c, max_outer, max_inner, stats);
  else topo2_run(c, Nx, Ny, vs[0], vs[1], e1, e2, max_outer, max_inner, stats);
  hipLaunchKernelGGL(k_planes_f64_to_f32, dim3(gb), dim3(256), 0, s, c.w.get(), d_soa,
                     (long long)N * nv);
  HIP_LAUNCH_OK();
  HIP_OK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Allocation-free fold mask (binder_fold_mask_3d_noalloc).  Its only scratch
// is k_jacobian3d's per-block partials (min, max, #(J < e1), #(J > e2)) for at
// most kFoldBlocks blocks, held in a file-scope device array.  Host code
// never passes the array's name to a kernel: it is a host-side stand-in, so
// the entry point looks up each device's own address once with
// hipGetSymbolAddress and uses that.
// ---------------------------------------------------------------------------
constexpr int kFoldBlocks = 4096;
constexpr int kMaxScratchDevices = 64;
constexpr long long kFoldNoallocMaxVoxels = 2147483647LL;  // 2^31 - 1

}  // namespace binder

static __device__ __align__(256) double scratch[binder::kFoldBlocks * 4];

namespace binder {
namespace {

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call made with that device current.
double *fold_scratch() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  BINDER_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BINDER_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  return g_scratch_ptr[dev];
}

// Fixed-order sum of the blocks' fold counts (integers held in FP64: exact).
__global__ __launch_bounds__(256) void k_fold_count(const double *part, int blocks,
                                                    long long *count) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < blocks; b += 256) s += part[b * 4 + 2];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) *count = (long long)red[0];
}

}  // namespace
}  // namespace binder

using namespace binder;

extern "C" {

int binder_fold_mask_3d_noalloc(const double *field_planes, int Nx, int Ny, int Nz,
                                double vx, double vy, double vz, double e1,
                                unsigned char *mask, long long *count, void *stream) {
  return guarded([&] {
    BINDER_REQUIRE(field_planes && count, "null argument");
    BINDER_REQUIRE(Nx > 0 && Ny > 0 && Nz > 0, "Dimensions must be positive");
    const long long nv = (long long)Nx * Ny * Nz;
    BINDER_REQUIRE(nv <= kFoldNoallocMaxVoxels,
                   "binder_fold_mask_3d_noalloc: at most 2^31 - 1 voxels");
    double *part = fold_scratch();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int blocks = (int)std::min<long long>((nv + 255) / 256, kFoldBlocks);
    hipLaunchKernelGGL(k_jacobian3d, dim3(blocks), dim3(256), 0, s,
                       Field{field_planes, field_planes + nv, field_planes + 2 * nv}, Nx,
                       Ny, Nz, 1.0 / (vx * vy * vz), e1, INFINITY, (double *)nullptr,
                       (uint8_t *)mask, part);
    HIP_LAUNCH_OK();
    hipLaunchKernelGGL(k_fold_count, dim3(1), dim3(256), 0, s, (const double *)part, blocks,
                       count);
    HIP_LAUNCH_OK();
  });
}

int binder_jacobian_determinant_3d(const double *field, int Nx, int Ny, int Nz,
                                   double vx, double vy, double vz,
                                   double *out) {
  return guarded([&] {
    BINDER_REQUIRE(field && out, "Field pointer cannot be null");
    BINDER_REQUIRE(Nx > 0 && Ny > 0 && Nz > 0, "Dimensions must be positive");
    OwnedStream st;
    const long long nv = (long long)Nx * Ny * Nz;
    DBuf<double> w, J, part;
    upload_soa3(field, nv, w, st.s);
    J.alloc(nv);
    const int blocks = (int)std::min<long long>((nv + 255) / 256, 4096);
    part.alloc((size_t)blocks * 4);
    const double vol_scale = 1.0 / (vx * vy * vz);
    hipLaunchKernelGGL(k_jacobian3d, dim3(blocks), dim3(256), 0, st.s,
                       Field{w.get(), w.get() + nv, w.get() + 2 * nv}, Nx, Ny,
                       Nz, vol_scale, 0.0, 0.0, J.get(), (uint8_t *)nullptr,
                       part.get());
    HIP_LAUNCH_OK();
    J.download(out, nv, st.s);
    HIP_OK(hipStreamSynchronize(st.s));
  });
}

int binder_jacobian_determinant_2d(const double *field, int Nx, int Ny,
                                   double vx, double vy, double *out) {
  return guarded([&] {
    BINDER_REQUIRE(field && out, "Field pointer cannot be null");
    BINDER_REQUIRE(Nx > 0 && Ny > 0, "Dimensions must be positive");
    OwnedStream st;
    const long long nv = (long long)Nx * Ny;
    std::vector<double> h((size_t)2 * nv);
    for (long long i = 0; i < nv; ++i) {
      h[i] = field[2 * i];
      h[nv + i] = field[2 * i + 1];
    }
    DBuf<double> w, J;
    w.upload(h.data(), h.size(), st.s);
    J.alloc(nv);
    const double vol_scale = 1.0 / (vx * vy);
    hipLaunchKernelGGL(k_jacobian2d, dim3((unsigned)((nv + 255) / 256)),
                       dim3(256), 0, st.s, w.get(), w.get() + nv, Nx, Ny,
                       vol_scale, J.get());
    HIP_LAUNCH_OK();
    J.download(out, nv, st.s);
    HIP_OK(hipStreamSynchronize(st.s));
  });
}

int binder_fold_mask_3d(const double *field, int Nx, int Ny, int Nz, double vx,
                        double vy, double vz, double e1, unsigned char *mask,
                        long long *count) {
  return guarded([&] {
    BINDER_REQUIRE(field && count, "null argument");
    BINDER_REQUIRE(Nx > 0 && Ny > 0 && Nz > 0, "Dimensions must be positive");
    OwnedStream st;
    const long long nv = (long long)Nx * Ny * Nz;
    DBuf<double> w, part;
    DBuf<uint8_t> m;
    upload_soa3(field, nv, w, st.s);
    m.alloc(nv);
    const int blocks = (int)std::min<long long>((nv + 255) / 256, 4096);
    part.alloc((size_t)blocks * 4);
    hipLaunchKernelGGL(k_jacobian3d, dim3(blocks), dim3(256), 0, st.s,
                       Field{w.get(), w.get() + nv, w.get() + 2 * nv}, Nx, Ny,
                       Nz, 1.0 / (vx * vy * vz), e1, INFINITY, (double *)nullptr,
                       m.get(), part.get());
    HIP_LAUNCH_OK();
    std::vector<double> hp((size_t)blocks * 4);
    part.download(hp.data(), hp.size(), st.s);
    if (mask) m.download(mask, nv, st.s);
    HIP_OK(hipStreamSynchronize(st.s));
    double lo = 0;
    for (int b = 0; b < blocks; ++b) lo += hp[b * 4 + 2];
    *count = (long long)lo;
  });
}

int binder_topology_correction_3d(const double *h_init, double e1, double e2,
                                  int max_outer_it, int max_it, int threads,
                                  const double *voxel_sizes, int Nx_v, int Ny_v,
                                  int Nz_v, double *out, double *stats) {
  (void)threads;
  return guarded([&] {
    BINDER_REQUIRE(h_init && voxel_sizes && out, "null argument");
    BINDER_REQUIRE(Nx_v >= 2 && Ny_v >= 2 && Nz_v >= 2,
                   "topology correction needs at least 2 voxels per axis");
    OwnedStream st;
    TopoCtx c;
    topo_init(c, st.s, Nx_v, Ny_v, Nz_v, voxel_sizes, e1, e2);
    upload_soa3(h_init, c.nv, c.w, c.s);
    topo3_run(c, max_outer_it, max_it, stats);
    std::vector<double> h((size_t)3 * c.nv);
    c.w.download(h.data(), h.size(), c.s);
    HIP_OK(hipStreamSynchronize(c.s));
    for (long long i = 0; i < c.nv; ++i)
      for (int k = 0; k < 3; ++k) out[i * 3 + k] = h[k * c.nv + i];
  });
}

int binder_topology_correction_2d(const double *h_init, double e1, double e2,
                                  int max_outer_it, int max_it, int threads,
                                  const double *voxel_sizes, int Nx_v, int Ny_v,
                                  double *out, double *stats) {
  (void)threads;
  return guarded([&] {
    BINDER_REQUIRE(h_init && voxel_sizes && out, "null argument");
    BINDER_REQUIRE(Nx_v >= 2 && Ny_v >= 2,
                   "topology correction needs at least 2 voxels per axis");
    OwnedStream st;
    TopoCtx c;
    const double vs3[3] = {voxel_sizes[0], 1.0, voxel_sizes[1]};
    topo_init(c, st.s, Nx_v, 1, Ny_v, vs3, e1, e2);
    const long long nv = c.nv;
    {
      std::vector<double> h((size_t)3 * nv, 0.0);
      for (long long i = 0; i < nv; ++i) {
        h[i] = h_init[2 * i];
        h[nv + i] = h_init[2 * i + 1];
      }
      c.w.upload(h.data(), h.size(), c.s);
      HIP_OK(hipStreamSynchronize(c.s));
    }
    topo2_run(c, Nx_v, Ny_v, voxel_sizes[0], voxel_sizes[1], e1, e2, max_outer_it, max_it,
              stats);
    std::vector<double> h((size_t)2 * nv);
    c.w.download(h.data(), h.size(), c.s);
    HIP_OK(hipStreamSynchronize(c.s));
    for (long long i = 0; i < nv; ++i) {
      out[2 * i] = h[i];
      out[2 * i + 1] = h[nv + i];
    }
  });
}

}  // extern "C"
