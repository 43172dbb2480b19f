// ORACLE -- test infrastructure, not product code.
//
// FP64 CPU restatement of BINDER's Jacobian determinant
// (BINDER/src/cpu/utils.cpp:10-345) and the 3D and 2D topology corrections
// (BINDER/src/cpu/topology_correction_3D.cpp:17-607,
// BINDER/src/cpu/topology_correction_2D.cpp:17-430).  The expression trees
// and corner orders follow the reference so that, compiled without FMA
// contraction, the Jacobian agrees bit for bit.  The correction follows the
// reference's loop order, including the sequential in-place gradient
// limiting, with the periodic Poisson integration done by this oracle's own
// mixed-radix FFT (the reference uses FFTW).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "fft.h"
#include "oracle.h"

namespace orc {

// 3x3 determinant by first-row cofactor expansion, the operand order used by
// every corner in utils.cpp:156-334 and topology_correction_3D.cpp:286-307,
// :538-569:  a*(e*i - f*h) - b*(d*i - f*g) + c*(d*h - e*g)
// with rows (a,b,c) = x component, (d,e,f) = y, (g,h,i) = z and columns the
// differences along x, y, z.
static inline double det3(double a, double b, double c, double d, double e,
                          double f, double g, double h, double i) {
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// compute_jacobian_determinant_3D (utils.cpp:100-345).  Corner order fff,
// bff, fbf, ffb, bbf, bfb, fbb, bbb; 'f' = w(p+e) - w(p), 'b' = w(p) - w(p-e).
static void jacobian3d(const double *w, int Nx, int Ny, int Nz, double vx,
                       double vy, double vz, double *J) {
  const double vol_scale = 1.0 / (vx * vy * vz);
  static const int corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                    {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  auto at = [&](int x, int y, int z) -> const double * {
    return w + (((long long)x * Ny + y) * Nz + z) * 3;
  };
#pragma omp parallel for schedule(static)
  for (int x = 0; x < Nx; ++x)
    for (int y = 0; y < Ny; ++y)
      for (int z = 0; z < Nz; ++z) {
        const double *p = at(x, y, z);
        double total = 0.0;
        int valid = 0;
        for (int c = 0; c < 8; ++c) {
          const int bx = corners[c][0], by = corners[c][1], bz = corners[c][2];
          if (bx ? x == 0 : x >= Nx - 1) continue;
          if (by ? y == 0 : y >= Ny - 1) continue;
          if (bz ? z == 0 : z >= Nz - 1) continue;
          double dx[3], dy[3], dz[3];  // difference of component k along axis
          const double *q;
          q = bx ? at(x - 1, y, z) : at(x + 1, y, z);
          for (int k = 0; k < 3; ++k) dx[k] = bx ? p[k] - q[k] : q[k] - p[k];
          q = by ? at(x, y - 1, z) : at(x, y + 1, z);
          for (int k = 0; k < 3; ++k) dy[k] = by ? p[k] - q[k] : q[k] - p[k];
          q = bz ? at(x, y, z - 1) : at(x, y, z + 1);
          for (int k = 0; k < 3; ++k) dz[k] = bz ? p[k] - q[k] : q[k] - p[k];
          const double Jc = det3(dx[0], dy[0], dz[0], dx[1], dy[1], dz[1], dx[2],
                                 dy[2], dz[2]) * vol_scale;
          total += Jc;
          valid++;
        }
        J[((long long)x * Ny + y) * Nz + z] = valid > 0 ? total / valid : 1.0;
      }
}

// compute_jacobian_determinant_2D (utils.cpp:10-98): corners ff, bf, fb, bb.
static void jacobian2d(const double *w, int Nx, int Ny, double vx, double vy,
                       double *J) {
  const double vol_scale = 1.0 / (vx * vy);
  static const int corners[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
  for (int x = 0; x < Nx; ++x)
    for (int y = 0; y < Ny; ++y) {
      const double *p = w + ((long long)x * Ny + y) * 2;
      double total = 0.0;
      int valid = 0;
      for (int c = 0; c < 4; ++c) {
        const int bx = corners[c][0], by = corners[c][1];
        if (bx ? x == 0 : x >= Nx - 1) continue;
        if (by ? y == 0 : y >= Ny - 1) continue;
        const double *qx = w + ((long long)(bx ? x - 1 : x + 1) * Ny + y) * 2;
        const double *qy = w + ((long long)x * Ny + (by ? y - 1 : y + 1)) * 2;
        const double X1 = bx ? p[0] - qx[0] : qx[0] - p[0];
        const double Y1 = bx ? p[1] - qx[1] : qx[1] - p[1];
        const double X2 = by ? p[0] - qy[0] : qy[0] - p[0];
        const double Y2 = by ? p[1] - qy[1] : qy[1] - p[1];
        total += (X1 * Y2 - X2 * Y1) * vol_scale;
        valid++;
      }
      J[(long long)x * Ny + y] = valid > 0 ? total / valid : 1.0;
    }
}

// ---------------------------------------------------------------------------
// topology_correction_3D (topology_correction_3D.cpp:17-607)
// ---------------------------------------------------------------------------
struct Topo {
  int Nx, Ny, Nz;
  long long nv;
  double e1, e2, vs[3];
  std::vector<double> w;  // physical positions [nv,3]

  // vertex of each of the 8 corner Jacobians of a cell, in J_corner order
  // fff, bff, fbf, ffb, fbb, bfb, bbf, bbb (topology_correction_3D.cpp:573-580,
  // offsets get_jac_off_3D :452-486).
  static constexpr int V[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

  const double *at(int x, int y, int z) const {
    return &w[(((long long)x * Ny + y) * Nz + z) * 3];
  }

  // compute_corner_jacobians_3D (:488-607): each corner is the determinant of
  // the three cell-edge differences meeting at its vertex.
  void corner_jacobians(std::vector<double> &Jc, long long &nlow,
                        long long &nhigh) const {
    const double vol_scale = 1.0 / (vs[0] * vs[1] * vs[2]);
    nlow = nhigh = 0;
    const long long cells = (long long)(Nx - 1) * (Ny - 1) * (Nz - 1);
    Jc.resize(cells * 8);
    for (int x = 0; x < Nx - 1; ++x)
      for (int y = 0; y < Ny - 1; ++y)
        for (int z = 0; z < Nz - 1; ++z) {
          const long long cell = ((long long)x * (Ny - 1) + y) * (Nz - 1) + z;
          for (int c = 0; c < 8; ++c) {
            const int cx = V[c][0], cy = V[c][1], cz = V[c][2];
            const double *ex1 = at(x + 1, y + cy, z + cz), *ex0 = at(x, y + cy, z + cz);
            const double *ey1 = at(x + cx, y + 1, z + cz), *ey0 = at(x + cx, y, z + cz);
            const double *ez1 = at(x + cx, y + cy, z + 1), *ez0 = at(x + cx, y + cy, z);
            double J = det3(ex1[0] - ex0[0], ey1[0] - ey0[0], ez1[0] - ez0[0],
                            ex1[1] - ex0[1], ey1[1] - ey0[1], ez1[1] - ez0[1],
                            ex1[2] - ex0[2], ey1[2] - ey0[2], ez1[2] - ez0[2]);
            J *= vol_scale;
            Jc[cell * 8 + c] = J;
            if (J < e1) nlow++;
            if (J > e2) nhigh++;
          }
        }
  }

  // compute_gradients_3D (:172-245): periodic forward differences of
  // component n along axis a, divided by the component's voxel size;
  // grad[(voxel*3 + n)*3 + a].
  void gradients(std::vector<double> &g) const {
    g.resize(nv * 9);
    for (int x = 0; x < Nx; ++x)
      for (int y = 0; y < Ny; ++y)
        for (int z = 0; z < Nz; ++z) {
          const int x2 = x + 1 > Nx - 1 ? x + 1 - Nx : x + 1;
          const int y2 = y + 1 > Ny - 1 ? y + 1 - Ny : y + 1;
          const int z2 = z + 1 > Nz - 1 ? z + 1 - Nz : z + 1;
          const double *p = at(x, y, z);
          const double *nb[3] = {at(x2, y, z), at(x, y2, z), at(x, y, z2)};
          double *o = &g[(((long long)x * Ny + y) * Nz + z) * 9];
          for (int n = 0; n < 3; ++n)
            for (int a = 0; a < 3; ++a) o[n * 3 + a] = (nb[a][n] - p[n]) / vs[n];
        }
  }

  // limit_gradients_3D (:247-332): sequential, in place, x->y->z->corner.
  // The blend loop stops at alpha = 1 (the reference would loop forever if
  // the identity itself were out of [e1,e2]).
  void limit(std::vector<double> &g, const std::vector<double> &Jc) const {
    const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int x = 0; x < Nx - 1; ++x)
      for (int y = 0; y < Ny - 1; ++y)
        for (int z = 0; z < Nz - 1; ++z) {
          const long long cell = ((long long)x * (Ny - 1) + y) * (Nz - 1) + z;
          for (int c = 0; c < 8; ++c) {
            const double jv = Jc[cell * 8 + c];
            if (!(jv < e1 || jv > e2)) continue;
            long long loc[3];
            for (int a = 0; a < 3; ++a) {
              int px = x + V[c][0], py = y + V[c][1], pz = z + V[c][2];
              if (a == 0) px = x;
              if (a == 1) py = y;
              if (a == 2) pz = z;
              loc[a] = ((long long)px * Ny + py) * Nz + pz;
            }
            double J[9], Jn[9];
            for (int n1 = 0; n1 < 3; ++n1)
              for (int n2 = 0; n2 < 3; ++n2) J[n1 * 3 + n2] = g[loc[n2] * 9 + n1 * 3 + n2];
            double alpha = 0.0;
            double det = det3(J[0], J[1], J[2], J[3], J[4], J[5], J[6], J[7], J[8]);
            int guard = 0;
            while ((det > e2 || det < e1) && guard < 64) {
              alpha += 0.1;
              if (alpha > 1.0) alpha = 1.0;
              for (int k = 0; k < 9; ++k) Jn[k] = (1 - alpha) * J[k] + alpha * I[k];
              det = det3(Jn[0], Jn[1], Jn[2], Jn[3], Jn[4], Jn[5], Jn[6], Jn[7], Jn[8]);
              if (alpha >= 1.0) ++guard;
            }
            alpha += 0.1;
            if (alpha > 1.0) alpha = 1.0;
            for (int n1 = 0; n1 < 3; ++n1)
              for (int n2 = 0; n2 < 3; ++n2)
                g[loc[n2] * 9 + n1 * 3 + n2] =
                    (1 - alpha) * J[n1 * 3 + n2] + alpha * I[n1 * 3 + n2];
          }
        }
  }

  // compute_mean_coordinates (:334-358)
  void mean(double m[3]) const {
    double s[3] = {0, 0, 0};
    for (long long i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) s[k] += w[i * 3 + k];
    for (int k = 0; k < 3; ++k) m[k] = s[k] / (double)(Nx * Ny * Nz);
  }

  // In-place 3D complex DFT of a [Nx,Ny,Nz] array (sign -1 forward).
  void fft3(std::vector<cd> &a, int sign) const {
    const int dims[3] = {Nx, Ny, Nz};
    const long long st[3] = {(long long)Ny * Nz, Nz, 1};
    for (int axis = 0; axis < 3; ++axis) {
      const int n = dims[axis];
      if (n == 1) continue;
      const long long lines = nv / n;
#pragma omp parallel
      {
        FFT1D f(n);
        std::vector<cd> in(n), out(n), work(n + 8);
#pragma omp for schedule(static)
        for (long long l = 0; l < lines; ++l) {
          long long base;
          if (axis == 2) base = l * Nz;
          else if (axis == 1) base = (l / Nz) * Ny * Nz + (l % Nz);
          else base = l;
          for (int j = 0; j < n; ++j) in[j] = a[base + j * st[axis]];
          f.run(in.data(), out.data(), sign, work.data());
          for (int j = 0; j < n; ++j) a[base + j * st[axis]] = out[j];
        }
      }
    }
  }

  // integrate_gradient_field_3D (:361-449): periodic least-squares
  // integration conj(e^{iw}-1).g_hat / sum(2 - 2 cos w), DC set to zero.
  void integrate(const std::vector<double> &g, const double m[3]) {
    std::vector<cd> fx(nv), fy(nv), fz(nv), out(nv);
    for (int n = 0; n < 3; ++n) {
      for (long long i = 0; i < nv; ++i) {
        fx[i] = cd(g[i * 9 + n * 3 + 0], 0.0);
        fy[i] = cd(g[i * 9 + n * 3 + 1], 0.0);
        fz[i] = cd(g[i * 9 + n * 3 + 2], 0.0);
      }
      fft3(fx, -1);
      fft3(fy, -1);
      fft3(fz, -1);
      for (int z = 0; z < Nz; ++z) {
        const double cz = std::cos(2.0 * M_PI * z / (Nz * 1.0));
        const double sz = std::sin(2.0 * M_PI * z / (Nz * 1.0));
        for (int y = 0; y < Ny; ++y) {
          const double cy = std::cos(2.0 * M_PI * y / (Ny * 1.0));
          const double sy = std::sin(2.0 * M_PI * y / (Ny * 1.0));
          for (int x = 0; x < Nx; ++x) {
            const double cx = std::cos(2.0 * M_PI * x / (Nx * 1.0));
            const double sx = std::sin(2.0 * M_PI * x / (Nx * 1.0));
            const long long i = ((long long)x * Ny + y) * Nz + z;
            const double norm = 6.0 - 2.0 * cz - 2.0 * cx - 2.0 * cy;
            if (std::fabs(norm) > 1e-10) {
              const double re = fx[i].real() * (cx - 1) + fx[i].imag() * sx +
                                fy[i].real() * (cy - 1) + fy[i].imag() * sy +
                                fz[i].real() * (cz - 1) + fz[i].imag() * sz;
              const double im = fx[i].imag() * (cx - 1) - fx[i].real() * sx +
                                fy[i].imag() * (cy - 1) - fy[i].real() * sy +
                                fz[i].imag() * (cz - 1) - fz[i].real() * sz;
              out[i] = cd(re / norm, im / norm);
            } else {
              out[i] = cd(0.0, 0.0);
            }
          }
        }
      }
      fft3(out, +1);
      for (long long i = 0; i < nv; ++i) w[i * 3 + n] = out[i].real() / (Nx * Ny * Nz * 1.0);
    }
    for (long long i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) {
        w[i * 3 + k] *= vs[k];
        w[i * 3 + k] += m[k];
      }
  }

  // correct_topology (:17-143).  stats = {outer, inner total, Jmin, Jmax,
  // #below e1, #above e2}.
  void run(int max_outer, int max_inner, double *stats) {
    for (long long i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) w[i * 3 + k] *= vs[k];  // to_physical
    std::vector<double> J(nv);
    auto jstats = [&](double &jmin, double &jmax, long long &nb, long long &na) {
      jacobian3d(w.data(), Nx, Ny, Nz, vs[0], vs[1], vs[2], J.data());
      jmin = *std::min_element(J.begin(), J.end());
      jmax = *std::max_element(J.begin(), J.end());
      nb = std::count_if(J.begin(), J.end(), [&](double v) { return v < e1; });
      na = std::count_if(J.begin(), J.end(), [&](double v) { return v > e2; });
    };
    double jmin, jmax;
    long long nb, na;
    jstats(jmin, jmax, nb, na);
    int outer = 0, inner_total = 0;
    if (!(jmin >= e1 && jmax <= e2)) {
      std::vector<double> Jc, g;
      while ((jmin < e1 || jmax > e2) && outer < max_outer) {
        long long lo, hi;
        corner_jacobians(Jc, lo, hi);
        int inner = 0;
        while ((lo > 0 || hi > 0) && inner < max_inner) {
          double m[3];
          gradients(g);
          limit(g, Jc);
          mean(m);
          integrate(g, m);
          corner_jacobians(Jc, lo, hi);
          inner++;
        }
        inner_total += inner;
        jstats(jmin, jmax, nb, na);
        outer++;
      }
    }
    for (long long i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) w[i * 3 + k] /= vs[k];  // to_grid
    if (stats) {
      stats[0] = outer;
      stats[1] = inner_total;
      stats[2] = jmin;
      stats[3] = jmax;
      stats[4] = (double)nb;
      stats[5] = (double)na;
    }
  }
};

constexpr int Topo::V[8][3];

// ---------------------------------------------------------------------------
// topology_correction_2D (topology_correction_2D.cpp:17-430)
// ---------------------------------------------------------------------------
struct Topo2 {
  int Nx, Ny;
  long long nv;
  double e1, e2, vs[2];
  std::vector<double> w;  // physical positions [nv,2]

  // vertex offsets of the two difference columns of each corner, in the
  // J_corner order ff, bf, fb, bb (get_jac_off_2D, :394-410): column a
  // (difference along axis a) is read at (x + xoff[a], y + yoff[a])
  static constexpr int XOFF[4][2] = {{0, 0}, {0, 1}, {0, 0}, {0, 1}};
  static constexpr int YOFF[4][2] = {{0, 0}, {0, 0}, {1, 0}, {1, 0}};

  const double *at(int x, int y) const { return &w[((long long)x * Ny + y) * 2]; }

  // compute_corner_jacobians_2D (:156-200): Jff, Jbf, Jfb, Jbb per cell;
  // counts are of cells with any corner below e1 / above e2.
  void corner_jacobians(std::vector<double> &Jc, long long &nlow,
                        long long &nhigh) const {
    const double vol_scale = 1.0 / (vs[0] * vs[1]);
    nlow = nhigh = 0;
    Jc.resize((size_t)(Nx - 1) * (Ny - 1) * 4);
    for (int x = 0; x < Nx - 1; ++x)
      for (int y = 0; y < Ny - 1; ++y) {
        const double *p00 = at(x, y), *p01 = at(x, y + 1), *p10 = at(x + 1, y),
                     *p11 = at(x + 1, y + 1);
        const double wx00 = p00[0], wx01 = p01[0], wx10 = p10[0], wx11 = p11[0];
        const double wy00 = p00[1], wy01 = p01[1], wy10 = p10[1], wy11 = p11[1];
        const double J[4] = {
            ((wx10 - wx00) * (wy01 - wy00) - (wx01 - wx00) * (wy10 - wy00)) * vol_scale,
            ((wx10 - wx00) * (wy11 - wy10) - (wx11 - wx10) * (wy10 - wy00)) * vol_scale,
            ((wx11 - wx01) * (wy01 - wy00) - (wx01 - wx00) * (wy11 - wy01)) * vol_scale,
            ((wx11 - wx01) * (wy11 - wy10) - (wx11 - wx10) * (wy11 - wy01)) * vol_scale};
        const long long cell = (long long)x * (Ny - 1) + y;
        bool lo = false, hi = false;
        for (int c = 0; c < 4; ++c) {
          Jc[cell * 4 + c] = J[c];
          lo = lo || J[c] < e1;
          hi = hi || J[c] > e2;
        }
        nlow += lo;
        nhigh += hi;
      }
  }

  // compute_gradients_2D (:202-245): periodic forward differences,
  // g[(voxel*2 + n)*2 + a] = (w_n(p + e_a) - w_n(p)) / vs_n
  void gradients(std::vector<double> &g) const {
    g.resize((size_t)nv * 4);
    for (int x = 0; x < Nx; ++x)
      for (int y = 0; y < Ny; ++y) {
        const int x2 = x + 1 > Nx - 1 ? x + 1 - Nx : x + 1;
        const int y2 = y + 1 > Ny - 1 ? y + 1 - Ny : y + 1;
        const double *p = at(x, y), *px = at(x2, y), *py = at(x, y2);
        double *o = &g[((long long)x * Ny + y) * 4];
        for (int n = 0; n < 2; ++n) {
          o[n * 2 + 0] = (px[n] - p[n]) / vs[n];
          o[n * 2 + 1] = (py[n] - p[n]) / vs[n];
        }
      }
  }

  // limit_gradients_2D (:247-318): sequential, in place, x -> y -> corner.
  // The blend loop stops at alpha = 1 (the reference would loop forever if
  // the identity itself were out of [e1, e2]).
  void limit(std::vector<double> &g, const std::vector<double> &Jc) const {
    for (int x = 0; x < Nx - 1; ++x)
      for (int y = 0; y < Ny - 1; ++y) {
        const long long cell = (long long)x * (Ny - 1) + y;
        for (int c = 0; c < 4; ++c) {
          const double jv = Jc[cell * 4 + c];
          if (!(jv < e1 || jv > e2)) continue;
          long long loc[2];
          for (int a = 0; a < 2; ++a)
            loc[a] = (long long)(x + XOFF[c][a]) * Ny + (y + YOFF[c][a]);
          double J[4];
          for (int n1 = 0; n1 < 2; ++n1)
            for (int n2 = 0; n2 < 2; ++n2) J[n1 * 2 + n2] = g[loc[n2] * 4 + n1 * 2 + n2];
          const double I[4] = {1, 0, 0, 1};
          double alpha = 0.0;
          double det = J[0] * J[3] - J[1] * J[2];
          int guard = 0;
          while ((det > e2 || det < e1) && guard < 64) {
            alpha += 0.1;
            if (alpha > 1.0) alpha = 1.0;
            double Jn[4];
            for (int k = 0; k < 4; ++k) Jn[k] = (1 - alpha) * J[k] + alpha * I[k];
            det = Jn[0] * Jn[3] - Jn[1] * Jn[2];
            if (alpha >= 1.0) ++guard;
          }
          alpha += 0.1;
          if (alpha > 1.0) alpha = 1.0;
          for (int n1 = 0; n1 < 2; ++n1)
            for (int n2 = 0; n2 < 2; ++n2)
              g[loc[n2] * 4 + n1 * 2 + n2] =
                  (1 - alpha) * J[n1 * 2 + n2] + alpha * I[n1 * 2 + n2];
        }
      }
  }

  // compute_mean_coordinates (:412-430)
  void mean(double m[2]) const {
    double s0 = 0.0, s1 = 0.0;
    for (long long i = 0; i < nv; ++i) {
      s0 += w[i * 2 + 0];
      s1 += w[i * 2 + 1];
    }
    m[0] = s0 / (Nx * Ny);
    m[1] = s1 / (Nx * Ny);
  }

  void fft2(std::vector<cd> &a, int sign) const {
    const int dims[2] = {Nx, Ny};
    for (int axis = 0; axis < 2; ++axis) {
      const int n = dims[axis];
      if (n == 1) continue;
      const long long lines = nv / n;
      const long long st = axis == 0 ? Ny : 1;
      FFT1D f(n);
      std::vector<cd> in(n), out(n), work(n + 8);
      for (long long l = 0; l < lines; ++l) {
        const long long base = axis == 1 ? l * Ny : l;
        for (int j = 0; j < n; ++j) in[j] = a[base + j * st];
        f.run(in.data(), out.data(), sign, work.data());
        for (int j = 0; j < n; ++j) a[base + j * st] = out[j];
      }
    }
  }

  // integrate_gradient_field_2D (:320-392)
  void integrate(const std::vector<double> &g, const double m[2]) {
    std::vector<cd> fx(nv), fy(nv), out(nv);
    for (int n = 0; n < 2; ++n) {
      for (long long i = 0; i < nv; ++i) {
        fx[i] = cd(g[i * 4 + n * 2 + 0], 0.0);
        fy[i] = cd(g[i * 4 + n * 2 + 1], 0.0);
      }
      fft2(fx, -1);
      fft2(fy, -1);
      for (int x = 0; x < Nx; ++x)
        for (int y = 0; y < Ny; ++y) {
          const long long i = (long long)x * Ny + y;
          const double cy = std::cos(2.0 * M_PI * y / Ny), sy = std::sin(2.0 * M_PI * y / Ny);
          const double cx = std::cos(2.0 * M_PI * x / Nx), sx = std::sin(2.0 * M_PI * x / Nx);
          const double norm = 4.0 - 2.0 * cy - 2.0 * cx;
          if (norm > 1e-10) {
            const double re = (fx[i].real() * (cx - 1.0) + fx[i].imag() * sx +
                               fy[i].real() * (cy - 1.0) + fy[i].imag() * sy) / norm;
            const double im = (fx[i].imag() * (cx - 1.0) - fx[i].real() * sx +
                               fy[i].imag() * (cy - 1.0) - fy[i].real() * sy) / norm;
            out[i] = cd(re, im);
          } else {
            out[i] = cd(0.0, 0.0);
          }
        }
      fft2(out, +1);
      for (long long i = 0; i < nv; ++i) {
        double v = out[i].real() * (1.0 / (Nx * Ny));
        v *= vs[n];
        v += m[n];
        w[i * 2 + n] = v;
      }
    }
  }

  void run(int max_outer, int max_inner, double *stats) {
    for (long long i = 0; i < nv; ++i)
      for (int k = 0; k < 2; ++k) w[i * 2 + k] *= vs[k];
    std::vector<double> J(nv);
    auto jstats = [&](double &jmin, double &jmax, long long &nb, long long &na) {
      jacobian2d(w.data(), Nx, Ny, vs[0], vs[1], J.data());
      jmin = *std::min_element(J.begin(), J.end());
      jmax = *std::max_element(J.begin(), J.end());
      nb = std::count_if(J.begin(), J.end(), [&](double v) { return v < e1; });
      na = std::count_if(J.begin(), J.end(), [&](double v) { return v > e2; });
    };
    double jmin, jmax;
    long long nb, na;
    jstats(jmin, jmax, nb, na);
    int outer = 0, inner_total = 0;
    if (!(jmin >= e1 && jmax <= e2)) {
      std::vector<double> Jc, g;
      while ((jmin < e1 || jmax > e2) && outer < max_outer) {
        long long lo, hi;
        corner_jacobians(Jc, lo, hi);
        int inner = 0;
        while ((lo > 0 || hi > 0) && inner < max_inner) {
          double m[2];
          gradients(g);
          limit(g, Jc);
          mean(m);
          integrate(g, m);
          corner_jacobians(Jc, lo, hi);
          inner++;
        }
        inner_total += inner;
        jstats(jmin, jmax, nb, na);
        outer++;
      }
    }
    for (long long i = 0; i < nv; ++i)
      for (int k = 0; k < 2; ++k) w[i * 2 + k] /= vs[k];
    if (stats) {
      stats[0] = outer;
      stats[1] = inner_total;
      stats[2] = jmin;
      stats[3] = jmax;
      stats[4] = (double)nb;
      stats[5] = (double)na;
    }
  }
};

constexpr int Topo2::XOFF[4][2];
constexpr int Topo2::YOFF[4][2];

}  // namespace orc

using namespace orc;

extern "C" {

void orc_jacobian_3d(const double *field, int Nx, int Ny, int Nz, double vx,
                     double vy, double vz, double *out) {
  jacobian3d(field, Nx, Ny, Nz, vx, vy, vz, out);
}

void orc_jacobian_2d(const double *field, int Nx, int Ny, double vx, double vy,
                     double *out) {
  jacobian2d(field, Nx, Ny, vx, vy, out);
}

void orc_topology_3d(const double *h_init, double e1, double e2, int max_outer,
                     int max_inner, const double *vs, int Nx, int Ny, int Nz,
                     double *out, double *stats) {
  Topo t;
  t.Nx = Nx;
  t.Ny = Ny;
  t.Nz = Nz;
  t.nv = (long long)Nx * Ny * Nz;
  t.e1 = e1;
  t.e2 = e2;
  for (int k = 0; k < 3; ++k) t.vs[k] = vs[k];
  t.w.assign(h_init, h_init + t.nv * 3);
  t.run(max_outer, max_inner, stats);
  std::memcpy(out, t.w.data(), sizeof(double) * t.nv * 3);
}

// Corner Jacobians of a grid-unit field (for the topology kernel tests).
void orc_corner_jacobians_3d(const double *field, int Nx, int Ny, int Nz,
                             const double *vs, double e1, double e2,
                             double *Jc_out, long long *counts) {
  Topo t;
  t.Nx = Nx;
  t.Ny = Ny;
  t.Nz = Nz;
  t.nv = (long long)Nx * Ny * Nz;
  t.e1 = e1;
  t.e2 = e2;
  for (int k = 0; k < 3; ++k) t.vs[k] = vs[k];
  t.w.assign(field, field + t.nv * 3);
  std::vector<double> Jc;
  long long lo, hi;
  t.corner_jacobians(Jc, lo, hi);
  std::memcpy(Jc_out, Jc.data(), sizeof(double) * Jc.size());
  counts[0] = lo;
  counts[1] = hi;
}

// The three steps of one inner iteration, exposed separately so that tests
// can pin each one (fields here are in physical units, as inside the loop).
static Topo make_topo(int Nx, int Ny, int Nz, const double *vs, double e1,
                      double e2) {
  Topo t;
  t.Nx = Nx;
  t.Ny = Ny;
  t.Nz = Nz;
  t.nv = (long long)Nx * Ny * Nz;
  t.e1 = e1;
  t.e2 = e2;
  for (int k = 0; k < 3; ++k) t.vs[k] = vs[k];
  return t;
}

void orc_topo_gradients(const double *w_phys, int Nx, int Ny, int Nz,
                        const double *vs, double *g_out) {
  Topo t = make_topo(Nx, Ny, Nz, vs, 0.0, 0.0);
  t.w.assign(w_phys, w_phys + t.nv * 3);
  std::vector<double> g;
  t.gradients(g);
  std::memcpy(g_out, g.data(), sizeof(double) * g.size());
}

void orc_topo_limit(double *g, const double *Jc, int Nx, int Ny, int Nz,
                    double e1, double e2) {
  const double ones[3] = {1, 1, 1};
  Topo t = make_topo(Nx, Ny, Nz, ones, e1, e2);
  std::vector<double> gv(g, g + t.nv * 9);
  std::vector<double> jc(Jc, Jc + (long long)(Nx - 1) * (Ny - 1) * (Nz - 1) * 8);
  t.limit(gv, jc);
  std::memcpy(g, gv.data(), sizeof(double) * gv.size());
}

void orc_topo_integrate(const double *g, const double *mean3, int Nx, int Ny,
                        int Nz, const double *vs, double *w_out) {
  Topo t = make_topo(Nx, Ny, Nz, vs, 0.0, 0.0);
  t.w.assign(t.nv * 3, 0.0);
  std::vector<double> gv(g, g + t.nv * 9);
  t.integrate(gv, mean3);
  std::memcpy(w_out, t.w.data(), sizeof(double) * t.nv * 3);
}

void orc_topology_2d(const double *h_init, double e1, double e2, int max_outer,
                     int max_inner, const double *vs, int Nx, int Ny,
                     double *out, double *stats) {
  Topo2 t;
  t.Nx = Nx;
  t.Ny = Ny;
  t.nv = (long long)Nx * Ny;
  t.e1 = e1;
  t.e2 = e2;
  t.vs[0] = vs[0];
  t.vs[1] = vs[1];
  t.w.assign(h_init, h_init + t.nv * 2);
  t.run(max_outer, max_inner, stats);
  std::memcpy(out, t.w.data(), sizeof(double) * t.nv * 2);
}

static Topo2 make_topo2(int Nx, int Ny, const double *vs, double e1, double e2) {
  Topo2 t;
  t.Nx = Nx;
  t.Ny = Ny;
  t.nv = (long long)Nx * Ny;
  t.e1 = e1;
  t.e2 = e2;
  t.vs[0] = vs[0];
  t.vs[1] = vs[1];
  return t;
}

void orc_corner_jacobians_2d(const double *field, int Nx, int Ny, const double *vs,
                             double e1, double e2, double *Jc_out, long long *counts) {
  Topo2 t = make_topo2(Nx, Ny, vs, e1, e2);
  t.w.assign(field, field + t.nv * 2);
  std::vector<double> Jc;
  long long lo, hi;
  t.corner_jacobians(Jc, lo, hi);
  std::memcpy(Jc_out, Jc.data(), sizeof(double) * Jc.size());
  counts[0] = lo;
  counts[1] = hi;
}

void orc_topo2_gradients(const double *w_phys, int Nx, int Ny, const double *vs,
                         double *g_out) {
  Topo2 t = make_topo2(Nx, Ny, vs, 0.0, 0.0);
  t.w.assign(w_phys, w_phys + t.nv * 2);
  std::vector<double> g;
  t.gradients(g);
  std::memcpy(g_out, g.data(), sizeof(double) * g.size());
}

void orc_topo2_limit(double *g, const double *Jc, int Nx, int Ny, double e1, double e2) {
  const double ones[2] = {1, 1};
  Topo2 t = make_topo2(Nx, Ny, ones, e1, e2);
  std::vector<double> gv(g, g + t.nv * 4);
  std::vector<double> jc(Jc, Jc + (long long)(Nx - 1) * (Ny - 1) * 4);
  t.limit(gv, jc);
  std::memcpy(g, gv.data(), sizeof(double) * gv.size());
}

void orc_topo2_integrate(const double *g, const double *mean2, int Nx, int Ny,
                         const double *vs, double *w_out) {
  Topo2 t = make_topo2(Nx, Ny, vs, 0.0, 0.0);
  t.w.assign(t.nv * 2, 0.0);
  std::vector<double> gv(g, g + t.nv * 4);
  t.integrate(gv, mean2);
  std::memcpy(w_out, t.w.data(), sizeof(double) * t.nv * 2);
}

}  // extern "C"
