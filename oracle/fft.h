// Oracle FFT/DCT helpers (test infrastructure only: used by the CPU oracle,
// never by the HIP product path).
//
// Mixed-radix complex FFT (decimation in time, any length; radices 2..7 and a
// direct DFT for larger prime factors) and the orthonormal DCT-II / DCT-III
// built on it with Makhoul's N-point reordering (J. Makhoul, "A fast cosine
// transform in one and two dimensions", IEEE TASSP 28(1), 1980).
//
// The orthonormal DCT-II is FFTW's REDFT10 scaled by sqrt(1/(2N)) and an extra
// 1/sqrt(2) at k = 0; its inverse is the orthonormal DCT-III.  This is the
// convention BINDER/src/cpu/DCT.cpp:83-175 establishes around FFTW.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace orc {

using cd = std::complex<double>;

class FFT1D {
 public:
  explicit FFT1D(int n) : n_(n), tw_(n) {
    for (int j = 0; j < n; ++j) {
      const double a = -2.0 * M_PI * double(j) / double(n);
      tw_[j] = cd(std::cos(a), std::sin(a));
    }
    int m = n;
    while (m > 1) {
      int p = 2;
      while (p * p <= m && m % p) ++p;
      if (m % p) p = m;
      factors_.push_back(p);
      m /= p;
    }
  }
  int size() const { return n_; }
  // out = sum_j in[j] exp(sign * 2 pi i j k / n), sign = -1 (forward) or +1.
  void run(const cd *in, cd *out, int sign, cd *work) const {
    rec(in, out, n_, 1, 0, sign, work);
  }

 private:
  // Recursive mixed-radix DIT step on a length-n subsequence with stride s.
  void rec(const cd *in, cd *out, int n, int s, int fi, int sign,
           cd *work) const {
    if (n == 1) {
      out[0] = in[0];
      return;
    }
    const int p = factors_[fi];
    const int m = n / p;
    for (int r = 0; r < p; ++r) rec(in + r * s, out + r * m, m, s * p, fi + 1, sign, work);
    const int tstride = n_ / n;  // twiddle table step for this level
    cd *tmp = work;              // p entries
    for (int k = 0; k < m; ++k) {
      for (int q = 0; q < p; ++q) {
        cd acc = 0.0;
        const int kk = k + q * m;
        for (int r = 0; r < p; ++r) {
          const long long e = (long long)((long long)kk * r % n) * tstride;
          cd w = tw_[(int)(e % n_)];
          if (sign > 0) w = std::conj(w);
          acc += out[r * m + k] * w;
        }
        tmp[q] = acc;
      }
      for (int q = 0; q < p; ++q) out[k + q * m] = tmp[q];
    }
  }
  int n_;
  std::vector<cd> tw_;
  std::vector<int> factors_;
};

// Orthonormal DCT-II / DCT-III of one contiguous line via an N-point FFT.
class DCT1D {
 public:
  explicit DCT1D(int n) : n_(n), fft_(n), v_(n), V_(n), work_(n + 8) {
    ph_.resize(n);
    for (int k = 0; k < n; ++k) {
      const double a = -M_PI * double(k) / (2.0 * n);
      ph_[k] = cd(std::cos(a), std::sin(a));
    }
    scale_ = std::sqrt(1.0 / (2.0 * n));
  }
  // In-place orthonormal DCT-II of x[0..n).
  void forward(double *x) {
    const int n = n_;
    if (n == 1) return;  // REDFT10 of length 1 is 2*x0; ortho scale is 1/2
    for (int i = 0; 2 * i < n; ++i) v_[i] = cd(x[2 * i], 0.0);
    for (int i = 0; 2 * i + 1 < n; ++i) v_[n - 1 - i] = cd(x[2 * i + 1], 0.0);
    fft_.run(v_.data(), V_.data(), -1, work_.data());
    for (int k = 0; k < n; ++k) {
      double X = 2.0 * (ph_[k] * V_[k]).real();  // FFTW REDFT10 value
      double f = scale_;
      if (k == 0) f /= std::sqrt(2.0);
      x[k] = X * f;
    }
  }
  // In-place orthonormal DCT-III (inverse of forward()).
  void inverse(double *c) {
    const int n = n_;
    if (n == 1) return;
    std::vector<double> &X = tmpX_;
    X.resize(n);
    for (int k = 0; k < n; ++k) {
      double f = scale_;
      if (k == 0) f /= std::sqrt(2.0);
      X[k] = c[k] / f;
    }
    // Z_k = e^{-i pi k / 2N} V_k = (X_k - i X_{N-k}) / 2, X_N := 0
    for (int k = 0; k < n; ++k) {
      const double xr = X[k];
      const double xi = (k == 0) ? 0.0 : X[n - k];
      cd Z(0.5 * xr, -0.5 * xi);
      V_[k] = std::conj(ph_[k]) * Z;
    }
    fft_.run(V_.data(), v_.data(), +1, work_.data());
    const double inv_n = 1.0 / n;
    for (int i = 0; 2 * i < n; ++i) c[2 * i] = v_[i].real() * inv_n;
    for (int i = 0; 2 * i + 1 < n; ++i) c[2 * i + 1] = v_[n - 1 - i].real() * inv_n;
  }

 private:
  int n_;
  FFT1D fft_;
  std::vector<cd> v_, V_, work_, ph_;
  std::vector<double> tmpX_;
  double scale_;
};

}  // namespace orc
