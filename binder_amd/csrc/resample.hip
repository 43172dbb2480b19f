// Resampling of a deformation field between pyramid levels, FP64 on the GPU.
//
// Same result as binder_amd.resample.resize (the published behaviour of
// skimage.transform.resize(order=1) that BINDER/FastRegistration.py:566-568
// applies to the field at every resolution change), up to rounding order:
//   1. when an axis shrinks (n_in > n_out) a Gaussian anti-aliasing filter
//      with sigma = (n_in / n_out - 1) / 2, truncated at radius
//      int(4 sigma + 0.5), weights exp(-x^2 / (2 sigma^2)) normalised to one,
//      applied axis by axis (x, then y, then z);
//   2. linear interpolation on the full pixel extent: output index o reads
//      input coordinate (o + 0.5) * n_in / n_out - 0.5;
//   3. both steps extend the input by mirroring about the edge samples
//      (... c b | a b c d | c b ...);
//   4. the result is clipped to the input's [min, max] and multiplied by a
//      scale factor (the pyramid's current / new resolution).
// The field is [nx][ny][nz][comps] (components innermost), as at the C ABI.
#include <algorithm>

#include "common.hpp"

namespace binder {
namespace {

// Mirror an index into [0, n) (period 2(n-1)); n == 1 maps everything to 0.
__device__ __forceinline__ int mirror_index(int i, int n) {
  if (n == 1) return 0;
  const int period = 2 * (n - 1);
  i %= period;
  if (i < 0) i += period;
  return i < n ? i : period - i;
}

// One axis of the separable Gaussian: out = sum_t w[t] in[mirror(p + t - r)]
// along `axis` of a [nx][ny][nz][comps] array.
__global__ __launch_bounds__(256) void k_gauss_axis(const double *__restrict__ in,
                                                    double *__restrict__ out, int nx,
                                                    int ny, int nz, int comps, int axis,
                                                    const double *__restrict__ w,
                                                    int radius) {
  const long long total = (long long)nx * ny * nz * comps;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const long long step = axis == 0 ? (long long)ny * nz * comps
                                   : (axis == 1 ? (long long)nz * comps : (long long)comps);
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += stride) {
    const long long v = e / comps;
    const int z = (int)(v % nz), y = (int)((v / nz) % ny), x = (int)(v / nz / ny);
    const int p = axis == 0 ? x : (axis == 1 ? y : z);
    const long long base = e - (long long)p * step;
    double acc = 0.0;
    for (int t = -radius; t <= radius; ++t)
      acc += w[t + radius] * in[base + (long long)mirror_index(p + t, n) * step];
    out[e] = acc;
  }
}

struct Lerp {
  int i0, i1;
  double w;  // weight of i1
};

__device__ __forceinline__ Lerp lerp_at(int o, int n_in, int n_out) {
  const double c = ((double)o + 0.5) * ((double)n_in / (double)n_out) - 0.5;
  const double f = floor(c);
  Lerp l;
  l.w = c - f;
  l.i0 = mirror_index((int)f, n_in);
  l.i1 = mirror_index((int)f + 1, n_in);
  return l;
}

__global__ __launch_bounds__(256) void k_zoom(const double *__restrict__ in, int nx,
                                              int ny, int nz, int comps,
                                              double *__restrict__ out, int nxo,
                                              int nyo, int nzo, double lo, double hi,
                                              double scale) {
  const long long total = (long long)nxo * nyo * nzo * comps;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += stride) {
    const int c = (int)(e % comps);
    const long long v = e / comps;
    const int zo = (int)(v % nzo), yo = (int)((v / nzo) % nyo), xo = (int)(v / nzo / nyo);
    const Lerp lx = lerp_at(xo, nx, nxo), ly = lerp_at(yo, ny, nyo), lz = lerp_at(zo, nz, nzo);
    auto at = [&](int x, int y, int z) {
      return in[(((long long)x * ny + y) * nz + z) * comps + c];
    };
    const double z00 = (1.0 - lz.w) * at(lx.i0, ly.i0, lz.i0) + lz.w * at(lx.i0, ly.i0, lz.i1);
    const double z01 = (1.0 - lz.w) * at(lx.i0, ly.i1, lz.i0) + lz.w * at(lx.i0, ly.i1, lz.i1);
    const double z10 = (1.0 - lz.w) * at(lx.i1, ly.i0, lz.i0) + lz.w * at(lx.i1, ly.i0, lz.i1);
    const double z11 = (1.0 - lz.w) * at(lx.i1, ly.i1, lz.i0) + lz.w * at(lx.i1, ly.i1, lz.i1);
    const double y0 = (1.0 - ly.w) * z00 + ly.w * z01;
    const double y1 = (1.0 - ly.w) * z10 + ly.w * z11;
    double r = (1.0 - lx.w) * y0 + lx.w * y1;
    r = r < lo ? lo : (r > hi ? hi : r);
    out[e] = r * scale;
  }
}

int grid_for(long long total) {
  long long b = (total + 255) / 256;
  return (int)std::min<long long>(std::max<long long>(b, 1), 8192);
}

}  // namespace
}  // namespace binder

using namespace binder;

extern "C" int binder_resize_field(const double *in, int nx, int ny, int nz,
                                   int comps, int nxo, int nyo, int nzo,
                                   double scale, double *out) {
  return guarded([&] {
    BINDER_REQUIRE(in && out, "null argument");
    BINDER_REQUIRE(nx > 0 && ny > 0 && nz > 0 && comps > 0 && nxo > 0 && nyo > 0 &&
                       nzo > 0,
                   "sizes must be positive");
    const long long nin = (long long)nx * ny * nz * comps;
    const long long nout = (long long)nxo * nyo * nzo * comps;
    double lo = in[0], hi = in[0];
    for (long long i = 1; i < nin; ++i) {
      lo = std::min(lo, in[i]);
      hi = std::max(hi, in[i]);
    }
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DBuf<double> a, b, o, w;
    a.upload(in, (size_t)nin, s);
    const int nin_ax[3] = {nx, ny, nz}, nout_ax[3] = {nxo, nyo, nzo};
    double *cur = a.get();
    for (int axis = 0; axis < 3; ++axis) {
      if (nin_ax[axis] <= nout_ax[axis]) continue;
      const double sigma = ((double)nin_ax[axis] / nout_ax[axis] - 1.0) / 2.0;
      const int radius = (int)(4.0 * sigma + 0.5);
      std::vector<double> wh(2 * radius + 1);
      double sum = 0.0;
      for (int t = -radius; t <= radius; ++t) {
        wh[t + radius] = std::exp(-0.5 / (sigma * sigma) * (double)t * t);
        sum += wh[t + radius];
      }
      for (double &v : wh) v /= sum;
      w.upload(wh.data(), wh.size(), s);
      if (!b.get()) b.alloc((size_t)nin);
      double *dst = cur == a.get() ? b.get() : a.get();
      hipLaunchKernelGGL(k_gauss_axis, dim3(grid_for(nin)), dim3(256), 0, s, cur, dst,
                         nx, ny, nz, comps, axis, w.get(), radius);
      HIP_LAUNCH_OK();
      // the weights buffer is re-uploaded for the next axis: wait first
      HIP_OK(hipStreamSynchronize(s));
      cur = dst;
    }
    o.alloc((size_t)nout);
    hipLaunchKernelGGL(k_zoom, dim3(grid_for(nout)), dim3(256), 0, s, cur, nx, ny, nz,
                       comps, o.get(), nxo, nyo, nzo, lo, hi, scale);
    HIP_LAUNCH_OK();
    o.download(out, (size_t)nout, s);
    HIP_OK(hipStreamSynchronize(s));
  });
}
