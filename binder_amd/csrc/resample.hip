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
//   4. the result is clipped to the input's [min, max] (over all components)
//      and multiplied by a scale factor (the pyramid's current / new
//      resolution).
// Two layouts share the kernels: the C ABI's [nx][ny][nz][comps] doubles
// (components innermost) and a registration handle's [comps][nx][ny][nz]
// float planes.  Element (voxel v, component c) sits at c * cs + v * vs.
#include <algorithm>

#include "common.hpp"
#include "resample.hpp"

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

struct Layout {
  long long cs, vs;  // component and voxel strides
};

// Element number e of a grid-stride loop -> (voxel, component), ordered as
// the layout stores them.
__device__ __forceinline__ void split_element(long long e, int comps, long long nvox,
                                              Layout L, long long &v, int &c) {
  if (L.vs == 1) {  // planes
    c = (int)(e / nvox);
    v = e - (long long)c * nvox;
  } else {  // components innermost
    v = e / comps;
    c = (int)(e - v * comps);
  }
}

// One axis of the separable Gaussian: out = sum_t w[t] in[mirror(p + t - r)]
// along `axis`.  Same layout for in and out.
__global__ __launch_bounds__(256) void k_gauss_axis(const double *__restrict__ in,
                                                    double *__restrict__ out, int nx,
                                                    int ny, int nz, int comps, Layout L,
                                                    int axis, const double *__restrict__ w,
                                                    int radius) {
  const long long total = (long long)nx * ny * nz * comps;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const long long vstep = axis == 0 ? (long long)ny * nz : (axis == 1 ? (long long)nz : 1ll);
  const long long nvox = (long long)nx * ny * nz;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += stride) {
    // consecutive threads take consecutive addresses in either layout
    long long v;
    int c;
    split_element(e, comps, nvox, L, v, c);
    const int z = (int)(v % nz), y = (int)((v / nz) % ny), x = (int)(v / nz / ny);
    const int p = axis == 0 ? x : (axis == 1 ? y : z);
    const long long base = c * L.cs + (v - (long long)p * vstep) * L.vs;
    double acc = 0.0;
    for (int t = -radius; t <= radius; ++t)
      acc += w[t + radius] * in[base + (long long)mirror_index(p + t, n) * vstep * L.vs];
    out[c * L.cs + v * L.vs] = acc;
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

// Linear zoom, clip to [range[0], range[1]], scale; OUT is double or float.
template <typename OUT>
__global__ __launch_bounds__(256) void k_zoom(const double *__restrict__ in, int nx, int ny,
                                              int nz, int comps, Layout Li,
                                              OUT *__restrict__ out, int nxo, int nyo,
                                              int nzo, Layout Lo,
                                              const double *__restrict__ range,
                                              double scale) {
  const long long total = (long long)nxo * nyo * nzo * comps;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const double lo = range[0], hi = range[1];
  const long long nvox = (long long)nxo * nyo * nzo;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += stride) {
    long long v;
    int c;
    split_element(e, comps, nvox, Lo, v, c);
    const int zo = (int)(v % nzo), yo = (int)((v / nzo) % nyo), xo = (int)(v / nzo / nyo);
    const Lerp lx = lerp_at(xo, nx, nxo), ly = lerp_at(yo, ny, nyo), lz = lerp_at(zo, nz, nzo);
    const double *inc = in + c * Li.cs;
    auto at = [&](int x, int y, int z) {
      return inc[(((long long)x * ny + y) * nz + z) * Li.vs];
    };
    const double z00 = (1.0 - lz.w) * at(lx.i0, ly.i0, lz.i0) + lz.w * at(lx.i0, ly.i0, lz.i1);
    const double z01 = (1.0 - lz.w) * at(lx.i0, ly.i1, lz.i0) + lz.w * at(lx.i0, ly.i1, lz.i1);
    const double z10 = (1.0 - lz.w) * at(lx.i1, ly.i0, lz.i0) + lz.w * at(lx.i1, ly.i0, lz.i1);
    const double z11 = (1.0 - lz.w) * at(lx.i1, ly.i1, lz.i0) + lz.w * at(lx.i1, ly.i1, lz.i1);
    const double y0 = (1.0 - ly.w) * z00 + ly.w * z01;
    const double y1 = (1.0 - ly.w) * z10 + ly.w * z11;
    double r = (1.0 - lx.w) * y0 + lx.w * y1;
    r = r < lo ? lo : (r > hi ? hi : r);
    out[c * Lo.cs + v * Lo.vs] = (OUT)(r * scale);
  }
}

__global__ __launch_bounds__(256) void k_f32_to_f64(const float *__restrict__ in,
                                                    double *__restrict__ out, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (double)in[i];
}

// Min and max of n doubles: per-block partials, then one block folds them.
__global__ __launch_bounds__(256) void k_minmax_part(const double *__restrict__ in,
                                                     long long n, double *__restrict__ part) {
  __shared__ double lo[256], hi[256];
  double a = INFINITY, b = -INFINITY;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = in[i];
    a = v < a ? v : a;
    b = v > b ? v : b;
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lo[0];
    part[2 * blockIdx.x + 1] = hi[0];
  }
}

__global__ __launch_bounds__(256) void k_minmax_final(const double *__restrict__ part,
                                                      int nparts, double *__restrict__ range) {
  __shared__ double lo[256], hi[256];
  double a = INFINITY, b = -INFINITY;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    a = fmin(a, part[2 * i]);
    b = fmax(b, part[2 * i + 1]);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    range[0] = lo[0];
    range[1] = hi[0];
  }
}

int grid_for(long long total) {
  long long b = (total + 255) / 256;
  return (int)std::min<long long>(std::max<long long>(b, 1), 8192);
}

std::vector<double> gauss_weights(int n_in, int n_out, int &radius) {
  const double sigma = ((double)n_in / n_out - 1.0) / 2.0;
  radius = (int)(4.0 * sigma + 0.5);
  std::vector<double> wh(2 * radius + 1);
  double sum = 0.0;
  for (int t = -radius; t <= radius; ++t) {
    wh[t + radius] = std::exp(-0.5 / (sigma * sigma) * (double)t * t);
    sum += wh[t + radius];
  }
  for (double &v : wh) v /= sum;
  return wh;
}

// The index grid as a [nx][ny][nz][comps] field: component c of voxel
// (x, y, z) is its index along axis c.
__global__ __launch_bounds__(256) void k_index_grid(double *__restrict__ out, int nx, int ny,
                                                    int nz, int comps) {
  const long long n = (long long)nx * ny * nz * comps;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long v = e / comps;
    const int c = (int)(e - v * comps);
    const int z = (int)(v % nz);
    const int y = (int)((v / nz) % ny);
    const int x = (int)(v / ((long long)nz * ny));
    out[e] = (double)(c == 0 ? x : c == 1 ? y : z);
  }
}

// The resized index grid from its resized axes: component c of output
// voxel (x, y, z) is the resized ramp of axis c at that voxel's index along
// c (axes = [x ramp (nxo) | y ramp (nyo) | z ramp (nzo)]).
__global__ __launch_bounds__(256) void k_grid_from_axes(const double *__restrict__ axes,
                                                        int nxo, int nyo, int nzo, int comps,
                                                        double *__restrict__ out) {
  const long long n = (long long)nxo * nyo * nzo * comps;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long v = e / comps;
    const int c = (int)(e - v * comps);
    const int z = (int)(v % nzo);
    const int y = (int)((v / nzo) % nyo);
    const int x = (int)(v / ((long long)nzo * nyo));
    out[e] = c == 0 ? axes[x] : (c == 1 ? axes[nxo + y] : axes[nxo + nyo + z]);
  }
}

// Gaussian passes on `a` (layout L, scratch b), then the zoom into `out`.
// `range` is a device pointer to {min, max}.  Synchronises the stream.
template <typename OUT>
void resize_device(DBuf<double> &a, DBuf<double> &b, Layout L, int nx, int ny, int nz,
                   int comps, OUT *out, Layout Lo, int nxo, int nyo, int nzo,
                   const double *range, double scale, hipStream_t s) {
  const long long nin = (long long)nx * ny * nz * comps;
  const long long nout = (long long)nxo * nyo * nzo * comps;
  const int nin_ax[3] = {nx, ny, nz}, nout_ax[3] = {nxo, nyo, nzo};
  DBuf<double> w;
  double *cur = a.get();
  for (int axis = 0; axis < 3; ++axis) {
    if (nin_ax[axis] <= nout_ax[axis]) continue;
    int radius = 0;
    const std::vector<double> wh = gauss_weights(nin_ax[axis], nout_ax[axis], radius);
    w.upload(wh.data(), wh.size(), s);
    if (!b.get()) b.alloc((size_t)nin);
    double *dst = cur == a.get() ? b.get() : a.get();
    hipLaunchKernelGGL(k_gauss_axis, dim3(grid_for(nin)), dim3(256), 0, s, cur, dst, nx, ny,
                       nz, comps, L, axis, w.get(), radius);
    HIP_LAUNCH_OK();
    // the weights buffer is re-uploaded for the next axis: wait first
    HIP_OK(hipStreamSynchronize(s));
    cur = dst;
  }
  hipLaunchKernelGGL(k_zoom<OUT>, dim3(grid_for(nout)), dim3(256), 0, s, cur, nx, ny, nz,
                     comps, L, out, nxo, nyo, nzo, Lo, range, scale);
  HIP_LAUNCH_OK();
  HIP_OK(hipStreamSynchronize(s));
}

}  // namespace

void resize_planes_device(const float *src, int nx, int ny, int nz, int comps, float *dst,
                          int nxo, int nyo, int nzo, double scale, hipStream_t s) {
  const long long nin = (long long)nx * ny * nz;
  const long long nout = (long long)nxo * nyo * nzo;
  DBuf<double> a, b, part, range;
  a.alloc((size_t)comps * nin);
  hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_for(comps * nin)), dim3(256), 0, s, src, a.get(),
                     comps * nin);
  HIP_LAUNCH_OK();
  const int nparts = std::min(grid_for(comps * nin), 1024);
  part.alloc((size_t)2 * nparts);
  range.alloc(2);
  hipLaunchKernelGGL(k_minmax_part, dim3(nparts), dim3(256), 0, s, a.get(), comps * nin,
                     part.get());
  HIP_LAUNCH_OK();
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, part.get(), nparts,
                     range.get());
  HIP_LAUNCH_OK();
  resize_device<float>(a, b, Layout{nin, 1}, nx, ny, nz, comps, dst, Layout{nout, 1}, nxo,
                       nyo, nzo, range.get(), scale, s);
}

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
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DBuf<double> a, b, o, part, range;
    a.upload(in, (size_t)nin, s);
    // the clipping range, min and max over every value, reduced on the device
    const int nparts = std::min(grid_for(nin), 1024);
    part.alloc((size_t)2 * nparts);
    range.alloc(2);
    hipLaunchKernelGGL(k_minmax_part, dim3(nparts), dim3(256), 0, s, a.get(), nin, part.get());
    HIP_LAUNCH_OK();
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, part.get(), nparts,
                       range.get());
    HIP_LAUNCH_OK();
    o.alloc((size_t)nout);
    resize_device<double>(a, b, Layout{1, comps}, nx, ny, nz, comps, o.get(),
                          Layout{1, comps}, nxo, nyo, nzo, range.get(), scale, s);
    o.download(out, (size_t)nout, s);
    HIP_OK(hipStreamSynchronize(s));
  });
}

extern "C" int binder_resize_index_grid(int nx, int ny, int nz, int comps, int nxo, int nyo,
                                        int nzo, double scale, double *out) {
  return guarded([&] {
    BINDER_REQUIRE(out, "null argument");
    BINDER_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nxo > 0 && nyo > 0 && nzo > 0,
                   "sizes must be positive");
    BINDER_REQUIRE(comps == 2 || comps == 3, "comps must be 2 or 3");
    BINDER_REQUIRE(comps == 3 || nz == 1, "a 2-component grid has nz == 1");
    const long long nin = (long long)nx * ny * nz * comps;
    const long long nout = (long long)nxo * nyo * nzo * comps;
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DBuf<double> a, b, o, range;
    a.alloc((size_t)nin);
    (void)nin;
    // min and max over every value of the grid: 0 and the largest index
    int hi = std::max(nx, ny);
    if (comps == 3) hi = std::max(hi, nz);
    const double lohi[2] = {0.0, (double)(hi - 1)};
    range.upload(lohi, 2, s);
    // component c of the grid varies along axis c only, and the Gaussian
    // and the linear zoom along any other axis leave a constant constant:
    // each axis's index ramp is resized once, in one dimension, with the
    // same kernels, and broadcast (the full-size volume never exists)
    const int nin_ax[3] = {nx, ny, nz}, nout_ax[3] = {nxo, nyo, nzo};
    DBuf<double> axes;
    axes.alloc((size_t)nxo + nyo + nzo);
    axes.zero(s);
    int off = 0;
    for (int c = 0; c < comps; ++c) {
      a.alloc((size_t)nin_ax[c]);
      b.release();
      hipLaunchKernelGGL(k_index_grid, dim3(grid_for(nin_ax[c])), dim3(256), 0, s, a.get(),
                         nin_ax[c], 1, 1, 1);
      HIP_LAUNCH_OK();
      resize_device<double>(a, b, Layout{1, 1}, nin_ax[c], 1, 1, 1, axes.get() + off,
                            Layout{1, 1}, nout_ax[c], 1, 1, range.get(), scale, s);
      off += nout_ax[c];
    }
    o.alloc((size_t)nout);
    hipLaunchKernelGGL(k_grid_from_axes, dim3(grid_for(nout)), dim3(256), 0, s, axes.get(), nxo,
                       nyo, nzo, comps, o.get());
    HIP_LAUNCH_OK();
    o.download(out, (size_t)nout, s);
    HIP_OK(hipStreamSynchronize(s));
  });
}

// ---------------------------------------------------------------------------
// Host-side intensity binning (BINDER/utils.py:20-82, bin_signal_robust):
// the 1000-bin histogram and the final floor binning over all host cores.
// The arithmetic is numpy's, operation for operation, so the counts and bins
// equal numpy.histogram's and the reference's numpy binning exactly.
// ---------------------------------------------------------------------------
#include <thread>
#include <vector>

namespace {

template <class F>
void host_parallel(long long n, F &&fn) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 32) nt = 32;
  if (n < (1ll << 16)) nt = 1;
  std::vector<std::thread> th;
  const long long per = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const long long b = (long long)t * per, e = std::min<long long>(n, b + per);
    if (b >= e) break;
    th.emplace_back([&fn, t, b, e] { fn(t, b, e); });
  }
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" int binder_min_max(const double *a, long long n, double *mn, double *mx) {
  return binder::guarded([&] {
    BINDER_REQUIRE(a && mn && mx && n > 0, "null or empty input");
    std::vector<double> lo(32, INFINITY), hi(32, -INFINITY);
    std::vector<char> nan(32, 0);
    host_parallel(n, [&](unsigned t, long long b, long long e) {
      double l = INFINITY, h = -INFINITY;
      char bad = 0;
      for (long long i = b; i < e; ++i) {
        const double v = a[i];
        if (v != v) bad = 1;
        l = v < l ? v : l;
        h = v > h ? v : h;
      }
      lo[t] = l;
      hi[t] = h;
      nan[t] = bad;
    });
    double l = INFINITY, h = -INFINITY;
    bool bad = false;
    for (int t = 0; t < 32; ++t) {
      l = lo[t] < l ? lo[t] : l;
      h = hi[t] > h ? hi[t] : h;
      bad = bad || nan[t];
    }
    *mn = bad ? NAN : l;
    *mx = bad ? NAN : h;
  });
}

extern "C" int binder_histogram_uniform(const double *a, long long n, int nbins,
                                        const double *edges, long long *counts) {
  return binder::guarded([&] {
    BINDER_REQUIRE(a && edges && counts && nbins >= 1, "null argument");
    const double first = edges[0], last = edges[nbins];
    const double denom = last - first;
    std::vector<std::vector<long long>> part(32, std::vector<long long>(nbins, 0));
    host_parallel(n, [&](unsigned t, long long b, long long e) {
      long long *c = part[t].data();
      for (long long i = b; i < e; ++i) {
        const double x = a[i];
        if (!(x >= first && x <= last)) continue;  // outside the range, or NaN
        const double f = ((x - first) / denom) * (double)nbins;
        long long k = (long long)f;
        if (k == nbins) k -= 1;
        if (x < edges[k]) k -= 1;
        if (x >= edges[k + 1] && k != nbins - 1) k += 1;
        c[k] += 1;
      }
    });
    for (int k = 0; k < nbins; ++k) {
      long long s = 0;
      for (int t = 0; t < 32; ++t) s += part[t][k];
      counts[k] = s;
    }
  });
}

extern "C" int binder_bin_floor(const double *a, long long n, double lo, double hi,
                                double width, int nbins, int *out) {
  return binder::guarded([&] {
    BINDER_REQUIRE(a && out, "null argument");
    host_parallel(n, [&](unsigned, long long b, long long e) {
      for (long long i = b; i < e; ++i) {
        const double x = a[i];
        int v;
        if (x < lo) {
          v = 0;
        } else if (x > hi) {
          v = nbins - 1;
        } else {
          const double q = std::floor((x - lo) / width);
          // numpy's cast of a NaN or out-of-range value on x86-64
          v = std::fabs(q) < 2147483648.0 ? (int)q : INT32_MIN;
        }
        out[i] = v;
      }
    });
  });
}
