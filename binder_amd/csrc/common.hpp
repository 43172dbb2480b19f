// Shared host/device utilities of the gfx950 backend: error plumbing for the
// C ABI, device buffers, kernel timing on a stream, and the counter-based
// Philox4x32-10 generator.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/binder_c.h"

namespace binder {

// ----------------------------------------------------------------------------
// Errors: C++ exceptions inside the library, integer codes at the C ABI.
// ----------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define BINDER_REQUIRE(cond, msg)                                   \
  do {                                                              \
    if (!(cond)) throw ::binder::Error(BINDER_EINVAL, (msg));       \
  } while (0)

#define HIP_OK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::binder::Error(BINDER_EDEVICE, std::string(#expr) + ": " +       \
                                                hipGetErrorString(e_));       \
  } while (0)

// Check the launch that was just enqueued.
#define HIP_LAUNCH_OK() HIP_OK(hipGetLastError())

template <class F>
int guarded(F &&f) {
  try {
    f();
    set_last_error("");
    return BINDER_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return BINDER_EDEVICE;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return BINDER_EDEVICE;
  }
}

// ----------------------------------------------------------------------------
// Device buffer (owning).
// ----------------------------------------------------------------------------
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    HIP_OK(hipMalloc((void **)&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T *host, size_t count, hipStream_t s) {
    alloc(count);
    HIP_OK(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T *host, size_t count, hipStream_t s) const {
    HIP_OK(hipMemcpyAsync(host, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n) HIP_OK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  T *get() const { return p; }
};

// ----------------------------------------------------------------------------
// Kernel timing: HIP events recorded on the launching stream around every
// launch while enabled; resolved lazily.
// ----------------------------------------------------------------------------
class KernelTimer {
 public:
  ~KernelTimer() { clear(); }
  void enable(bool on) {
    clear();
    enabled_ = on;
  }
  bool enabled() const { return enabled_; }
  // Returns an index to pass to stop(), or -1 when disabled.
  int start(const char *name, hipStream_t s) {
    if (!enabled_) return -1;
    Rec r;
    r.name = name;
    HIP_OK(hipEventCreate(&r.a));
    HIP_OK(hipEventCreate(&r.b));
    HIP_OK(hipEventRecord(r.a, s));
    recs_.push_back(r);
    return (int)recs_.size() - 1;
  }
  void stop(int idx, hipStream_t s) noexcept {
    if (idx < 0) return;
    (void)hipEventRecord(recs_[idx].b, s);
  }
  // Aggregate per name (events must have completed: caller synchronizes).
  void summary(std::vector<std::string> &names, std::vector<double> &ms,
               std::vector<long long> &count) {
    names.clear();
    ms.clear();
    count.clear();
    for (auto &r : recs_) {
      float t = 0.f;
      HIP_OK(hipEventElapsedTime(&t, r.a, r.b));
      size_t k = 0;
      for (; k < names.size(); ++k)
        if (names[k] == r.name) break;
      if (k == names.size()) {
        names.push_back(r.name);
        ms.push_back(0.0);
        count.push_back(0);
      }
      ms[k] += t;
      count[k] += 1;
    }
  }
  void clear() {
    for (auto &r : recs_) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    recs_.clear();
  }

 private:
  struct Rec {
    const char *name;
    hipEvent_t a, b;
  };
  bool enabled_ = false;
  std::vector<Rec> recs_;
};

// RAII scope that times one launch.
struct TimedLaunch {
  KernelTimer *t;
  hipStream_t s;
  int idx;
  TimedLaunch(KernelTimer *t_, const char *name, hipStream_t s_)
      : t(t_), s(s_), idx(t_ ? t_->start(name, s_) : -1) {}
  ~TimedLaunch() {
    if (t) t->stop(idx, s);
  }
};

// ----------------------------------------------------------------------------
// Work-group placement.  MI355X dispatches consecutive work-groups of a grid
// round-robin over its 8 XCDs, each with a private L2.  For kernels that
// re-read a spatial neighbourhood (E-step gathers, Jacobian stencils) a
// physical block id is mapped to a logical one such that each XCD owns one
// contiguous range of logical blocks, keeping neighbourhood reuse in one L2.
// ----------------------------------------------------------------------------
constexpr int kNumXcd = 8;

__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
  if (nb % kNumXcd) return b;
  return (b % kNumXcd) * (nb / kNumXcd) + b / kNumXcd;
}

// ----------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): a stateless, counter-based generator.  Every random
// number used by the sampler is a pure function of (seed, stream, sweep, site),
// so no generator state is stored on the device.
// ----------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline void philox_mulhilo(uint32_t a, uint32_t b,
                                               uint32_t &hi, uint32_t &lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                               uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c.x, hi0, lo0);
    philox_mulhilo(M1, c.z, hi1, lo1);
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Streams (the fourth counter word) used by the sampler.
enum : uint32_t {
  kStreamEStep = 1u,     // categorical draw per voxel
  kStreamTheta = 2u,     // Dirichlet draws
  kStreamDeform = 3u,    // DCT-domain Gaussian noise (component in .z)
  kStreamNu = 4u,        // regularisation-weight Gamma draw
  kStreamNltSample = 5u  // standalone sampleDeformationField
};

// Uniform in (0,1) with 32 bits of resolution.
__host__ __device__ inline double u01_open(uint32_t v) {
  return ((double)v + 0.5) * (1.0 / 4294967296.0);
}
// Uniform in [0,1) with 24 bits (float).
__host__ __device__ inline float u01f(uint32_t v) {
  return (float)(v >> 8) * (1.0f / 16777216.0f);
}

// Two standard normals from two uniforms (Box-Muller).
__host__ __device__ inline void box_muller(uint32_t a, uint32_t b, double &n0,
                                           double &n1) {
  const double u1 = u01_open(a), u2 = u01_open(b);
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586476925286766559 * u2;
  n0 = r * cos(t);
  n1 = r * sin(t);
}

// Gamma(shape, 1) by Marsaglia & Tsang, "A simple method for generating gamma
// variables", ACM TOMS 26(3), 2000; shape < 1 via the boost
// Gamma(a) = Gamma(a+1) * U^(1/a).  Counter c advances per attempt; the loop
// is bounded (acceptance is > 95 % per attempt for shape >= 1).
__host__ __device__ inline double gamma_mt(double shape, u32x4 c, uint32_t k0,
                                           uint32_t k1) {
  double boost = 1.0;
  if (shape < 1.0) {
    u32x4 r = philox4x32_10(u32x4{c.x, c.y, c.z ^ 0x80000000u, c.w}, k0, k1);
    boost = pow(u01_open(r.x), 1.0 / shape);
    shape += 1.0;
  }
  const double d = shape - 1.0 / 3.0;
  const double cc = 1.0 / sqrt(9.0 * d);
  double last = d;
  for (uint32_t attempt = 0; attempt < 64; ++attempt) {
    u32x4 r = philox4x32_10(u32x4{c.x, c.y, c.z + (attempt << 8), c.w}, k0, k1);
    double n0, n1;
    box_muller(r.x, r.y, n0, n1);
    const double v1 = 1.0 + cc * n0;
    if (v1 <= 0.0) continue;
    const double v = v1 * v1 * v1;
    const double u = u01_open(r.z);
    last = d * v;
    // the paper's squeeze u < 1 - 0.0331 x^4 accepts most draws without the
    // two logarithms; it only accepts draws the exact test accepts
    const double x2 = n0 * n0;
    if (u < 1.0 - 0.0331 * x2 * x2) return d * v * boost;
    if (log(u) < 0.5 * x2 + d - d * v + d * log(v)) return d * v * boost;
  }
  return last * boost;
}

}  // namespace binder
