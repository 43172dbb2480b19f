// Device-side resampling of a handle's [comps][nx][ny][nz] float field to a
// new grid with the pyramid's resize semantics (see resample.hip).
#pragma once
#include "common.hpp"

namespace binder {

// dst = clip(zoom(gauss(src)), min(src), max(src)) * scale, FP64 inside.
// Synchronises the stream before returning.
void resize_planes_device(const float *src, int nx, int ny, int nz, int comps, float *dst,
                          int nxo, int nyo, int nzo, double scale, hipStream_t s);

}  // namespace binder
