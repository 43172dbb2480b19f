// Topology correction of a field that already lives on the device.
#pragma once
#include "common.hpp"

namespace binder {

// In-place topology correction (3D: topology_correction_3D.cpp:17-143,
// N == 2: topology_correction_2D.cpp:17-140) of a grid-unit f32 field in
// [N][Nx*Ny*Nz] planes (Nz == 1 for N == 2), computed in FP64 and written
// back rounded to f32.  vs: N voxel sizes.  stats as in
// binder_topology_correction_3d.  Synchronises `s`.
void topology_on_device(float *d_soa, int N, int Nx, int Ny, int Nz, const double *vs,
                        double e1, double e2, int max_outer, int max_inner, hipStream_t s,
                        double *stats);

}  // namespace binder
