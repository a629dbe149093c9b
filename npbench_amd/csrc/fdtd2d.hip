// 2-D finite-difference time domain of polybench's fdtd_2d
// (npbench/benchmarks/polybench/fdtd_2d/fdtd_2d_numpy.py:4-11), FP64,
// bitwise the NumPy reference.
//
// A time step is the reference's four in-place statements, which read
// fields the previous statements wrote.  Two launches keep that order without
// races: fdtd_e sets ey's first row to fict[t] and updates ey (rows 1..) and
// ex (columns 1..) from the old hz -- it writes neither field another thread
// reads -- and fdtd_h then updates hz from the new ex and ey.  Each thread
// owns one point; lanes run along the contiguous axis, so every load and
// store is coalesced and the neighbours come from lines the wave has in
// cache.  All three fields stay resident in device memory; a time step moves
// about 9 field-sized passes.
#include <hip/hip_runtime.h>

#include "common.h"

namespace npbx {
namespace {

constexpr int kRowThreads = 256;

// ey[0, :] = fict[t] (:7); ey[i, j] -= 0.5 (hz[i, j] - hz[i-1, j]) for i >= 1
// (:8); ex[i, j] -= 0.5 (hz[i, j] - hz[i, j-1]) for j >= 1 (:9).
__global__ __launch_bounds__(kRowThreads) void fdtd_e(double* __restrict__ ex,
                                                      double* __restrict__ ey,
                                                      const double* __restrict__ hz,
                                                      const double* __restrict__ fict, int t,
                                                      int nx, int ny) {
  const int j = static_cast<int>(blockIdx.x) * kRowThreads + threadIdx.x;
  const int i = static_cast<int>(blockIdx.y);
  if (j >= ny) return;
  const int64_t at = static_cast<int64_t>(i) * ny + j;
  // neighbour indices stay inside the array on the first row and column,
  // where their values are not used
  const double h = hz[at], h_up = hz[i > 0 ? at - ny : at], h_left = hz[j > 0 ? at - 1 : at];
  const double ey_new = ey[at] - 0.5 * (h - h_up);
  ey[at] = i == 0 ? fict[t] : ey_new;
  if (j >= 1) ex[at] = ex[at] - 0.5 * (h - h_left);
}

// hz[i, j] -= 0.7 (((ex[i, j+1] - ex[i, j]) + ey[i+1, j]) - ey[i, j]) for
// i < nx-1, j < ny-1 (:10-11).
__global__ __launch_bounds__(kRowThreads) void fdtd_h(const double* __restrict__ ex,
                                                      const double* __restrict__ ey,
                                                      double* __restrict__ hz, int nx, int ny) {
  const int j = static_cast<int>(blockIdx.x) * kRowThreads + threadIdx.x;
  const int i = static_cast<int>(blockIdx.y);
  if (j >= ny - 1 || i >= nx - 1) return;
  const int64_t at = static_cast<int64_t>(i) * ny + j;
  hz[at] = hz[at] - 0.7 * (((ex[at + 1] - ex[at]) + ey[at + ny]) - ey[at]);
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_fdtd2d(int dtype, void* ex, void* ey, void* hz, const void* fict, int64_t NX,
                           int64_t NY, int64_t tmax, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(dtype == NPBX_F64, "fdtd2d: dtype %s not supported (f64)", dtype_name(dtype));
  NPBX_REQUIRE(NX >= 0 && NY >= 0 && NX <= 65535 && NY <= (int64_t(1) << 24) &&
                   NX * NY <= (int64_t(1) << 31),
               "fdtd2d: NX=%lld NY=%lld out of range", static_cast<long long>(NX),
               static_cast<long long>(NY));
  if (NX == 0 || NY == 0 || tmax <= 0) return NPBX_OK;
  NPBX_REQUIRE(ex && ey && hz && fict, "fdtd2d: null array");
  NPBX_REQUIRE(ex != ey && ey != hz && ex != hz, "fdtd2d: ex, ey and hz must be distinct");
  const hipStream_t s = as_stream(stream);
  const dim3 grid(static_cast<unsigned>(cdiv(NY, kRowThreads)), static_cast<unsigned>(NX));
  auto* e_x = static_cast<double*>(ex);
  auto* e_y = static_cast<double*>(ey);
  auto* h_z = static_cast<double*>(hz);
  const auto* f = static_cast<const double*>(fict);
  for (int64_t t = 0; t < tmax; ++t) {
    hipLaunchKernelGGL(fdtd_e, grid, dim3(kRowThreads), 0, s, e_x, e_y, h_z, f,
                       static_cast<int>(t), static_cast<int>(NX), static_cast<int>(NY));
    NPBX_LAUNCHED("fdtd_e");
    hipLaunchKernelGGL(fdtd_h, grid, dim3(kRowThreads), 0, s, e_x, e_y, h_z,
                       static_cast<int>(NX), static_cast<int>(NY));
    NPBX_LAUNCHED("fdtd_h");
  }
  return NPBX_OK;
}
