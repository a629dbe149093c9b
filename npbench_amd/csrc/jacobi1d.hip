// 1-D three-point Jacobi of polybench's jacobi_1d
// (npbench/benchmarks/polybench/jacobi_1d/jacobi_1d_numpy.py:4-8), FP64,
// bitwise the NumPy reference: B[1:-1] = 0.33333 * ((A[:-2] + A[1:-1]) + A[2:])
// and back, TSTEPS - 1 times.
//
// A grid of 32k points is a few hundred kilobytes and a sweep is a few
// microseconds of work, so one launch per sweep would be bound by launches.
// A launch here is a round of up to kDepth sweeps: each one-wave workgroup
// loads a tile of kTile points (its kSeg outputs and kDepth halo points on
// each side) into registers and sweeps it there kDepth times, the valid
// region shrinking by one point per side per sweep, then stores its outputs.  Rounds follow
// the same depth rule as the 2-D and 3-D temporal tiles (npbx_round_depths):
// the last round is a single sweep, so the last level lands in A and the one
// before it in B, as the reference leaves them.  The global boundary points
// are fixed per array: level l's are A's when l is even, B's when odd.
#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"

namespace npbx {
namespace {

constexpr int kLanes = 64;                    // one wave per tile
constexpr int kPerLane = 4;                   // consecutive points per lane
constexpr int kTile = kLanes * kPerLane;      // tile points, halo included
constexpr int kDepth = 64;                    // sweeps per round at most
constexpr int kSeg = kTile - 2 * kDepth;      // outputs per tile

struct J1Round {
  const double* X;  // level L0 (interior)
  double* Y;        // level L0 + depth (interior) is written here
  const double* A;  // boundary values: level l's are A's (l even) or B's
  const double* B;
  int n;
  int parity0;  // L0 & 1
  int depth;    // 1 .. kDepth
};

// Full-wave DPP shifts of a double (both 32-bit halves): lane l receives lane
// l-1's value (from_left) or lane l+1's (from_right); lanes 0 and 63 receive
// 0, which only ever reaches the tile's halo.
__device__ __forceinline__ double from_left(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// One wave per tile; lane l holds tile points l*kPerLane .. l*kPerLane+3 in
// registers and a sweep takes its two outer neighbours from the adjacent
// lanes by DPP: no LDS, no barrier.  Only the first and last tiles hold a
// global boundary point.
__global__ __launch_bounds__(kLanes) void jacobi1d_round(J1Round g) {
  const int lane = threadIdx.x, n = g.n;
  const int t0 = 1 + static_cast<int>(blockIdx.x) * kSeg - kDepth;  // global index of tile 0
  const int p0 = t0 + lane * kPerLane;  // global index of this lane's first point
  const bool edge = t0 <= 0 || t0 + kTile > n - 1;  // the tile holds index 0 or n-1
  const double a_lo = g.A[0], b_lo = g.B[0], a_hi = g.A[n - 1], b_hi = g.B[n - 1];
  auto fix = [&](double (&v)[kPerLane], int par) {
#pragma unroll
    for (int e = 0; e < kPerLane; ++e) {
      if (p0 + e == 0) v[e] = par ? b_lo : a_lo;
      if (p0 + e == n - 1) v[e] = par ? b_hi : a_hi;
    }
  };
  double v[kPerLane];
#pragma unroll
  for (int e = 0; e < kPerLane; ++e) v[e] = g.X[min(max(p0 + e, 0), n - 1)];
  if (edge) fix(v, g.parity0);
  for (int l = 0; l < g.depth; ++l) {
    const double left = from_left(v[kPerLane - 1]);
    const double right = from_right(v[0]);
    double nv[kPerLane];
#pragma unroll
    for (int e = 0; e < kPerLane; ++e) {
      const double lo = e > 0 ? v[e - 1] : left;
      const double hi = e < kPerLane - 1 ? v[e + 1] : right;
      nv[e] = 0.33333 * ((lo + v[e]) + hi);
    }
    if (edge) fix(nv, (g.parity0 + l + 1) & 1);
#pragma unroll
    for (int e = 0; e < kPerLane; ++e) v[e] = nv[e];
  }
#pragma unroll
  for (int e = 0; e < kPerLane; ++e) {
    const int idx = lane * kPerLane + e;
    const int p = p0 + e;
    if (idx >= kDepth && idx < kTile - kDepth && p >= 1 && p <= n - 2) g.Y[p] = v[e];
  }
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_jacobi1d(int dtype, void* A, void* B, int64_t N, int64_t tsteps,
                             void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(dtype == NPBX_F64, "jacobi1d: dtype %s not supported (f64)", dtype_name(dtype));
  NPBX_REQUIRE(N >= 0 && N <= (int64_t(1) << 30), "jacobi1d: N=%lld out of range",
               static_cast<long long>(N));
  const int64_t nsweeps = 2 * std::max<int64_t>(tsteps - 1, 0);
  if (N < 3 || nsweeps == 0) return NPBX_OK;
  NPBX_REQUIRE(A && B && A != B, "jacobi1d: A and B must be distinct arrays");
  // at most ceil((nsweeps - 1) / kDepth) + 1 rounds of depth <= kDepth, + 1
  std::vector<int32_t> depths(static_cast<size_t>(cdiv(nsweeps, kDepth) + 2));
  int32_t count = 0;
  NPBX_TRY(npbx_round_depths(nsweeps, kDepth, depths.data(),
                             static_cast<int32_t>(depths.size()), &count));
  depths.resize(static_cast<size_t>(count));
  const hipStream_t s = as_stream(stream);
  const unsigned nb = static_cast<unsigned>(cdiv(N - 2, kSeg));
  J1Round g;
  g.A = static_cast<const double*>(A);
  g.B = static_cast<const double*>(B);
  g.n = static_cast<int>(N);
  // round 1 reads A (level 0); each round writes the other array
  double* X = static_cast<double*>(A);
  int64_t level = 0;
  for (int32_t d : depths) {
    g.X = X;
    g.Y = X == A ? static_cast<double*>(B) : static_cast<double*>(A);
    g.parity0 = static_cast<int>(level & 1);
    g.depth = d;
    hipLaunchKernelGGL(jacobi1d_round, dim3(nb), dim3(kLanes), 0, s, g);
    NPBX_LAUNCHED("jacobi1d_round");
    X = g.Y;
    level += d;
  }
  return NPBX_OK;
}
