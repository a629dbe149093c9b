// 3-D 7-point explicit heat sweeps (heat_3d_numpy.py:4-20).
//
// Layout: N x N x N, C order (i slowest, k contiguous).  A sweep writes the
// interior [1, N-2]^3 of dst from src; boundary faces are never written.
//
// Plain sweep (2.5-D): a 256-thread workgroup covers a 64 (k, one wave,
// coalesced) x 4 (j) tile of columns and each thread marches along i over
// kRows planes, carrying the i-1 / i / i+1 values of its column in registers.
// The j+-1 and k+-1 neighbours of plane i come from the same or adjacent
// cache lines that neighbouring lanes and waves fetch (L1/L2 hits), so each
// src element crosses HBM about once per sweep.
//
// Arithmetic follows the reference: per axis 0.125*((x[+1] - 2*c) + x[-1]),
// the axis terms summed i, j, k left to right, then + c.
#include "common.h"

namespace npbx {
namespace {

constexpr int kLanes = 64;
constexpr int kCols = 4;
constexpr int kRows = 16;

template <class C>
__device__ __forceinline__ C heat_point(C c, C ip, C im, C jp, C jm, C kp, C km) {
  const C c2 = static_cast<C>(2.0) * c;
  const C q = static_cast<C>(0.125);
  const C ti = q * ((ip - c2) + im);
  const C tj = q * ((jp - c2) + jm);
  const C tk = q * ((kp - c2) + km);
  return ((ti + tj) + tk) + c;
}

template <class S>
__global__ __launch_bounds__(kLanes* kCols) void heat3d_sweep(const S* __restrict__ src,
                                                             S* __restrict__ dst, int n) {
  using T = Storage<S>;
  using C = typename T::C;
  const int k = 1 + blockIdx.x * kLanes + threadIdx.x;
  const int j = 1 + blockIdx.y * kCols + threadIdx.y;
  const int i0 = 1 + blockIdx.z * kRows;
  if (k > n - 2 || j > n - 2) return;
  const int i_end = min(i0 + kRows, n - 1);
  const int64_t plane = static_cast<int64_t>(n) * n;
  int64_t p = static_cast<int64_t>(i0) * plane + static_cast<int64_t>(j) * n + k;
  C below = T::load(src[p - plane]);  // i - 1
  C centre = T::load(src[p]);
  for (int i = i0; i < i_end; ++i, p += plane) {
    const C above = T::load(src[p + plane]);  // i + 1
    const C jp = T::load(src[p + n]);
    const C jm = T::load(src[p - n]);
    const C kp = T::load(src[p + 1]);
    const C km = T::load(src[p - 1]);
    dst[p] = T::store(heat_point<C>(centre, above, below, jp, jm, kp, km));
    below = centre;
    centre = above;
  }
}

template <class S>
int run_sweeps(S* A, S* B, int n, int64_t first, int64_t nsweeps, hipStream_t stream) {
  const dim3 block(kLanes, kCols);
  const dim3 grid(static_cast<unsigned>(cdiv(n - 2, kLanes)),
                  static_cast<unsigned>(cdiv(n - 2, kCols)),
                  static_cast<unsigned>(cdiv(n - 2, kRows)));
  for (int64_t s = first; s < first + nsweeps; ++s) {
    const bool even = (s % 2) == 0;
    const S* src = even ? A : B;
    S* dst = even ? B : A;
    hipLaunchKernelGGL(heat3d_sweep<S>, grid, block, 0, stream, src, dst, n);
    NPBX_LAUNCHED("heat3d_sweep");
  }
  return NPBX_OK;
}

int heat_entry(int dtype, void* A, void* B, int64_t N, int64_t first, int64_t nsweeps, int tblock,
               void* stream) {
  NPBX_REQUIRE(N >= 0 && N <= 4096, "heat3d: N=%lld out of range", static_cast<long long>(N));
  NPBX_REQUIRE(nsweeps >= 0 && first >= 0, "heat3d: negative sweep count");
  NPBX_REQUIRE(tblock >= 0, "heat3d: negative tblock");
  if (N < 3 || nsweeps == 0) return NPBX_OK;
  NPBX_REQUIRE(A && B, "heat3d: null array");
  NPBX_REQUIRE(A != B, "heat3d: A and B must be distinct arrays");
  hipStream_t s = as_stream(stream);
  return dispatch_dtype(dtype, [&](auto tag) {
    using S = decltype(tag);
    return run_sweeps<S>(static_cast<S*>(A), static_cast<S*>(B), static_cast<int>(N), first,
                         nsweeps, s);
  });
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_heat3d(int dtype, void* A, void* B, int64_t N, int64_t tsteps, int tblock,
                           void* stream) {
  const int64_t iters = tsteps > 1 ? tsteps - 1 : 0;  // for t in range(1, TSTEPS)
  return npbx::heat_entry(dtype, A, B, N, 0, 2 * iters, tblock, stream);
}

extern "C" int npbx_heat3d_sweeps(int dtype, void* A, void* B, int64_t N, int64_t first,
                                  int64_t nsweeps, int tblock, void* stream) {
  return npbx::heat_entry(dtype, A, B, N, first, nsweeps, tblock, stream);
}
