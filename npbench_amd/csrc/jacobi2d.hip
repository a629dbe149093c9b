// 2-D 5-point Jacobi sweeps (jacobi_2d_numpy.py:4-10).
//
// Layout: N x N row-major, j (columns) contiguous.  A sweep writes the
// interior [1, N-2]^2 of dst from src; boundaries are never written.
//
// Plain sweep: a 256-thread workgroup is 64 lanes along j (one wave per
// row strip, coalesced 64-element row segments) x 4 row strips; each thread
// marches down ROWS rows keeping the rows above/at/below its point in
// registers, so every src element is fetched from HBM about once.  The j+-1
// neighbours come from the same cache lines (L1 hits).
//
// Arithmetic follows the reference order: ((((c + W) + E) + S) + N) * 0.2,
// with W=[i,j-1], E=[i,j+1], S=[i+1,j], N=[i-1,j]; the library is compiled
// with -ffp-contract=off so the FP64 path is bitwise equal to NumPy.
#include "common.h"

namespace npbx {
namespace {

constexpr int kLanes = 64;
constexpr int kStrips = 4;
constexpr int kRows = 8;

template <class S>
__global__ __launch_bounds__(kLanes* kStrips) void jacobi2d_sweep(const S* __restrict__ src,
                                                                 S* __restrict__ dst, int n) {
  using T = Storage<S>;
  using C = typename T::C;
  const int j = 1 + blockIdx.x * kLanes + threadIdx.x;
  const int i0 = 1 + (blockIdx.y * kStrips + threadIdx.y) * kRows;
  if (j > n - 2 || i0 > n - 2) return;
  const int i_end = min(i0 + kRows, n - 1);
  int64_t row = static_cast<int64_t>(i0) * n;
  C north = T::load(src[row - n + j]);
  C centre = T::load(src[row + j]);
  for (int i = i0; i < i_end; ++i, row += n) {
    const C south = T::load(src[row + n + j]);
    const C west = T::load(src[row + j - 1]);
    const C east = T::load(src[row + j + 1]);
    C acc = centre + west;
    acc = acc + east;
    acc = acc + south;
    acc = acc + north;
    dst[row + j] = T::store(static_cast<C>(0.2) * acc);
    north = centre;
    centre = south;
  }
}

template <class S>
int run_sweeps(S* A, S* B, int n, int64_t first, int64_t nsweeps, hipStream_t stream) {
  const dim3 block(kLanes, kStrips);
  const dim3 grid(static_cast<unsigned>(cdiv(n - 2, kLanes)),
                  static_cast<unsigned>(cdiv(n - 2, kStrips * kRows)));
  for (int64_t s = first; s < first + nsweeps; ++s) {
    const bool even = (s % 2) == 0;
    const S* src = even ? A : B;
    S* dst = even ? B : A;
    hipLaunchKernelGGL(jacobi2d_sweep<S>, grid, block, 0, stream, src, dst, n);
    NPBX_LAUNCHED("jacobi2d_sweep");
  }
  return NPBX_OK;
}

int jacobi_entry(int dtype, void* A, void* B, int64_t N, int64_t first, int64_t nsweeps,
                 int tblock, void* stream) {
  NPBX_REQUIRE(N >= 0 && N <= (int64_t(1) << 20), "jacobi2d: N=%lld out of range",
               static_cast<long long>(N));
  NPBX_REQUIRE(nsweeps >= 0 && first >= 0, "jacobi2d: negative sweep count");
  NPBX_REQUIRE(tblock >= 0, "jacobi2d: negative tblock");
  if (N < 3 || nsweeps == 0) return NPBX_OK;
  NPBX_REQUIRE(A && B, "jacobi2d: null array");
  NPBX_REQUIRE(A != B, "jacobi2d: A and B must be distinct arrays");
  NPBX_REQUIRE(N * N <= (int64_t(1) << 31), "jacobi2d: N*N exceeds 2^31 elements");
  hipStream_t s = as_stream(stream);
  return dispatch_dtype(dtype, [&](auto tag) {
    using S = decltype(tag);
    return run_sweeps<S>(static_cast<S*>(A), static_cast<S*>(B), static_cast<int>(N), first,
                         nsweeps, s);
  });
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_jacobi2d(int dtype, void* A, void* B, int64_t N, int64_t tsteps, int tblock,
                             void* stream) {
  const int64_t iters = tsteps > 1 ? tsteps - 1 : 0;  // for t in range(1, TSTEPS)
  return npbx::jacobi_entry(dtype, A, B, N, 0, 2 * iters, tblock, stream);
}

extern "C" int npbx_jacobi2d_sweeps(int dtype, void* A, void* B, int64_t N, int64_t first,
                                    int64_t nsweeps, int tblock, void* stream) {
  return npbx::jacobi_entry(dtype, A, B, N, first, nsweeps, tblock, stream);
}
