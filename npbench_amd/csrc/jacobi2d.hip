// 2-D 5-point Jacobi sweeps (jacobi_2d_numpy.py:4-10).
//
// Layout: N x N row-major, j (columns) contiguous.  A sweep writes the
// interior [1, N-2]^2 of dst from src; boundaries are never written.
//
// Plain sweep, vector form: each lane owns V consecutive j (one 16-byte
// load: 4 FP32, 2 FP64, 8 BF16); a 256-thread workgroup is 64 lanes x 4 row
// strips and each thread marches down kRowsVec rows (16, or 4 when 16 would
// leave the device with fewer than four workgroups per CU, as a 1000^2 FP64
// grid would) keeping the rows above/at/below its points in registers; the
// loads of kUnroll rows are issued before any is used.  The j+-1 edge scalars are L1 hits; tiles are
// numbered XCD by XCD so neighbouring strips share an L2.  Every src element
// crosses HBM about once per sweep.
//
// Scalar form (rows not 16-byte aligned): one j per lane, same marching.
//
// Arithmetic follows the reference order: ((((c + W) + E) + S) + N) * 0.2,
// with W=[i,j-1], E=[i,j+1], S=[i+1,j], N=[i-1,j]; the library is compiled
// with -ffp-contract=off so the FP64 path is bitwise equal to NumPy.
#include "common.h"
#include "points.h"
#include "vecio.h"

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
    dst[row + j] = T::store(jacobi_point<C>(centre, west, east, south, north));
    north = centre;
    centre = south;
  }
}

constexpr int kUnroll = 4;

template <class S, int kRowsVec>
__global__ __launch_bounds__(kLanes* kStrips) void jacobi2d_sweep_vec(const S* __restrict__ src,
                                                                     S* __restrict__ dst, int n,
                                                                     int tiles_j) {
  using T = Storage<S>;
  using C = typename T::C;
  constexpr int V = Vec16<S>::N;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tj = t % tiles_j;
  const int ti = t / tiles_j;
  const int j0 = (tj * kLanes + threadIdx.x) * V;
  const int i0 = 1 + (ti * kStrips + threadIdx.y) * kRowsVec;
  if (j0 >= n || i0 > n - 2) return;
  const int i_end = min(i0 + kRowsVec, n - 1);
  const bool has_w = j0 > 0, has_e = j0 + V < n;
  const bool full = j0 >= 1 && j0 + V <= n - 1;

  C north[V], centre[V];
  load16(src + static_cast<int64_t>(i0 - 1) * n + j0, north);
  load16(src + static_cast<int64_t>(i0) * n + j0, centre);
  for (int i = i0; i < i_end; i += kUnroll) {
    C south[kUnroll][V], w[kUnroll], e[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t row = static_cast<int64_t>(min(i + u, n - 2)) * n + j0;
      load16(src + row + n, south[u]);
      w[u] = has_w ? T::load(src[row - 1]) : C(0);
      e[u] = has_e ? T::load(src[row + V]) : C(0);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (i + u >= i_end) break;
      C out[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const C west = v == 0 ? w[u] : centre[v - 1];
        const C east = v == V - 1 ? e[u] : centre[v + 1];
        out[v] = jacobi_point<C>(centre[v], west, east, south[u][v], north[v]);
      }
      S* o = dst + static_cast<int64_t>(i + u) * n + j0;
      if (full) {
        store16(o, out);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (j0 + v >= 1 && j0 + v <= n - 2) o[v] = T::store(out[v]);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        north[v] = centre[v];
        centre[v] = south[u][v];
      }
    }
  }
}

template <class S>
int run_sweeps(S* A, S* B, int n, int64_t first, int64_t nsweeps, hipStream_t stream) {
  constexpr int V = Vec16<S>::N;
  const dim3 block(kLanes, kStrips);
  const bool vec = n % V == 0 && aligned16(A) && aligned16(B);
  const int tiles_j = static_cast<int>(cdiv(n, kLanes * V));
  // rows per thread: 16, or 4 when 16 gives fewer than ~4 workgroups per CU
  const bool short_rows = int64_t(tiles_j) * cdiv(n - 2, kStrips * 16) < 1024;
  const int rows = short_rows ? 4 : 16;
  const int tiles_i = static_cast<int>(cdiv(n - 2, kStrips * rows));
  const unsigned nblocks = static_cast<unsigned>(int64_t(tiles_j) * tiles_i);
  const dim3 grid(static_cast<unsigned>(cdiv(n - 2, kLanes)),
                  static_cast<unsigned>(cdiv(n - 2, kStrips * kRows)));
  for (int64_t s = first; s < first + nsweeps; ++s) {
    const bool even = (s % 2) == 0;
    const S* src = even ? A : B;
    S* dst = even ? B : A;
    if (vec && short_rows) {
      hipLaunchKernelGGL((jacobi2d_sweep_vec<S, 4>), dim3(nblocks), block, 0, stream, src, dst,
                         n, tiles_j);
      NPBX_LAUNCHED("jacobi2d_sweep_vec");
    } else if (vec) {
      hipLaunchKernelGGL((jacobi2d_sweep_vec<S, 16>), dim3(nblocks), block, 0, stream, src, dst,
                         n, tiles_j);
      NPBX_LAUNCHED("jacobi2d_sweep_vec");
    } else {
      hipLaunchKernelGGL(jacobi2d_sweep<S>, grid, block, 0, stream, src, dst, n);
      NPBX_LAUNCHED("jacobi2d_sweep");
    }
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
  if (tblock >= 2)
    return tiled_sweeps(0, dtype, A, B, static_cast<int>(N), first, nsweeps, tblock, 0, s);
  NPBX_REQUIRE(dtype_size(dtype) != 0, "jacobi2d: unknown dtype %d", dtype);
  const int deep = jacobi_deep_depth();
  if (deep >= 2 && nsweeps >= 2)
    return jacobi_deep_sweeps(dtype, A, B, static_cast<int>(N), first, nsweeps, deep, s);
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
