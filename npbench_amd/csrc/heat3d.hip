// 3-D 7-point explicit heat sweeps (heat_3d_numpy.py:4-20).
//
// Layout: N x N x N, C order (i slowest, k contiguous).  A sweep writes the
// interior [1, N-2]^3 of dst from src; boundary faces are never written.
//
// Plain sweep (2.5-D), vector form: each lane owns V consecutive k (one
// 16-byte load: 4 FP32, 2 FP64, 8 BF16), a wave spans 64*V k, a 256-thread
// workgroup is one wave-row x 4 j-rows, and each thread marches along i over
// kPlanes planes carrying the i-1 / i / i+1 vectors in registers.  Loads for
// kUnroll planes (i+1, j+-1 vectors and the two k-edge scalars) are issued
// before any of them is used, so every lane keeps several 16-byte HBM
// requests in flight.  The j+-1 rows and k edges are L1/L2 hits.  Tiles are
// numbered XCD by XCD (xcd_tile) so neighbours share an L2.  Each src element
// crosses HBM about once per sweep.
//
// Scalar form (rows not 16-byte aligned): one k per lane, same marching.
//
// Arithmetic (points.h): the FP64 drop-in follows the reference order, per
// axis 0.125*((x[+1] - 2*c) + x[-1]), the axis terms summed i, j, k left to
// right, then + c; FP32 arithmetic uses the 7-operation form of the same map.
#include "common.h"
#include "points.h"
#include "vecio.h"

namespace npbx {
namespace {

constexpr int kLanes = 64;
constexpr int kCols = 4;
constexpr int kRows = 16;

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

// kPlanes = planes per workgroup (i segment); kUnroll = planes whose loads
// are issued together.  Both are picked per call from the tuning knob
// "sweep.shape" (sweep_shape below).
template <class S, int kPlanes, int kUnroll>
__global__ __launch_bounds__(kLanes* kCols) void heat3d_sweep_vec(const S* __restrict__ src,
                                                                 S* __restrict__ dst, int n,
                                                                 int tiles_k, int tiles_j) {
  using T = Storage<S>;
  using C = typename T::C;
  constexpr int V = Vec16<S>::N;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tk = t % tiles_k;
  const int tj = (t / tiles_k) % tiles_j;
  const int ti = t / (tiles_k * tiles_j);
  const int k0 = (tk * kLanes + threadIdx.x) * V;
  const int j = 1 + tj * kCols + threadIdx.y;
  const int i0 = 1 + ti * kPlanes;
  if (k0 >= n || j > n - 2) return;
  const int i_end = min(i0 + kPlanes, n - 1);
  const int64_t plane = static_cast<int64_t>(n) * n;
  const int64_t col = static_cast<int64_t>(j) * n + k0;
  const bool has_km = k0 > 0, has_kp = k0 + V < n;
  const bool full = k0 >= 1 && k0 + V <= n - 1;  // all V points interior in k

  C below[V], centre[V];
  load16(src + col + (i0 - 1) * plane, below);
  load16(src + col + i0 * plane, centre);
  for (int i = i0; i < i_end; i += kUnroll) {
    C above[kUnroll][V], jp[kUnroll][V], jm[kUnroll][V], km[kUnroll], kp[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      // clamp so the loads of planes past the segment end stay in bounds
      const int64_t q = col + min(i + u, n - 2) * plane;
      load16(src + q + plane, above[u]);
      load16(src + q + n, jp[u]);
      load16(src + q - n, jm[u]);
      km[u] = has_km ? T::load(src[q - 1]) : C(0);
      kp[u] = has_kp ? T::load(src[q + V]) : C(0);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (i + u >= i_end) break;
      C out[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const C left = v == 0 ? km[u] : centre[v - 1];
        const C right = v == V - 1 ? kp[u] : centre[v + 1];
        out[v] = heat_point<C>(centre[v], above[u][v], below[v], jp[u][v], jm[u][v], right, left);
      }
      S* o = dst + col + static_cast<int64_t>(i + u) * plane;
      if (full) {
        store16(o, out);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (k0 + v >= 1 && k0 + v <= n - 2) o[v] = T::store(out[v]);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        below[v] = centre[v];
        centre[v] = above[u][v];
      }
    }
  }
}

template <class S, int kPlanes, int kUnroll>
int run_vec(S* A, S* B, int n, int64_t first, int64_t nsweeps, hipStream_t stream) {
  constexpr int V = Vec16<S>::N;
  const dim3 block(kLanes, kCols);
  const int tiles_k = static_cast<int>(cdiv(n, kLanes * V));
  const int tiles_j = static_cast<int>(cdiv(n - 2, kCols));
  const int tiles_i = static_cast<int>(cdiv(n - 2, kPlanes));
  const int64_t nblocks = int64_t(tiles_k) * tiles_j * tiles_i;
  NPBX_REQUIRE(nblocks < (int64_t(1) << 31), "heat3d: grid too large");
  for (int64_t s = first; s < first + nsweeps; ++s) {
    const bool even = (s % 2) == 0;
    hipLaunchKernelGGL((heat3d_sweep_vec<S, kPlanes, kUnroll>),
                       dim3(static_cast<unsigned>(nblocks)), block, 0, stream, even ? A : B,
                       even ? B : A, n, tiles_k, tiles_j);
    NPBX_LAUNCHED("heat3d_sweep_vec");
  }
  return NPBX_OK;
}

template <class S>
int run_sweeps(S* A, S* B, int n, int64_t first, int64_t nsweeps, hipStream_t stream) {
  constexpr int V = Vec16<S>::N;
  if (n % V == 0 && aligned16(A) && aligned16(B)) {
    switch (sweep_shape()) {
      case 1: return run_vec<S, 16, 8>(A, B, n, first, nsweeps, stream);
      case 2: return run_vec<S, 32, 8>(A, B, n, first, nsweeps, stream);
      case 3: return run_vec<S, 8, 4>(A, B, n, first, nsweeps, stream);
      default: return run_vec<S, 16, 4>(A, B, n, first, nsweeps, stream);
    }
  }
  const dim3 block(kLanes, kCols);
  const dim3 grid(static_cast<unsigned>(cdiv(n - 2, kLanes)),
                  static_cast<unsigned>(cdiv(n - 2, kCols)),
                  static_cast<unsigned>(cdiv(n - 2, kRows)));
  for (int64_t s = first; s < first + nsweeps; ++s) {
    const bool even = (s % 2) == 0;
    hipLaunchKernelGGL(heat3d_sweep<S>, grid, block, 0, stream, even ? A : B, even ? B : A, n);
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
  if (tblock >= 2)
    return tiled_sweeps(1, dtype, A, B, static_cast<int>(N), first, nsweeps, tblock, 0, s);
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
