// 2-D Jacobi sweeps (jacobi_2d_numpy.py:4-10), temporal-blocked in LDS: a
// launch advances every grid point by up to kDeep sweeps.
//
// A workgroup owns an output block of kBI rows x kBJ columns and stages the
// block plus a halo of d cells (d = the launch's depth) in LDS: d + kBI + d
// rows of at most 64 columns, one wave-wide row, lanes along j.  Each sweep
// recomputes the staged region shrunk by one cell per side into the other
// LDS buffer, each of the workgroup's 4, 8 or 16 waves marching a band of
// consecutive rows, and the last writes only the block to global memory: one read
// of the region and one write of the block per d sweeps, and one launch
// instead of d.  Region points on the grid's boundary take, at every level,
// the value that level's boundary holds in the reference (A's at even
// levels, B's at odd: the ping-pong never writes boundaries); points past
// the grid are never read by a point that is computed.  Each level's values
// are rounded through the storage type, as the plain sweep's stores round
// them, and each point is jacobi_point -- the reference's
// ((((c + W) + E) + S) + N) * 0.2 -- so results are bitwise the plain
// sweeps' (FP64: the NumPy reference's).
//
// Rounds follow sweep_rounds (tiles.hip): depths <= kDeep covering all but
// the last sweep, their count of the parity that leaves the last two levels
// in A and B as the reference does, then a depth-1 round.
#include <hip/hip_runtime.h>

#include <atomic>
#include <vector>

#include "common.h"
#include "points.h"
#include "vecio.h"

namespace npbx {
namespace {

constexpr int kDeep = 8;                // most sweeps per launch
constexpr int kRJ = 64;                 // staged columns: one wave
constexpr int kBJ = kRJ - 2 * kDeep;    // output columns per block
constexpr int kBI = 64;                 // output rows per block
constexpr int kRI = kBI + 2 * kDeep;    // staged rows at the deepest launch

std::atomic<int> g_jacobi_deep{kDeep};  // depth per launch; 0 or 1 = plain sweeps
std::atomic<int> g_jacobi_deep_waves{0};  // waves per workgroup: 4, 8, 16 or 0 = by shape

// Full-wave DPP shifts: lane l receives lane l-1's (from_left) or lane l+1's
// (from_right) value; lanes 0 and 63 receive 0, which only lanes outside a
// sweep's columns ever take.
__device__ __forceinline__ int dpp_shr(int x) {
  return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_shl(int x) {
  return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ float lane_from_left(float x) {
  return __int_as_float(dpp_shr(__float_as_int(x)));
}
__device__ __forceinline__ float lane_from_right(float x) {
  return __int_as_float(dpp_shl(__float_as_int(x)));
}
__device__ __forceinline__ double lane_from_left(double x) {
  return __hiloint2double(dpp_shr(__double2hiint(x)), dpp_shr(__double2loint(x)));
}
__device__ __forceinline__ double lane_from_right(double x) {
  return __hiloint2double(dpp_shl(__double2hiint(x)), dpp_shl(__double2loint(x)));
}

template <class S, int NW>
__global__ __launch_bounds__(64 * NW) void jacobi2d_deep(const S* __restrict__ X,
                                                              S* __restrict__ Y,
                                                              const S* __restrict__ A,
                                                              const S* __restrict__ B, int n,
                                                              int parity0, int depth,
                                                              int tiles_j) {
  using St = Storage<S>;
  using C = typename St::C;
  __shared__ C buf[2][kRI * kRJ];
  const int lane = threadIdx.x % 64;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);  // wave-uniform
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ti = tile / tiles_j, tj = tile % tiles_j;
  const int d = depth;
  const int ri = kBI + 2 * d, rj = kBJ + 2 * d;
  const int oi = 1 + ti * kBI - d, oj = 1 + tj * kBJ - d;  // grid point of region (0, 0)
  const int gj = oj + lane;
  const bool col_in = lane < rj && gj >= 0 && gj < n;
  const bool col_bnd = gj == 0 || gj == n - 1;
  auto bnd = [&](int level) { return ((parity0 + level) & 1) ? B : A; };

  // Level 0 into buf[0]: interior points from X, boundary points from the
  // level's array; and buf[1]'s boundary points from the other array.  A
  // buffer holds levels of one parity only (level s in buf[s & 1]) and the
  // sweeps write interior points only, so every level finds its boundary in
  // place.  Addresses are clamped into the grid, so the loads are
  // unconditional (values loaded for points past the grid are never read).
  // All of a wave's loads are issued before the first LDS store, so their
  // latencies overlap.
  const bool touches = oi <= 0 || oj <= 0 || oi + ri >= n || oj + rj >= n;
  const int gjc = min(max(gj, 0), n - 1);
  constexpr int kLoadRows = (kRI + NW - 1) / NW;
  C v[kLoadRows];
#pragma unroll
  for (int k = 0; k < kLoadRows; ++k) {
    const int r = w + k * NW;
    if (r < ri) {
      const int gi = oi + r;
      const int64_t at = static_cast<int64_t>(min(max(gi, 0), n - 1)) * n + gjc;
      const bool edge = col_bnd || gi == 0 || gi == n - 1;
      v[k] = St::load((edge ? bnd(0) : X)[at]);
    }
  }
#pragma unroll
  for (int k = 0; k < kLoadRows; ++k) {
    const int r = w + k * NW;
    if (r < ri) buf[0][r * kRJ + lane] = v[k];
  }
  if (touches) {
#pragma unroll
    for (int k = 0; k < kLoadRows; ++k) {
      const int r = w + k * NW;
      if (r < ri) {
        const int gi = oi + r;
        const int64_t at = static_cast<int64_t>(min(max(gi, 0), n - 1)) * n + gjc;
        v[k] = St::load(bnd(1)[at]);
      }
    }
#pragma unroll
    for (int k = 0; k < kLoadRows; ++k) {
      const int r = w + k * NW;
      const int gi = oi + r;
      if (r < ri && (col_bnd || gi == 0 || gi == n - 1)) buf[1][r * kRJ + lane] = v[k];
    }
  }
  __syncthreads();

  // Each wave marches a band of consecutive rows, keeping the rows above and
  // at its point in registers: one LDS read (the row below) per point, the
  // west and east neighbours from the adjacent lanes (full-wave DPP shifts),
  // and the next row's read issued before the current row is computed.
  // Every lane takes part in the reads and shifts (lanes outside the
  // sweep's columns supply neighbours); only lanes on interior points inside
  // the sweep's region write.
  for (int s = 1; s <= d; ++s) {
    const C* __restrict__ src = buf[(s - 1) & 1];
    C* __restrict__ dst = buf[s & 1];
    const bool lane_on = col_in && !col_bnd && lane >= s && lane < rj - s;
    const int rows = ri - 2 * s;
    const int h = (rows + NW - 1) / NW;
    // the wave's band, clipped to the grid's interior rows
    const int r0 = max(s + w * h, 1 - oi), r1 = min(min(s + (w + 1) * h, ri - s), n - 1 - oi);
    if (r0 < r1) {
      C north = src[(r0 - 1) * kRJ + lane];
      C centre = src[r0 * kRJ + lane];
      C south = src[(r0 + 1) * kRJ + lane];
      for (int r = r0; r < r1; ++r) {
        const C next = src[min(r + 2, ri - 1) * kRJ + lane];  // row r+2, for the next step
        const C west = lane_from_left(centre), east = lane_from_right(centre);
        if (lane_on)
          dst[r * kRJ + lane] = St::load(St::store(jacobi_point<C>(centre, west, east, south,
                                                                   north)));
        north = centre;
        centre = south;
        south = next;
      }
    }
    __syncthreads();
  }

  // the block's interior points of level d
  const C* fin = buf[d & 1];
  const bool lane_out = lane >= d && lane < d + kBJ && gj >= 1 && gj <= n - 2;
  for (int r = d + w; r < d + kBI; r += NW) {
    const int gi = oi + r;
    if (!lane_out || gi < 1 || gi > n - 2) continue;
    Y[static_cast<int64_t>(gi) * n + gj] = St::store(fin[r * kRJ + lane]);
  }
}

}  // namespace

int jacobi_deep_depth() { return g_jacobi_deep.load(); }
void set_jacobi_deep_depth(int d) { g_jacobi_deep.store(d); }
int jacobi_deep_waves() { return g_jacobi_deep_waves.load(); }
void set_jacobi_deep_waves(int w) { g_jacobi_deep_waves.store(w); }
int jacobi_deep_max() { return kDeep; }

// Sweeps [first, first + nsweeps) of one grid through the LDS-blocked kernel
// at depth <= T (2 .. kDeep) per launch.
int jacobi_deep_sweeps(int dtype, void* A, void* B, int n, int64_t first, int64_t nsweeps, int T,
                       hipStream_t stream) {
  NPBX_REQUIRE(T >= 2 && T <= kDeep, "jacobi.deep=%d outside [2, %d]", T, kDeep);
  const int tiles_i = static_cast<int>(cdiv(n - 2, kBI));
  const int tiles_j = static_cast<int>(cdiv(n - 2, kBJ));
  const int64_t ntiles = int64_t(tiles_i) * tiles_j;
  NPBX_REQUIRE(ntiles < (int64_t(1) << 31), "jacobi2d: too many tiles");
  // By shape: 16 waves while the grid has at most two blocks per CU or the
  // staged region is FP64 (80 KiB of LDS: two workgroups per CU), else 8
  // (four 40 KiB workgroups per CU, so the same 32 waves in smaller barrier
  // groups).
  int waves = jacobi_deep_waves();
  if (waves == 0) waves = (ntiles <= 2 * 256 || dtype == NPBX_F64) ? 16 : 8;
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using S = decltype(tag);
    S* a = static_cast<S*>(A);
    S* b = static_cast<S*>(B);
    int64_t level = first;
    // the first round reads the array that holds level `first`; each round
    // writes the other array and the next reads it
    S* X = (first & 1) ? b : a;
    for (int depth : sweep_rounds(nsweeps, T)) {
      S* Y = X == a ? b : a;
      const int parity = static_cast<int>(level & 1);
      const unsigned blocks = static_cast<unsigned>(ntiles);
      if (waves == 4)
        hipLaunchKernelGGL((jacobi2d_deep<S, 4>), dim3(blocks), dim3(256), 0, stream, X, Y, a, b,
                           n, parity, depth, tiles_j);
      else if (waves == 8)
        hipLaunchKernelGGL((jacobi2d_deep<S, 8>), dim3(blocks), dim3(512), 0, stream, X, Y, a, b,
                           n, parity, depth, tiles_j);
      else
        hipLaunchKernelGGL((jacobi2d_deep<S, 16>), dim3(blocks), dim3(1024), 0, stream, X, Y, a,
                           b, n, parity, depth, tiles_j);
      NPBX_LAUNCHED("jacobi2d_deep");
      X = Y;
      level += depth;
    }
    return NPBX_OK;
  });
}

}  // namespace npbx
