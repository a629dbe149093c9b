// 2-D Gauss-Seidel of polybench's seidel_2d
// (npbench/benchmarks/polybench/seidel_2d/seidel_2d_numpy.py:4-13), FP64,
// bitwise the NumPy reference.
//
// A sweep updates rows 1..N-2 in order: row i adds the sum of its seven
// neighbours -- row i-1 already swept, its own right neighbour and row i+1
// not yet -- and then runs x_j = (x_j + x_{j-1}) / 9 along j, a sequential
// recurrence.  Row i of sweep t needs row i-1 of sweep t and row i+1 of
// sweep t-1, so every row with i + 2t = s can run in slot s once slot s-1
// is done, and no two rows of a slot are neighbours.  One workgroup walks
// the slots with a barrier between them; each wave takes rows of the slot:
// its lanes form the neighbour sums (coalesced, in the reference's order)
// into the wave's LDS row, then one lane runs the recurrence from LDS and
// stores the row, with the short division by 9 (its reciprocal prepared
// once) and the division operator for a row where the short form could not
// vouch.  A launch runs at most kSweepsPerLaunch sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "division.h"

namespace npbx {
namespace {

constexpr int kThreadsS = 1024;
constexpr int kWavesS = kThreadsS / 64;
constexpr int kMaxN = 1280;  // a row per wave in LDS: 16 x 1280 doubles = 160 KiB
constexpr int kSweepsPerLaunch = 64;
constexpr int kScanUnroll = 8;

// Orders the wave's LDS writes before its own lanes' reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kThreadsS) void seidel2d_slots(double* __restrict__ A, int n,
                                                            int sweeps) {
  __shared__ double rows[kWavesS][kMaxN];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* lb = rows[wave];
  double lo = 0.0, hi = INFINITY;
  const Divisor nine = prepare_divisor(9.0, lo, hi);
  const int last_slot = (n - 3) + 2 * (sweeps - 1);
  for (int slot = 0; slot <= last_slot; ++slot) {
    // rows (t, i) with (i - 1) + 2 t = slot, t in [0, sweeps), i in [1, n-2]
    const int t_lo = max(0, (slot - (n - 3) + 1) / 2);
    const int t_hi = min(sweeps - 1, slot / 2);
    for (int t = t_hi - wave; t >= t_lo; t -= kWavesS) {
      const int i = slot - 2 * t + 1;
      const double* up = A + static_cast<int64_t>(i - 1) * n;
      double* row = A + static_cast<int64_t>(i) * n;
      const double* dn = A + static_cast<int64_t>(i + 1) * n;
      // A[i, 1:-1] += (seven neighbours, left to right) (:8-10)
      for (int j = 1 + lane; j < n - 1; j += 64) {
        double s = up[j - 1] + up[j];
        s = s + up[j + 1];
        s = s + row[j + 1];
        s = s + dn[j - 1];
        s = s + dn[j];
        s = s + dn[j + 1];
        lb[j] = row[j] + s;
      }
      wave_lds_sync();
      // A[i, j] = (A[i, j] + A[i, j-1]) / 9 for j = 1 .. n-2 (:11-13); the
      // inputs stay in LDS, so a row the short form cannot vouch for is
      // redone with the division operator
      if (lane == 0) {
        double x = row[0];
        double mn = INFINITY, mx = 0.0;
        int j = 1;
        for (; j + kScanUnroll <= n - 1; j += kScanUnroll) {
          double in[kScanUnroll];  // loaded ahead of the dependent chain
#pragma unroll
          for (int k = 0; k < kScanUnroll; ++k) in[k] = lb[j + k];
#pragma unroll
          for (int k = 0; k < kScanUnroll; ++k) {
            double m;
            x = short_division(in[k] + x, nine.d, nine.y, m);
            mn = fmin(mn, fabs(m));
            mx = fmax(mx, fabs(m));
            row[j + k] = x;
          }
        }
        for (; j < n - 1; ++j) {
          double m;
          x = short_division(lb[j] + x, nine.d, nine.y, m);
          mn = fmin(mn, fabs(m));
          mx = fmax(mx, fabs(m));
          row[j] = x;
        }
        if (!(mn >= lo && mx < hi)) {
          x = row[0];
          for (int j = 1; j < n - 1; ++j) {
            x = (lb[j] + x) / 9.0;
            row[j] = x;
          }
        }
      }
      wave_lds_sync();  // the next row of this wave rewrites lb
    }
    __syncthreads();
  }
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_seidel2d(int dtype, void* A, int64_t N, int64_t tsteps, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(dtype == NPBX_F64, "seidel2d: dtype %s not supported (f64)", dtype_name(dtype));
  NPBX_REQUIRE(N >= 0 && N <= kMaxN, "seidel2d: N=%lld out of range (0 .. %d)",
               static_cast<long long>(N), kMaxN);
  NPBX_REQUIRE(tsteps <= (int64_t(1) << 30), "seidel2d: TSTEPS out of range");
  const int64_t sweeps = tsteps - 1;  // range(0, TSTEPS - 1)
  if (N < 3 || sweeps <= 0) return NPBX_OK;
  NPBX_REQUIRE(A != nullptr, "seidel2d: null array");
  const hipStream_t s = as_stream(stream);
  for (int64_t done = 0; done < sweeps; done += kSweepsPerLaunch) {
    const int k = static_cast<int>(std::min<int64_t>(kSweepsPerLaunch, sweeps - done));
    hipLaunchKernelGGL(seidel2d_slots, dim3(1), dim3(kThreadsS), 0, s, static_cast<double*>(A),
                       static_cast<int>(N), k);
    NPBX_LAUNCHED("seidel2d_slots");
  }
  return NPBX_OK;
}
