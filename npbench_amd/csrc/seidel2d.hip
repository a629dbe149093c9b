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
// the slots with a barrier between them; a group of at most kSweepsPerLaunch
// sweeps is one slot schedule, so a slot holds at most that many rows, and
// their recurrences run side by side in the lanes of one wave (the chain's
// cost is paid once per slot, not once per row).  A launch runs a range of
// one group's slots holding a bounded number of slot-blocks (knob
// "seidel.blocks"): all state is in A, so where the slots are split changes
// nothing, and no launch runs longer than about that budget whatever N is.  Per block of kBlockJ columns: the
// workgroup forms every row's inputs (the neighbour sums, coalesced, in the
// reference's order) into LDS, wave 0 runs the recurrences through the
// block with the short division by 9 (its reciprocal prepared once; a lane
// that cannot vouch for its block redoes it with the division operator), and
// the workgroup stores the block, coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "common.h"
#include "division.h"

namespace npbx {
namespace {

constexpr int kThreadsS = 1024;
constexpr int kSweepsPerLaunch = 64;  // also the most rows a slot holds
constexpr int kBlockJ = 256;          // columns per block: 256 x 64 doubles = 128 KiB of LDS
constexpr int kMaxN = 1 << 14;

// Slot-blocks (slots x 256-column blocks) per launch: about 18 us each for a
// full slot, so the default keeps a launch under about 40 ms.
std::atomic<int> g_seidel_blocks{2048};

// The reference's neighbour sum plus centre for row i at column j (:8-10).
__device__ __forceinline__ double row_input(const double* up, const double* row,
                                            const double* dn, int j) {
  double s = up[j - 1] + up[j];
  s = s + up[j + 1];
  s = s + row[j + 1];
  s = s + dn[j - 1];
  s = s + dn[j];
  s = s + dn[j + 1];
  return row[j] + s;
}

__global__ __launch_bounds__(kThreadsS) void seidel2d_slots(double* __restrict__ A, int n,
                                                            int sweeps, int slot_begin,
                                                            int slot_end) {
  __shared__ double blk[kBlockJ * kSweepsPerLaunch];  // [column in block][row of the slot]
  const int tid = threadIdx.x;
  double lo = 0.0, hi = INFINITY;
  const Divisor nine = prepare_divisor(9.0, lo, hi);
  for (int slot = slot_begin; slot < slot_end; ++slot) {
    // rows (t, i) with (i - 1) + 2 t = slot, t in [0, sweeps), i in [1, n-2];
    // slot row r is t = t_hi - r
    const int t_lo = max(0, (slot - (n - 3) + 1) / 2);
    const int t_hi = min(sweeps - 1, slot / 2);
    const int count = t_hi - t_lo + 1;
    double x = 0.0;  // lane r of wave 0: the recurrence's carry of slot row r
    if (tid < count) x = A[static_cast<int64_t>(slot - 2 * (t_hi - tid) + 1) * n];
    for (int jb = 1; jb < n - 1; jb += kBlockJ) {
      const int bw = min(kBlockJ, n - 1 - jb);
      // 1. inputs of every row of the slot for columns jb .. jb+bw-1, lanes
      //    along j (coalesced), 4 rows at a time over the workgroup
      for (int r = tid / kBlockJ; r < count; r += kThreadsS / kBlockJ) {
        const int jj = tid % kBlockJ;
        if (jj < bw) {
          const int i = slot - 2 * (t_hi - r) + 1;
          const double* row = A + static_cast<int64_t>(i) * n;
          blk[jj * kSweepsPerLaunch + r] = row_input(row - n, row, row + n, jb + jj);
        }
      }
      __syncthreads();
      // 2. the recurrences x_j = (in_j + x_{j-1}) / 9 (:11-13), one lane per
      //    row, results over the inputs; a lane whose |m| left the exact range
      //    recomputes its inputs (its row in memory still holds the old
      //    values) and redoes the block with the division operator
      if (tid < count) {
        const int r = tid;
        const double x0 = x;
        double mn = INFINITY, mx = 0.0;
        for (int jj = 0; jj < bw; ++jj) {
          double m;
          x = short_division(blk[jj * kSweepsPerLaunch + r] + x, nine.d, nine.y, m);
          mn = fmin(mn, fabs(m));
          mx = fmax(mx, fabs(m));
          blk[jj * kSweepsPerLaunch + r] = x;
        }
        if (!(mn >= lo && mx < hi)) {
          const int i = slot - 2 * (t_hi - r) + 1;
          const double* row = A + static_cast<int64_t>(i) * n;
          x = x0;
          for (int jj = 0; jj < bw; ++jj) {
            x = (row_input(row - n, row, row + n, jb + jj) + x) / 9.0;
            blk[jj * kSweepsPerLaunch + r] = x;
          }
        }
      }
      __syncthreads();
      // 3. the block's results back to the rows, coalesced
      for (int r = tid / kBlockJ; r < count; r += kThreadsS / kBlockJ) {
        const int jj = tid % kBlockJ;
        if (jj < bw) {
          const int i = slot - 2 * (t_hi - r) + 1;
          A[static_cast<int64_t>(i) * n + jb + jj] = blk[jj * kSweepsPerLaunch + r];
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace

int seidel_blocks() { return g_seidel_blocks.load(); }
void set_seidel_blocks(int blocks) { g_seidel_blocks.store(blocks); }

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
  const int n = static_cast<int>(N);
  const int blocks_per_slot = static_cast<int>(cdiv(n - 2, kBlockJ));
  const int slots_per_launch = std::max(1, g_seidel_blocks.load() / blocks_per_slot);
  for (int64_t done = 0; done < sweeps; done += kSweepsPerLaunch) {
    const int k = static_cast<int>(std::min<int64_t>(kSweepsPerLaunch, sweeps - done));
    const int slots = (n - 3) + 2 * (k - 1) + 1;  // slots 0 .. (n-3) + 2(k-1)
    for (int s0 = 0; s0 < slots; s0 += slots_per_launch) {
      const int s1 = std::min(slots, s0 + slots_per_launch);
      hipLaunchKernelGGL(seidel2d_slots, dim3(1), dim3(kThreadsS), 0, s, static_cast<double*>(A),
                         n, k, s0, s1);
      NPBX_LAUNCHED("seidel2d_slots");
    }
  }
  return NPBX_OK;
}
