// Alternating-direction implicit time steps of polybench's ADI
// (npbench/benchmarks/polybench/adi/adi_numpy.py:6-52), FP64, bitwise the
// NumPy reference.
//
// One time step is a column sweep -- for every column i of u a tridiagonal
// solve along the rows, written to the scratch v -- and a row sweep -- for
// every row i of v a solve along the columns, written back to u.  Each solve
// is the reference's recurrence: a forward pass
//     q[j] = (rhs[j] - a q[j-1]) / (a p[j-1] + b)
// and a back substitution x[j] = p[j] x[j+1] + q[j].  p and the denominators
// do not depend on the data and are the same for every line, so they are
// computed once per call (adi_tables) with the reference's own operations.
//
// One thread owns one line (a column in the first half, a row in the
// second).  A line is a serial chain of divisions, so the time is the
// chain's latency.  Three things keep the chain short:
//  * the steps of a 16-step chunk have no branches between them, so the
//    right-hand sides and their LDS reads are scheduled off the chain (steps
//    past the end compute values nothing reads, into padded scratch);
//  * the division q = n / den uses the reciprocal the IEEE division sequence
//    derives from den alone (precomputed per step), leaving three dependent
//    operations on the chain; it is the same sequence, so the same bits,
//    whenever that sequence does not rescale its operands, which each step
//    checks off the chain -- a chunk where any lane fails the check is redone
//    with the division operator;
//  * memory stays off it: the column sweep reads rows of u and writes rows of
//    v (lanes at consecutive addresses); the row sweep, whose columns are
//    strided in memory, stages 16-column tiles of v through LDS (loaded
//    row-contiguous by the whole workgroup) and gathers its output columns
//    in LDS tiles the workgroup stores row-contiguous; the forward passes
//    keep q in a scratch array laid out [step][line], read back one chunk
//    ahead by the back substitutions.
//
// Grids of up to 256 lines run every time step in one workgroup and one
// launch, a barrier between the halves; larger ones launch each half over
// ceil(lines / 256) workgroups (one wave per SIMD either way: two waves on a
// SIMD would share its issue slots, and the chain is issue-bound there).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "division.h"

namespace npbx {
namespace {

constexpr int kCW = 16;  // steps per LDS tile and per q prefetch chunk

// The reference's scalars (adi_numpy.py:12-25), computed in the same order.
struct AdiCoef {
  double a, b, c, d, e, f;
};

AdiCoef adi_coef(int64_t tsteps, int64_t n) {
  const double DX = 1.0 / static_cast<double>(n);
  const double DY = 1.0 / static_cast<double>(n);
  const double DT = 1.0 / static_cast<double>(tsteps);
  const double B1 = 2.0, B2 = 1.0;
  const double mul1 = B1 * DT / (DX * DX);
  const double mul2 = B2 * DT / (DY * DY);
  AdiCoef k;
  k.a = -mul1 / 2.0;
  k.b = 1.0 + mul2;
  k.c = k.a;
  k.d = -mul2 / 2.0;
  k.e = 1.0 + mul2;
  k.f = k.d;
  return k;
}

// Tables of one sweep, tl >= n entries each (entry j for step j = 1 .. n-2;
// entries past n-2 repeat entry n-2, so a 16-step chunk reads a contiguous
// block without clamping).
struct SweepTab {
  const double* p;    // p[j]
  const double* den;  // the forward denominators, lo p[j-1] + mid
  const double* y;    // the division sequence's reciprocal of den[j]
  const double* ok;   // ok[0] = 1 when every den[j] is in den_in_range;
                      // ok[1], ok[2]: quotient estimates m = n y with
                      // ok[1] <= |m| < ok[2] never make it rescale
};

struct AdiArgs {
  double* u;   // (n, n), in place
  double* v;   // (n + 16, ldv) scratch: the column sweep's result, rows past
               // n-1 take the writes of steps past the end
  double* qs;  // (rows, ld) scratch: q of step j of line l at j*ld + l
  SweepTab cols, rows;
  // right-hand sides, as the reference groups them:
  //   column sweep ((-d u[j,i-1] + (1+2d) u[j,i]) - f u[j,i+1]) - a q
  //   row sweep    ((-a v[i-1,j] + (1+2a) v[i,j]) - c v[i+1,j]) - d q
  double nd, c1, f, a;
  double na, c2, c, d;
  int n;
  int ld;
  int ldv;
  int tsteps;
};


// p[j] = -c / (a p[j-1] + b) with p[0] = 0 (adi_numpy.py:28,30) and
// p[j] = -f / (d p[j-1] + e) (:38,40), for j = 1 .. n-2 (lane 0: the column
// sweep's table, lane 1: the row sweep's), with each den's reciprocal; each
// table tl entries long.
__global__ void adi_tables(AdiCoef k, int n, int tl, double* __restrict__ tab) {
  const int lane = threadIdx.x;
  if (lane > 1) return;
  const double lo = lane == 0 ? k.a : k.d;
  const double mid = lane == 0 ? k.b : k.e;
  const double hi = lane == 0 ? k.c : k.f;
  double* p = tab + (lane * 4) * tl;
  double* den = p + tl;
  double* y = den + tl;
  double* ok = y + tl;
  double prev = 0.0, dmin = INFINITY;
  bool in_range = true;
  p[0] = den[0] = y[0] = 0.0;
  for (int j = 1; j < n - 1; ++j) {
    const double dj = lo * prev + mid;
    prev = -hi / dj;
    den[j] = dj;
    p[j] = prev;
    y[j] = division_reciprocal(dj);
    in_range = in_range && den_in_range(dj);
    dmin = fmin(dmin, fabs(dj));
  }
  for (int j = n - 1; j < tl; ++j) {
    p[j] = p[n - 2];
    den[j] = den[n - 2];
    y[j] = y[n - 2];
  }
  ok[0] = in_range ? 1.0 : 0.0;
  quotient_bounds(dmin, ok[1], ok[2]);
}

// LDS of one workgroup: two buffers, each a tile of at most (NT + 2) x
// (kCW + 1) doubles followed by the chunk's (den, y) pairs.
template <int NT>
struct AdiLds {
  static constexpr int TILE = (NT + 2) * (kCW + 1);  // even: the pairs stay 16-byte aligned
  static constexpr int BUF = TILE + 2 * kCW;
  static constexpr int ELEMS = 2 * BUF;
};

// The back substitutions read p through the constant address space: the
// indices are wave-uniform, so the reads are scalar loads into SGPRs.
typedef const __attribute__((address_space(4))) double* ConstTab;
__device__ __forceinline__ ConstTab as_const(const double* p) {
  return (ConstTab)(p);  // address-space cast: C-style casts only
}

// Entries j0 .. j0+kCW-1 of a table: one contiguous block.
__device__ __forceinline__ void table_chunk(ConstTab tab, int j0, double (&d)[kCW]) {
  const ConstTab b = tab + j0;
#pragma unroll
  for (int kk = 0; kk < kCW; ++kk) d[kk] = b[kk];
}

// Row j of a [row][line] scratch array: a wave-uniform base, so that the
// lanes' accesses are one scalar base plus a 32-bit lane offset.
__device__ __forceinline__ double* row_of(double* base, int j, int pitch) {
  return base + static_cast<uint32_t>(j * pitch);
}

// Where a line keeps q between its forward pass and its back substitution:
// step j at base[j * pitch] -- the [step][line] scratch in device memory, or,
// for one-wave workgroups on grids of at most kQRows steps, an LDS array of
// the same layout (no device-memory round trip per step).
struct QSlots {
  double* base;
  int pitch;
  __device__ __forceinline__ double& at(int j) const {
    return base[static_cast<uint32_t>(j * pitch)];
  }
};
constexpr int kQRows = 280;

// Hides the thread index, the side and the line base from loop-invariant code
// motion: the one-launch kernel runs both sweeps in a time-step loop, and
// addresses hoisted out of it would hold both sweeps' registers at once.
__device__ __forceinline__ void opaque(int& t, int& n, int& lb) {
  asm volatile("" : "+v"(t));
  asm volatile("" : "+s"(n), "+s"(lb));
}

// A workgroup barrier that orders LDS only.  The chunk loops' barriers guard
// the LDS tiles; with __syncthreads they would also wait for every global
// store in flight (the per-step stores of q and v), which the other threads
// never read within the sweep.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-sweep constants of the forward pass.
struct Fwd {
  double w;       // weight of q[j-1]
  bool fast;      // every den in range: try the short division
  double lo, hi;  // its numerator bounds
};

__device__ __forceinline__ Fwd forward_consts(const SweepTab& tab, double w) {
  Fwd f;
  f.w = w;
  f.fast = __builtin_amdgcn_readfirstlane(static_cast<int>(tab.ok[0] != 0.0)) != 0;
  f.lo = tab.ok[1];
  f.hi = tab.ok[2];
  return f;
}

// The forward pass over one chunk: q[j] = (rhs(jj) - w q[j-1]) / den[j] for
// steps j = j0 + jj, stored to row j of qs.  rhs(jj) reads the chunk's LDS
// tile, dy the chunk's (den, y) pairs.  Steps past n-2 compute values that
// nothing reads.
// The chunk's right-hand sides and table pairs are read from LDS and formed
// before the recurrence starts: none depends on q, and with them in registers
// the chain waits for nothing but itself.
template <class Rhs>
__device__ __forceinline__ void forward_chunk(double& q, const Fwd& f, const double* dy, int j0,
                                              const QSlots& qs, Rhs rhs) {
  double rv[kCW], den[kCW], y[kCW];
#pragma unroll
  for (int jj = 0; jj < kCW; ++jj) {
    const double2 d = *reinterpret_cast<const double2*>(dy + 2 * jj);
    den[jj] = d.x;
    y[jj] = d.y;
    rv[jj] = rhs(jj);
  }
  const double q0 = q;
  bool redo = !f.fast;
  if (f.fast) {
    // the quotient estimates' magnitude range over the chunk; lanes past
    // the last line and steps past n-2 take part too (their values come from
    // clamped loads of the grid, so they only rarely cause a redo).  A NaN
    // (from a NaN in u) is left out and stays NaN in both forms.
    double amin = INFINITY, amax = 0.0;
#pragma unroll
    for (int jj = 0; jj < kCW; ++jj) {
      const double num = rv[jj] - f.w * q;
      double m;
      q = short_division(num, den[jj], y[jj], m);
      qs.at(j0 + jj) = q;
      amin = fmin(amin, fabs(m));
      amax = fmax(amax, fabs(m));
    }
    redo = __any(!(amin >= f.lo && amax < f.hi));
  }
  if (redo) {
    q = q0;
#pragma unroll
    for (int jj = 0; jj < kCW; ++jj) {
      q = (rv[jj] - f.w * q) / den[jj];
      qs.at(j0 + jj) = q;
    }
  }
}

// Besides its kCW tile values, each of the lanes 0..63 loads one more value
// per chunk: lanes 0..31 the tile's last two columns or rows, lanes 32..63
// the chunk's table pairs (den of step j0 + (t-32), then y).  One load from
// a per-lane address, so that nothing is chosen by a branch (a branch here
// made the compiler keep the two values in a stack array).
__device__ __forceinline__ const double* pair_addr(const SweepTab& tab, int t, int j0) {
  const int e = (t - 2 * kCW) & (2 * kCW - 1);
  return (e < kCW ? tab.den : tab.y) + j0 + (e & (kCW - 1));
}
__device__ __forceinline__ void commit_pair(double* dy, int t, double val) {
  const int e = t - 2 * kCW;
  if (e >= 0 && e < 2 * kCW) dy[2 * (e & (kCW - 1)) + (e >= kCW ? 1 : 0)] = val;
}

// The back substitution over one chunk: x[j] = p[j] x[j+1] + q[j] for
// j = j0+kCW-1 down to j0, emit(kk, x) after each.  PARTIAL: the chunk runs
// past n-2 (the first one processed); its steps past the end leave x alone.
template <bool PARTIAL, class Emit>
__device__ __forceinline__ void back_chunk(double& x, const double (&q)[kCW],
                                           const double (&p)[kCW], int n, int j0, Emit emit) {
#pragma unroll
  for (int kk = kCW - 1; kk >= 0; --kk) {
    const double nx = p[kk] * x + q[kk];
    if constexpr (PARTIAL)
      x = j0 + kk <= n - 2 ? nx : x;
    else
      x = nx;
    emit(kk, x);
  }
}

// This thread's q of steps j0 .. j0+kCW-1.
__device__ __forceinline__ void q_chunk(const QSlots& qs, int j0, double (&q)[kCW]) {
#pragma unroll
  for (int kk = 0; kk < kCW; ++kk) q[kk] = qs.at(j0 + kk);
}

// Runs process(c, q, partial) for the chunks c = nchunks-1 down to 0, each
// with its q loaded three chunks ahead into one of four register sets used
// in turn (no copies; a chunk's 16 steps are too short to cover an L2 round
// trip); the first chunk processed is the partial one.
template <class Process>
__device__ __forceinline__ void back_chunks(int nchunks, const QSlots& qs, Process process) {
  double qr[4][kCW];
  auto load = [&](int c, double (&dst)[kCW]) {
    if (c >= 0) q_chunk(qs, 1 + c * kCW, dst);
  };
  int c = nchunks - 1;
  load(c, qr[0]);
  load(c - 1, qr[1]);
  load(c - 2, qr[2]);
  load(c - 3, qr[3]);
  process(c, qr[0], std::true_type{});
  load(c - 4, qr[0]);
  // chunk c - k sits in set k % 4
  for (c -= 1; c >= 0; c -= 4) {
    process(c, qr[1], std::false_type{});
    load(c - 4, qr[1]);
    if (c - 1 < 0) break;
    process(c - 1, qr[2], std::false_type{});
    load(c - 5, qr[2]);
    if (c - 2 < 0) break;
    process(c - 2, qr[3], std::false_type{});
    load(c - 6, qr[3]);
    if (c - 3 < 0) break;
    process(c - 3, qr[0], std::false_type{});
    load(c - 7, qr[0]);
  }
}

// Column sweep (adi_numpy.py:27-34) of the lines lb+1 .. lb+NT: column i
// of u solved along the rows into column i of v.
template <int NT, bool QLDS>
__device__ __forceinline__ void column_sweep(const AdiArgs& g, int lb, double* lds,
                                             double* qlds) {
  using L = AdiLds<NT>;
  constexpr int W = NT + 2;  // tile columns lb .. lb+NT+1
  int t = threadIdx.x, n = g.n;
  opaque(t, n, lb);
  const int i = lb + 1 + t;
  const int nchunks = (n - 2 + kCW - 1) / kCW;
  const Fwd f = forward_consts(g.cols, g.a);
  QSlots qs;
  if constexpr (QLDS)
    qs = {qlds + t, NT};
  else
    qs = {g.qs + lb + t, g.ld};
  // Tile of rows j0 .. j0+kCW-1 of u, columns lb .. lb+W-1, tile[row][col]:
  // thread t loads column lb+t of every row, threads 0..31 also the last two
  // columns (row j0 + t/2, column lb+NT + t%2).  Rows and columns are
  // clamped into the grid: what a clamped load reads only reaches steps past
  // n-2 and lanes past the last line.
  const int col = min(lb + t, n - 1);
  const int xr = t >> 1, xcol = min(lb + NT + (t & 1), n - 1);
  double pre[kCW], extra;
  auto fetch = [&](int c) {
    const int j0 = 1 + c * kCW;
#pragma unroll
    for (int m = 0; m < kCW; ++m) pre[m] = g.u[static_cast<int64_t>(min(j0 + m, n - 2)) * n + col];
    const double* xa = g.u + static_cast<int64_t>(min(j0 + (xr & (kCW - 1)), n - 2)) * n + xcol;
    extra = *(t < 2 * kCW ? xa : pair_addr(g.cols, t, j0));
  };
  double q = 1.0;  // q[0] = v[0, i] = 1 (:29-31)
  fetch(0);
  for (int c = 0; c < nchunks; ++c) {
    const int j0 = 1 + c * kCW;
    double* buf = lds + (c & 1) * L::BUF;
    double* dy = buf + L::TILE;
#pragma unroll
    for (int m = 0; m < kCW; ++m) buf[m * W + t] = pre[m];
    if (t < 2 * kCW) buf[xr * W + NT + (t & 1)] = extra;
    commit_pair(dy, t, extra);
    lds_barrier();
    if (c + 1 < nchunks) fetch(c + 1);
    forward_chunk(q, f, dy, j0, qs, [&](int jj) {
      const double* r = buf + jj * W + t;
      return (g.nd * r[0] + g.c1 * r[1]) - g.f * r[2];
    });
  }
  // back substitution (:32-34) with v[n-1, i] = 1.  Every lane writes (the
  // pitch holds the lanes past the last line); steps past the end write
  // rows n .. n+15.
  row_of(g.v, 0, g.ldv)[i] = 1.0;
  row_of(g.v, n - 1, g.ldv)[i] = 1.0;
  double x = 1.0;
  const ConstTab ptab = as_const(g.cols.p);
  back_chunks(nchunks, qs, [&](int c, const double (&qc)[kCW], auto partial) {
    const int j0 = 1 + c * kCW;
    double pj[kCW];
    table_chunk(ptab, j0, pj);
    back_chunk<decltype(partial)::value>(x, qc, pj, n, j0, [&](int kk, double xv) {
      const int j = j0 + kk;
      row_of(g.v, decltype(partial)::value && j > n - 2 ? j + 1 : j, g.ldv)[i] = xv;
    });
  });
}

// Row sweep (adi_numpy.py:36-46) of the lines lb+1 .. lb+NT: row i of v
// solved along the columns into row i of u.
template <int NT, bool QLDS>
__device__ __forceinline__ void row_sweep(const AdiArgs& g, int lb, double* lds, double* qlds) {
  using L = AdiLds<NT>;
  constexpr int P = kCW + 1;    // LDS row pitch (odd: lanes hit different banks)
  constexpr int RS = NT / kCW;  // tile rows one pass of the workgroup covers
  int t = threadIdx.x, n = g.n;
  opaque(t, n, lb);
  const int i = lb + 1 + t;
  const bool live = i <= n - 2;
  const int nchunks = (n - 2 + kCW - 1) / kCW;
  const Fwd f = forward_consts(g.rows, g.d);
  QSlots qs;
  if constexpr (QLDS)
    qs = {qlds + t, NT};
  else
    qs = {g.qs + lb + t, g.ld};
  // Tile of columns j0 .. j0+kCW-1 of v, rows lb .. lb+NT+1, tile[row][col]:
  // kCW consecutive threads read one row's run (thread t: column j0 + t%kCW
  // of rows lb + t/kCW + m*RS), threads 0..31 also the last two rows; rows
  // and columns clamped into the grid as in the column sweep.
  const int tj = t % kCW, tr = t / kCW;
  double pre[kCW], extra;
  auto fetch = [&](int c) {
    const int colj = min(1 + c * kCW + tj, n - 2);
#pragma unroll
    for (int m = 0; m < kCW; ++m) pre[m] = row_of(g.v, min(lb + tr + m * RS, n - 1), g.ldv)[colj];
    const double* xa = row_of(g.v, min(lb + NT + (tr & 1), n - 1), g.ldv) + colj;
    extra = *(t < 2 * kCW ? xa : pair_addr(g.rows, t, 1 + c * kCW));
  };
  double q = 1.0;  // q[0] = u[i, 0] = 1 (:37-39)
  fetch(0);
  for (int c = 0; c < nchunks; ++c) {
    const int j0 = 1 + c * kCW;
    double* buf = lds + (c & 1) * L::BUF;
    double* dy = buf + L::TILE;
#pragma unroll
    for (int m = 0; m < kCW; ++m) buf[(tr + m * RS) * P + tj] = pre[m];
    if (t < 2 * kCW) buf[(NT + tr) * P + tj] = extra;
    commit_pair(dy, t, extra);
    lds_barrier();
    if (c + 1 < nchunks) fetch(c + 1);
    forward_chunk(q, f, dy, j0, qs, [&](int jj) {
      const double* r = buf + t * P + jj;
      return (g.na * r[0] + g.c2 * r[P]) - g.c * r[2 * P];
    });
  }
  lds_barrier();  // the tiles are reused for the output below
  // back substitution (:44-46) with u[i, n-1] = 1; u[i, 0] = 1 (:37)
  if (live) {
    g.u[static_cast<int64_t>(i) * n] = 1.0;
    g.u[static_cast<int64_t>(i) * n + n - 1] = 1.0;
  }
  double x = 1.0;
  const ConstTab ptab = as_const(g.rows.p);
  const bool all_rows = lb + NT <= n - 2;  // every line of the workgroup is live
  back_chunks(nchunks, qs, [&](int c, const double (&qc)[kCW], auto partial) {
    const int j0 = 1 + c * kCW;
    double pj[kCW];
    table_chunk(ptab, j0, pj);
    double* ob = lds + (c & 1) * L::BUF;  // ob[line][column - j0]
    back_chunk<decltype(partial)::value>(x, qc, pj, n, j0,
                                         [&](int kk, double xv) { ob[t * P + kk] = xv; });
    lds_barrier();
    // rows lb+1 .. lb+NT, columns j0 .. j0+kCW-1 of u, kCW threads per row run
    const int colj = j0 + tj;
    double* urow = g.u + static_cast<int64_t>(lb + 1 + tr) * n + colj;
    const int64_t rstep = static_cast<int64_t>(RS) * n;
    if (!decltype(partial)::value && all_rows) {
#pragma unroll
      for (int m = 0; m < kCW; ++m) urow[m * rstep] = ob[(tr + m * RS) * P + tj];
    } else if (colj <= n - 2) {
#pragma unroll
      for (int m = 0; m < kCW; ++m)
        if (lb + 1 + tr + m * RS <= n - 2) urow[m * rstep] = ob[(tr + m * RS) * P + tj];
    }
  });
}

template <int NT>
__global__ __launch_bounds__(NT) void adi_fused(AdiArgs g) {
  __shared__ __align__(16) double lds[AdiLds<NT>::ELEMS];
  for (int s = 0; s < g.tsteps; ++s) {
    column_sweep<NT, false>(g, 0, lds, nullptr);
    __syncthreads();
    row_sweep<NT, false>(g, 0, lds, nullptr);
    __syncthreads();
  }
}

// One half step over ceil(lines / NT) workgroups; QLDS: q kept in LDS.
template <int NT, bool QLDS>
__global__ __launch_bounds__(NT) void adi_columns(AdiArgs g) {
  __shared__ __align__(16) double lds[AdiLds<NT>::ELEMS];
  __shared__ double qlds[QLDS ? kQRows * NT : 1];
  column_sweep<NT, QLDS>(g, static_cast<int>(blockIdx.x) * NT, lds, qlds);
}

template <int NT, bool QLDS>
__global__ __launch_bounds__(NT) void adi_rows(AdiArgs g) {
  __shared__ __align__(16) double lds[AdiLds<NT>::ELEMS];
  __shared__ double qlds[QLDS ? kQRows * NT : 1];
  row_sweep<NT, QLDS>(g, static_cast<int>(blockIdx.x) * NT, lds, qlds);
}

// Self-check of the short division: random (n, den) pairs, half of them with
// n / den next to a rounding midpoint (the cases a wrong rounding would
// show), compared bit for bit with the division operator wherever the
// no-rescale checks pass.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A double of random sign and mantissa with biased exponent in
// [e_lo, e_lo + e_span).
__device__ __forceinline__ double make_double(uint64_t bits, int e_lo, int e_span) {
  const uint64_t mant = bits & 0xFFFFFFFFFFFFFull;
  const uint64_t e = static_cast<uint64_t>(e_lo + static_cast<int>((bits >> 52) % e_span));
  const uint64_t sign = (bits >> 63) << 63;
  return __longlong_as_double(static_cast<long long>(sign | (e << 52) | mant));
}

__global__ __launch_bounds__(256) void adi_division_check(uint64_t seed, int64_t pairs,
                                                          unsigned long long* counts) {
  unsigned long long checked = 0, wrong = 0;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; k < pairs;
       k += static_cast<int64_t>(gridDim.x) * 256) {
    const uint64_t h0 = mix64(seed ^ (2 * static_cast<uint64_t>(k)));
    const uint64_t h1 = mix64(seed ^ (2 * static_cast<uint64_t>(k) + 1));
    const double den = make_double(h0, 1023 - 40, 80);
    double n;
    if (k & 1) {
      n = make_double(h1, 1023 - 300, 600);
    } else {
      // n = den * (q0 + half an ulp of q0), rounded once: n / den lies next
      // to the midpoint between q0 and its neighbour
      const double q0 = make_double(h1, 1023 - 200, 400);
      const double half = __longlong_as_double(
          static_cast<long long>(static_cast<uint64_t>(exponent_of(q0) - 53) << 52));
      n = __builtin_fma(den, q0, den * half);
    }
    if (!den_in_range(den)) continue;
    double lo, hi, m;
    quotient_bounds(fabs(den), lo, hi);
    const double q = short_division(n, den, division_reciprocal(den), m);
    if (!unscaled_division(m, lo, hi)) continue;
    ++checked;
    const double ieee = n / den;
    wrong += __double_as_longlong(q) != __double_as_longlong(ieee) ? 1 : 0;
  }
  atomicAdd(&counts[0], checked);
  atomicAdd(&counts[1], wrong);
}

// 0: the default (3), 1: every time step in one single-workgroup launch, 2:
// a launch per half step over 256-line workgroups, 3: the same over one-wave
// workgroups.  Measured on the MI355X, one wave per CU is the fastest: a
// line's chain leaves the SIMD idle most of the time, and waves sharing a CU
// queue behind each other's memory instructions.
std::atomic<int> g_adi_path{0};

template <int NT, bool QLDS>
int launch_halves(const AdiArgs& g, int nb, int64_t tsteps, hipStream_t stream) {
  for (int64_t s = 0; s < tsteps; ++s) {
    hipLaunchKernelGGL((adi_columns<NT, QLDS>), dim3(nb), dim3(NT), 0, stream, g);
    NPBX_LAUNCHED("adi_columns");
    hipLaunchKernelGGL((adi_rows<NT, QLDS>), dim3(nb), dim3(NT), 0, stream, g);
    NPBX_LAUNCHED("adi_rows");
  }
  return NPBX_OK;
}

// Launch shape and scratch layout of one call: v (the column sweep's output,
// (n + 16) rows of ldv), q ([step][line], qrows x ld) and the eight tables.
struct AdiPlan {
  int path, nt, nb;
  int64_t ld, ldv, qrows, tl, vel, qel, tel;
  size_t bytes() const { return static_cast<size_t>(vel + qel + tel) * sizeof(double); }
};

int plan_adi(int64_t n64, AdiPlan& pl) {
  const int n = static_cast<int>(n64);
  const int lines = n - 2;
  pl.path = g_adi_path.load();
  if (pl.path == 0) pl.path = 3;
  NPBX_REQUIRE(pl.path != 1 || lines <= 512,
               "adi: one-launch path holds at most 512 lines (N=%d)", n);
  pl.nt = pl.path == 1 ? (lines <= 256 ? 256 : 512) : pl.path == 3 ? 64 : 256;
  pl.nb = static_cast<int>(cdiv(lines, pl.nt));
  pl.ld = static_cast<int64_t>(pl.nb) * pl.nt;
  pl.ldv = pl.ld + 1;
  pl.qrows = cdiv(lines, kCW) * kCW + 1;
  pl.tl = pl.qrows + 1;  // table length: every chunk's block
  pl.vel = (n64 + 16) * pl.ldv;
  pl.qel = pl.qrows * pl.ld;
  pl.tel = 8 * pl.tl;
  return NPBX_OK;
}

// Runs the call with ws, pl.bytes() of device memory; what ws holds on entry
// does not reach u (the library workspace it normally is keeps whatever the
// previous call left).
int run_adi(double* u, int64_t n64, int64_t tsteps, const AdiPlan& pl, void* ws,
            hipStream_t stream) {
  const int n = static_cast<int>(n64);
  const int path = pl.path, nt = pl.nt, nb = pl.nb;
  const int64_t ld = pl.ld, ldv = pl.ldv, qrows = pl.qrows, tl = pl.tl;
  double* v = static_cast<double*>(ws);
  double* qs = v + pl.vel;
  double* tab = qs + pl.qel;
  const AdiCoef k = adi_coef(tsteps, n64);
  hipLaunchKernelGGL(adi_tables, dim3(1), dim3(64), 0, stream, k, n, static_cast<int>(tl), tab);
  NPBX_LAUNCHED("adi_tables");
  AdiArgs g;
  g.u = u;
  g.v = v;
  g.qs = qs;
  g.cols = {tab, tab + tl, tab + 2 * tl, tab + 3 * tl};
  g.rows = {tab + 4 * tl, tab + 5 * tl, tab + 6 * tl, tab + 7 * tl};
  g.nd = -k.d;
  g.c1 = 1.0 + 2.0 * k.d;
  g.f = k.f;
  g.a = k.a;
  g.na = -k.a;
  g.c2 = 1.0 + 2.0 * k.a;
  g.c = k.c;
  g.d = k.d;
  g.n = n;
  g.ld = static_cast<int>(ld);
  g.ldv = static_cast<int>(ldv);
  g.tsteps = static_cast<int>(tsteps);
  if (path == 1) {
    if (nt == 256)
      hipLaunchKernelGGL(adi_fused<256>, dim3(1), dim3(256), 0, stream, g);
    else
      hipLaunchKernelGGL(adi_fused<512>, dim3(1), dim3(512), 0, stream, g);
    NPBX_LAUNCHED("adi_fused");
    return NPBX_OK;
  }
  if (nt == 256) return launch_halves<256, false>(g, nb, tsteps, stream);
  if (qrows <= kQRows) return launch_halves<64, true>(g, nb, tsteps, stream);
  return launch_halves<64, false>(g, nb, tsteps, stream);
}

}  // namespace

int adi_path() { return g_adi_path.load(); }
void set_adi_path(int path) { g_adi_path.store(path); }

}  // namespace npbx

extern "C" int npbx_adi(int dtype, void* u, int64_t N, int64_t tsteps, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(dtype == NPBX_F64, "adi: dtype %s not supported (f64)", dtype_name(dtype));
  NPBX_REQUIRE(tsteps != 0, "adi: TSTEPS=0 (the reference divides by it)");
  NPBX_REQUIRE(N > 0, "adi: N=%lld must be positive", static_cast<long long>(N));
  NPBX_REQUIRE(N <= 32768 && tsteps <= (int64_t(1) << 30), "adi: N or TSTEPS out of range");
  if (N < 3 || tsteps < 0) return NPBX_OK;  // no interior line, or no time step
  NPBX_REQUIRE(u, "adi: null array");
  AdiPlan pl;
  NPBX_TRY(plan_adi(N, pl));
  void* ws = nullptr;
  NPBX_TRY(workspace(pl.bytes(), &ws));
  return run_adi(static_cast<double*>(u), N, tsteps, pl, ws, as_stream(stream));
}

// ---- allocation-free entry point ---------------------------------------------
// npbx_adi_noalloc runs npbx_adi's kernels with their scratch in this
// file-scope device array instead of the library workspace, so it never
// allocates.  Every GPU the library runs it on holds its own copy (16 MiB)
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: kernels get the device address resolved per device by
// hipGetSymbolAddress (adi_scratch_ptr), never the array's name.
namespace npbx {
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
// The largest N accepted: its scratch (v, q and the tables, AdiPlan) is
// 16,661,568 bytes on every launch path that takes N=1000.
constexpr int64_t kNoAllocMaxN = 1000;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch`, resolved once per device with
// that device current.
int adi_scratch_ptr(void** out) {
  int dev = 0;
  NPBX_HIP(hipGetDevice(&dev));
  NPBX_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    NPBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NPBX_REQUIRE(p != nullptr, "adi_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NPBX_OK;
}
}  // namespace

int adi_noalloc_setup() {
  void* p = nullptr;
  return adi_scratch_ptr(&p);
}
}  // namespace npbx

extern "C" int npbx_adi_noalloc(int dtype, void* u, int64_t N, int64_t tsteps, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(dtype == NPBX_F64, "adi_noalloc: dtype %s not supported (f64)",
               dtype_name(dtype));
  NPBX_REQUIRE(tsteps != 0, "adi_noalloc: TSTEPS=0 (the reference divides by it)");
  NPBX_REQUIRE(N > 0, "adi_noalloc: N=%lld must be positive", static_cast<long long>(N));
  NPBX_REQUIRE(N <= 32768 && tsteps <= (int64_t(1) << 30),
               "adi_noalloc: N or TSTEPS out of range");
  if (N < 3 || tsteps < 0) return NPBX_OK;
  NPBX_REQUIRE(u, "adi_noalloc: null array");
  NPBX_REQUIRE(N <= kNoAllocMaxN,
               "adi_noalloc: N=%lld exceeds %lld, the most its fixed %zu-byte scratch holds; "
               "use npbx_adi",
               static_cast<long long>(N), static_cast<long long>(kNoAllocMaxN),
               kNoAllocScratchBytes);
  AdiPlan pl;
  NPBX_TRY(plan_adi(N, pl));
  NPBX_REQUIRE(pl.bytes() <= kNoAllocScratchBytes,
               "adi_noalloc: N=%lld needs %zu bytes of scratch, more than its fixed %zu",
               static_cast<long long>(N), pl.bytes(), kNoAllocScratchBytes);
  void* ws = nullptr;
  NPBX_TRY(adi_scratch_ptr(&ws));
  return run_adi(static_cast<double*>(u), N, tsteps, pl, ws, as_stream(stream));
}

extern "C" int npbx_adi_division_check(int64_t pairs, uint64_t seed, int64_t* checked,
                                       int64_t* wrong) {
  using namespace npbx;
  NPBX_REQUIRE(pairs > 0 && pairs <= (int64_t(1) << 34) && checked && wrong,
               "adi_division_check: bad arguments");
  unsigned long long* counts = nullptr;
  NPBX_HIP(hipMalloc(reinterpret_cast<void**>(&counts), 2 * sizeof(unsigned long long)));
  unsigned long long host[2] = {0, 0};
  int rc = NPBX_OK;
  if (hipMemcpy(counts, host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(NPBX_EHIP, "adi_division_check: copy to the device failed");
  } else {
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(cdiv(pairs, 256), 8192));
    hipLaunchKernelGGL(adi_division_check, dim3(blocks), dim3(256), 0, nullptr, seed, pairs,
                       counts);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(host, counts, sizeof(host), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(NPBX_EHIP, "adi_division_check: the check kernel failed");
    } else {
      *checked = static_cast<int64_t>(host[0]);
      *wrong = static_cast<int64_t>(host[1]);
    }
  }
  (void)hipFree(counts);
  return rc;
}
