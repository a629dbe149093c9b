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
#include <type_traits>

#include "common.h"

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

// Tables of one sweep, n entries each (entry j for step j = 1 .. n-2).
struct SweepTab {
  const double* p;    // p[j]
  const double* den;  // the forward denominators, lo p[j-1] + mid
  const double* y;    // the division sequence's reciprocal of den[j]
  const double* ok;   // ok[0] = 1 when every den[j] is in den_in_range;
                      // ok[1], ok[2]: numerators n with ok[1] <= |n| < ok[2]
                      // never make the sequence rescale, for any den[j]
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

// The reciprocal that the compiler's IEEE FP64 division sequence computes
// from an unscaled denominator: the hardware estimate refined by two Newton
// steps.  With it the sequence ends m = n y, r = fma(-den, m, n),
// q = fma(r, y, m) (its div_fmas and div_fixup steps return their operand
// when nothing was rescaled and the result is a normal number).
__device__ __forceinline__ double division_reciprocal(double den) {
  double y = __builtin_amdgcn_rcp(den);
  double e = __builtin_fma(-den, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-den, y, 1.0);
  return __builtin_fma(y, e, y);
}

__device__ __forceinline__ double short_division(double n, double den, double y) {
  const double m = n * y;
  const double r = __builtin_fma(-den, m, n);
  return __builtin_fma(r, y, m);
}

__device__ __forceinline__ int exponent_of(double x) {
  return (__double2hiint(x) >> 20) & 0x7ff;
}

// den's side of the no-rescale condition: a normal number whose reciprocal
// is normal, with room for the numerator bounds below.
__device__ __forceinline__ bool den_in_range(double den) {
  const int e = exponent_of(den);
  return e >= 100 && e <= 1900;
}

// The numerator's side, for dens in range between dmin and dmax in
// magnitude: the division sequence rescales when n is zero, tiny (biased
// exponent <= 53), not finite, at least 2^768 times den (in exponent), or
// when the quotient is subnormal.  lo <= |n| < hi excludes all of them:
//   lo = max(2^-960, 2^-1020 dmax), hi = 2^766 dmin
// (a NaN fails both comparisons; hi may be infinite when dmin is large,
// where no finite n reaches the exponent gap).
__device__ __forceinline__ void numerator_bounds(double dmin, double dmax, double& lo,
                                                 double& hi) {
  lo = fmax(0x1p-960, 0x1p-1020 * dmax);
  hi = 0x1p766 * dmin;
}

__device__ __forceinline__ bool unscaled_division(double n, double lo, double hi) {
  const double a = fabs(n);
  return a >= lo && a < hi;
}

// p[j] = -c / (a p[j-1] + b) with p[0] = 0 (adi_numpy.py:28,30) and
// p[j] = -f / (d p[j-1] + e) (:38,40), for j = 1 .. n-2 (lane 0: the column
// sweep's table, lane 1: the row sweep's), with each den's reciprocal.
__global__ void adi_tables(AdiCoef k, int n, double* __restrict__ tab) {
  const int lane = threadIdx.x;
  if (lane > 1) return;
  const double lo = lane == 0 ? k.a : k.d;
  const double mid = lane == 0 ? k.b : k.e;
  const double hi = lane == 0 ? k.c : k.f;
  double* p = tab + (lane * 4) * n;
  double* den = p + n;
  double* y = den + n;
  double* ok = y + n;
  double prev = 0.0, dmin = INFINITY, dmax = 0.0;
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
    dmax = fmax(dmax, fabs(dj));
  }
  ok[0] = in_range ? 1.0 : 0.0;
  numerator_bounds(dmin, dmax, ok[1], ok[2]);
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

// Entries j0 .. j0+kCW-1 of a table (indices past n-2 read entry n-2).
__device__ __forceinline__ void table_chunk(ConstTab tab, int n, int j0, double (&d)[kCW]) {
#pragma unroll
  for (int kk = 0; kk < kCW; ++kk) d[kk] = tab[min(j0 + kk, n - 2)];
}

// Row j of a [row][line] scratch array: a wave-uniform base, so that the
// lanes' accesses are one scalar base plus a 32-bit lane offset.
__device__ __forceinline__ double* row_of(double* base, int j, int pitch) {
  return base + static_cast<uint32_t>(j * pitch);
}

// Hides the thread index, the side and the line base from loop-invariant code
// motion: the one-launch kernel runs both sweeps in a time-step loop, and
// addresses hoisted out of it would hold both sweeps' registers at once.
__device__ __forceinline__ void opaque(int& t, int& n, int& lb) {
  asm volatile("" : "+v"(t));
  asm volatile("" : "+s"(n), "+s"(lb));
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
template <class Rhs>
__device__ __forceinline__ void forward_chunk(double& q, const Fwd& f, const double* dy, int n,
                                              int j0, bool live, double* qs, int ld, int col,
                                              Rhs rhs) {
  const double q0 = q;
  bool redo = !f.fast;
  if (f.fast) {
    bool bad = false;
#pragma unroll
    for (int jj = 0; jj < kCW; ++jj) {
      const double2 d = *reinterpret_cast<const double2*>(dy + 2 * jj);
      const double num = rhs(jj) - f.w * q;
      q = short_division(num, d.x, d.y);
      row_of(qs, j0 + jj, ld)[col] = q;
      bad = bad || (live && j0 + jj <= n - 2 && !unscaled_division(num, f.lo, f.hi));
    }
    redo = __any(bad);
  }
  if (redo) {
    q = q0;
#pragma unroll
    for (int jj = 0; jj < kCW; ++jj) {
      q = (rhs(jj) - f.w * q) / dy[2 * jj];
      row_of(qs, j0 + jj, ld)[col] = q;
    }
  }
}

// The table pairs of chunk j0 for the LDS buffer: threads 32..63 each load
// one entry (den of step j0 + (t-32), then y), clamped to step n-2.
__device__ __forceinline__ double fetch_pair(const SweepTab& tab, int t, int n, int j0) {
  const int e = t - 32;
  const int j = min(j0 + (e & (kCW - 1)), n - 2);
  return (e < kCW ? tab.den : tab.y)[j];
}
__device__ __forceinline__ void commit_pair(double* dy, int t, double val) {
  const int e = t - 32;
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
__device__ __forceinline__ void q_chunk(double* qs, int ld, int col, int j0, double (&q)[kCW]) {
#pragma unroll
  for (int kk = 0; kk < kCW; ++kk) q[kk] = row_of(qs, j0 + kk, ld)[col];
}

// Column sweep (adi_numpy.py:27-34) of the lines lb+1 .. lb+NT: column i
// of u solved along the rows into column i of v.
template <int NT>
__device__ __forceinline__ void column_sweep(const AdiArgs& g, int lb, double* lds) {
  using L = AdiLds<NT>;
  constexpr int W = NT + 2;  // tile columns lb .. lb+NT+1
  int t = threadIdx.x, n = g.n;
  opaque(t, n, lb);
  const int i = lb + 1 + t;
  const bool live = i <= n - 2;
  const int nchunks = (n - 2 + kCW - 1) / kCW;
  const Fwd f = forward_consts(g.cols, g.a);
  // Tile of rows j0 .. j0+kCW-1 of u, columns lb .. lb+W-1, tile[row][col]:
  // thread t loads column lb+t of every row, threads 0..31 also the last two
  // columns (row j0 + t/2, column lb+NT + t%2).  Rows and columns are
  // clamped into the grid: what a clamped load reads only reaches steps past
  // n-2 and lanes past the last line.
  const int col = min(lb + t, n - 1);
  const int xr = t >> 1, xcol = min(lb + NT + (t & 1), n - 1);
  double pre[kCW], prex = 0.0, prey = 0.0;
  auto fetch = [&](int c) {
    const int j0 = 1 + c * kCW;
#pragma unroll
    for (int m = 0; m < kCW; ++m) pre[m] = g.u[static_cast<int64_t>(min(j0 + m, n - 2)) * n + col];
    if (t < 2 * kCW) prex = g.u[static_cast<int64_t>(min(j0 + xr, n - 2)) * n + xcol];
    else if (t < 4 * kCW) prey = fetch_pair(g.cols, t, n, j0);
  };
  double q = 1.0;  // q[0] = v[0, i] = 1 (:29-31)
  fetch(0);
  for (int c = 0; c < nchunks; ++c) {
    const int j0 = 1 + c * kCW;
    double* buf = lds + (c & 1) * L::BUF;
    double* dy = buf + L::TILE;
#pragma unroll
    for (int m = 0; m < kCW; ++m) buf[m * W + t] = pre[m];
    if (t < 2 * kCW) buf[xr * W + NT + (t & 1)] = prex;
    commit_pair(dy, t, prey);
    __syncthreads();
    if (c + 1 < nchunks) fetch(c + 1);
    forward_chunk(q, f, dy, n, j0, live, g.qs, g.ld, lb + t, [&](int jj) {
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
  double qn[kCW];
  const ConstTab ptab = as_const(g.cols.p);
  q_chunk(g.qs, g.ld, lb + t, 1 + (nchunks - 1) * kCW, qn);
  auto chunk = [&](int c, auto partial) {
    const int j0 = 1 + c * kCW;
    double qc[kCW], pj[kCW];
#pragma unroll
    for (int kk = 0; kk < kCW; ++kk) qc[kk] = qn[kk];
    table_chunk(ptab, n, j0, pj);
    if (c > 0) q_chunk(g.qs, g.ld, lb + t, j0 - kCW, qn);
    back_chunk<decltype(partial)::value>(x, qc, pj, n, j0, [&](int kk, double xv) {
      const int j = j0 + kk;
      row_of(g.v, decltype(partial)::value && j > n - 2 ? j + 1 : j, g.ldv)[i] = xv;
    });
  };
  chunk(nchunks - 1, std::true_type{});
  for (int c = nchunks - 2; c >= 0; --c) chunk(c, std::false_type{});
}

// Row sweep (adi_numpy.py:36-46) of the lines lb+1 .. lb+NT: row i of v
// solved along the columns into row i of u.
template <int NT>
__device__ __forceinline__ void row_sweep(const AdiArgs& g, int lb, double* lds) {
  using L = AdiLds<NT>;
  constexpr int P = kCW + 1;    // LDS row pitch (odd: lanes hit different banks)
  constexpr int RS = NT / kCW;  // tile rows one pass of the workgroup covers
  int t = threadIdx.x, n = g.n;
  opaque(t, n, lb);
  const int i = lb + 1 + t;
  const bool live = i <= n - 2;
  const int nchunks = (n - 2 + kCW - 1) / kCW;
  const Fwd f = forward_consts(g.rows, g.d);
  // Tile of columns j0 .. j0+kCW-1 of v, rows lb .. lb+NT+1, tile[row][col]:
  // kCW consecutive threads read one row's run (thread t: column j0 + t%kCW
  // of rows lb + t/kCW + m*RS), threads 0..31 also the last two rows; rows
  // and columns clamped into the grid as in the column sweep.
  const int tj = t % kCW, tr = t / kCW;
  double pre[kCW], prex = 0.0, prey = 0.0;
  auto fetch = [&](int c) {
    const int colj = min(1 + c * kCW + tj, n - 2);
#pragma unroll
    for (int m = 0; m < kCW; ++m) pre[m] = row_of(g.v, min(lb + tr + m * RS, n - 1), g.ldv)[colj];
    if (t < 2 * kCW) prex = row_of(g.v, min(lb + NT + tr, n - 1), g.ldv)[colj];
    else if (t < 4 * kCW) prey = fetch_pair(g.rows, t, n, 1 + c * kCW);
  };
  double q = 1.0;  // q[0] = u[i, 0] = 1 (:37-39)
  fetch(0);
  for (int c = 0; c < nchunks; ++c) {
    const int j0 = 1 + c * kCW;
    double* buf = lds + (c & 1) * L::BUF;
    double* dy = buf + L::TILE;
#pragma unroll
    for (int m = 0; m < kCW; ++m) buf[(tr + m * RS) * P + tj] = pre[m];
    if (t < 2 * kCW) buf[(NT + tr) * P + tj] = prex;
    commit_pair(dy, t, prey);
    __syncthreads();
    if (c + 1 < nchunks) fetch(c + 1);
    forward_chunk(q, f, dy, n, j0, live, g.qs, g.ld, lb + t, [&](int jj) {
      const double* r = buf + t * P + jj;
      return (g.na * r[0] + g.c2 * r[P]) - g.c * r[2 * P];
    });
  }
  __syncthreads();  // the tiles are reused for the output below
  // back substitution (:44-46) with u[i, n-1] = 1; u[i, 0] = 1 (:37)
  if (live) {
    g.u[static_cast<int64_t>(i) * n] = 1.0;
    g.u[static_cast<int64_t>(i) * n + n - 1] = 1.0;
  }
  double x = 1.0;
  double qn[kCW];
  const ConstTab ptab = as_const(g.rows.p);
  q_chunk(g.qs, g.ld, lb + t, 1 + (nchunks - 1) * kCW, qn);
  auto chunk = [&](int c, auto partial) {
    const int j0 = 1 + c * kCW;
    double qc[kCW], pj[kCW];
#pragma unroll
    for (int kk = 0; kk < kCW; ++kk) qc[kk] = qn[kk];
    table_chunk(ptab, n, j0, pj);
    if (c > 0) q_chunk(g.qs, g.ld, lb + t, j0 - kCW, qn);
    double* ob = lds + (c & 1) * L::BUF;  // ob[line][column - j0]
    back_chunk<decltype(partial)::value>(x, qc, pj, n, j0,
                                         [&](int kk, double xv) { ob[t * P + kk] = xv; });
    __syncthreads();
    // rows lb+1 .. lb+NT, columns j0 .. j0+kCW-1 of u, kCW threads per row run
    const int colj = j0 + tj;
    if (colj <= n - 2) {
#pragma unroll
      for (int m = 0; m < kCW; ++m) {
        const int rr = tr + m * RS;
        if (lb + 1 + rr <= n - 2)
          g.u[static_cast<int64_t>(lb + 1 + rr) * n + colj] = ob[rr * P + tj];
      }
    }
  };
  chunk(nchunks - 1, std::true_type{});
  for (int c = nchunks - 2; c >= 0; --c) chunk(c, std::false_type{});
}

template <int NT>
__global__ __launch_bounds__(NT) void adi_fused(AdiArgs g) {
  __shared__ __align__(16) double lds[AdiLds<NT>::ELEMS];
  for (int s = 0; s < g.tsteps; ++s) {
    column_sweep<NT>(g, 0, lds);
    __syncthreads();
    row_sweep<NT>(g, 0, lds);
    __syncthreads();
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void adi_columns(AdiArgs g) {
  __shared__ __align__(16) double lds[AdiLds<NT>::ELEMS];
  column_sweep<NT>(g, static_cast<int>(blockIdx.x) * NT, lds);
}

template <int NT>
__global__ __launch_bounds__(NT) void adi_rows(AdiArgs g) {
  __shared__ __align__(16) double lds[AdiLds<NT>::ELEMS];
  row_sweep<NT>(g, static_cast<int>(blockIdx.x) * NT, lds);
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
    double lo, hi;
    numerator_bounds(fabs(den), fabs(den), lo, hi);
    if (!unscaled_division(n, lo, hi)) continue;
    const double q = short_division(n, den, division_reciprocal(den));
    ++checked;
    const double ieee = n / den;
    wrong += __double_as_longlong(q) != __double_as_longlong(ieee) ? 1 : 0;
  }
  atomicAdd(&counts[0], checked);
  atomicAdd(&counts[1], wrong);
}

constexpr int kSplitNT = 256;
std::atomic<int> g_adi_path{0};  // 0: by size, 1: one launch, 2: a launch per half step

int run_adi(double* u, int64_t n64, int64_t tsteps, hipStream_t stream) {
  const int n = static_cast<int>(n64);
  const int lines = n - 2;
  int path = g_adi_path.load();
  if (path == 0) path = lines <= 256 ? 1 : 2;
  NPBX_REQUIRE(path == 2 || lines <= 512,
               "adi: one-launch path holds at most 512 lines (N=%d)", n);
  const int nt = path == 1 ? (lines <= 256 ? 256 : 512) : kSplitNT;
  const int nb = static_cast<int>(cdiv(lines, nt));
  const int64_t ld = static_cast<int64_t>(nb) * nt;
  const int64_t ldv = ld + 1;
  const int64_t qrows = cdiv(lines, kCW) * kCW + 1;
  const int64_t vel = (n64 + 16) * ldv, qel = qrows * ld, tel = 8 * n64;
  void* ws = nullptr;
  NPBX_TRY(workspace(static_cast<size_t>(vel + qel + tel) * sizeof(double), &ws));
  double* v = static_cast<double*>(ws);
  double* qs = v + vel;
  double* tab = qs + qel;
  const AdiCoef k = adi_coef(tsteps, n64);
  hipLaunchKernelGGL(adi_tables, dim3(1), dim3(64), 0, stream, k, n, tab);
  NPBX_LAUNCHED("adi_tables");
  AdiArgs g;
  g.u = u;
  g.v = v;
  g.qs = qs;
  g.cols = {tab, tab + n, tab + 2 * n, tab + 3 * n};
  g.rows = {tab + 4 * n, tab + 5 * n, tab + 6 * n, tab + 7 * n};
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
  for (int64_t s = 0; s < tsteps; ++s) {
    hipLaunchKernelGGL(adi_columns<kSplitNT>, dim3(nb), dim3(kSplitNT), 0, stream, g);
    NPBX_LAUNCHED("adi_columns");
    hipLaunchKernelGGL(adi_rows<kSplitNT>, dim3(nb), dim3(kSplitNT), 0, stream, g);
    NPBX_LAUNCHED("adi_rows");
  }
  return NPBX_OK;
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
  return run_adi(static_cast<double*>(u), N, tsteps, as_stream(stream));
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
