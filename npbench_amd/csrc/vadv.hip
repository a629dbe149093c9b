// Implicit vertical advection (vadv_numpy.py:9-78): one tridiagonal system
// per (i, j) column of K levels, solved by the Thomas algorithm in FP64.
//
// Layout: utens_stage, u_stage, u_pos, utens (I, J, K); wcon (I+1, J, K); C
// order, so column c = i*J + j holds its K values contiguously at c*K, and
// wcon's rows i and i+1 for that column sit at c*K and (c+J)*K.  Field
// storage is FP64 or FP32; the coefficients, the forward sweep and the back
// substitution are FP64 in both cases, in the reference's order of operations
// (reciprocal of the pivot, then multiply), so the FP64 path is bitwise equal
// to NumPy.
//
// Kernel: one wave (64 lanes) owns 64 consecutive columns, lane = column, so
// the 64 columns' data is one contiguous block of 64*K elements per field.
// The k axis is walked in chunks of KC = 16 levels: the wave loads a chunk of
// all six streams with lanes running along k (16-byte loads, 4 or 8 lanes per
// column segment, coalesced), forms wcon's column sum and the right-hand-side
// base there, stores those and u_stage transposed into LDS, and each lane then
// reads its own column from LDS.  The loads of chunk n+1 are issued before
// chunk n is solved, so HBM requests stay in flight during the dependent FP64
// pivot chain.  c'_k and d'_k go to a library workspace laid out [k][column]
// (coalesced 512-byte rows); the back substitution reads them in reverse, one
// chunk ahead, and stages its output through LDS for a coalesced store.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "common.h"
#include "division.h"
#include "vecio.h"

namespace npbx {
namespace {

constexpr double kBetM = 0.5;  // vadv_numpy.py:4
constexpr double kBetP = 0.5;  // vadv_numpy.py:5
constexpr int kWave = 64;
constexpr int kPad = kWave + 1;  // LDS row pitch (elements), breaks bank conflicts

// Levels per chunk; one column's chunk is 128 bytes (FP64) or 64 bytes (FP32).
constexpr int KC = 16;

// Element-granular access for a column-contiguous stream: W elements per
// load (16 bytes when VEC, else 1).  Load g = lane + 64*m of a chunk covers
// elements e = g*W .. g*W+W-1, i.e. column e / KC, levels e % KC .., so
// KC/W consecutive lanes read one column's segment.
template <class S, bool VEC>
struct Lanes {
  static constexpr int W = VEC ? 16 / static_cast<int>(sizeof(S)) : 1;
  static constexpr int LOADS = KC / W;  // loads per lane per stream per chunk
};

template <class S, bool VEC>
__device__ __forceinline__ void load_chunk(const S* __restrict__ p, int64_t c0, int ncol, int K,
                                           int k0, S (&r)[KC]) {
  using L = Lanes<S, VEC>;
  const int lane = threadIdx.x;
#pragma unroll
  for (int m = 0; m < L::LOADS; ++m) {
    const int e = (lane + kWave * m) * L::W;
    const int col = e / KC, kk = e % KC;
    const bool ok = col < ncol && k0 + kk < K;
    const S* q = p + (c0 + col) * K + k0 + kk;
    if constexpr (VEC) {
      uint4 raw = make_uint4(0, 0, 0, 0);
      if (ok) raw = *reinterpret_cast<const uint4*>(q);
      __builtin_memcpy(&r[m * L::W], &raw, 16);
    } else {
      r[m] = ok ? *q : S(0);
    }
  }
}

// LDS slot of element index e of a chunk: row = level, column = column.
__device__ __forceinline__ int slot(int e) { return (e % KC) * kPad + e / KC; }

template <class S, bool VEC>
__global__ __launch_bounds__(kWave) void vadv_wave(S* __restrict__ utens_stage,
                                                   const S* __restrict__ u_stage,
                                                   const S* __restrict__ wcon,
                                                   const S* __restrict__ u_pos,
                                                   const S* __restrict__ utens, double dtr, int J,
                                                   int K, int64_t ncols, double* __restrict__ cp,
                                                   double* __restrict__ dp) {
  using L = Lanes<S, VEC>;
  // Staged per chunk: s_k = wcon[i+1] + wcon[i], the d right-hand side base
  // ((dtr*u_pos + utens) + utens_stage) formed by the loading lane in the
  // reference's order, and u_stage.  The back substitution reuses base_lds
  // for u_pos and u_lds (as double) for its output.
  __shared__ double s_lds[KC * kPad];
  __shared__ double base_lds[KC * kPad];
  __shared__ double u_lds[KC * kPad];
  const int lane = threadIdx.x;
  // XCD-contiguous numbering: the waves of columns (i, j) and (i+1, j), which
  // both read wcon's row i+1, run on one XCD at about the same time and share
  // its L2.
  const int64_t c0 = static_cast<int64_t>(xcd_tile(blockIdx.x, gridDim.x)) * kWave;
  const int ncol = static_cast<int>(std::min<int64_t>(kWave, ncols - c0));
  const bool active = lane < ncol;
  const int64_t col = c0 + lane;
  const S* w_hi = wcon + static_cast<int64_t>(J) * K;
  const int nchunks = (K + KC - 1) / KC;

  // ---- forward sweep -------------------------------------------------------
  // Step m needs u[m+1] and s[m+1]; it is completed when level m+1 arrives.
  double u_prev = 0, u_cur = 0, s_cur = 0, base_cur = 0, c_prev = 0, d_prev = 0;
  S r_us[KC], r_u[KC], r_up[KC], r_ut[KC], r_wl[KC], r_wh[KC];
  load_chunk<S, VEC>(utens_stage, c0, ncol, K, 0, r_us);
  load_chunk<S, VEC>(u_stage, c0, ncol, K, 0, r_u);
  load_chunk<S, VEC>(u_pos, c0, ncol, K, 0, r_up);
  load_chunk<S, VEC>(utens, c0, ncol, K, 0, r_ut);
  load_chunk<S, VEC>(wcon, c0, ncol, K, 0, r_wl);
  load_chunk<S, VEC>(w_hi, c0, ncol, K, 0, r_wh);
  for (int n = 0; n < nchunks; ++n) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      const int e = (lane + kWave * (q / L::W)) * L::W + q % L::W;
      const int at = slot(e);
      s_lds[at] = static_cast<double>(r_wh[q]) + static_cast<double>(r_wl[q]);
      base_lds[at] = (dtr * static_cast<double>(r_up[q]) + static_cast<double>(r_ut[q])) +
                     static_cast<double>(r_us[q]);
      u_lds[at] = static_cast<double>(r_u[q]);
    }
    __syncthreads();
    if (n + 1 < nchunks) {
      const int k1 = (n + 1) * KC;
      load_chunk<S, VEC>(utens_stage, c0, ncol, K, k1, r_us);
      load_chunk<S, VEC>(u_stage, c0, ncol, K, k1, r_u);
      load_chunk<S, VEC>(u_pos, c0, ncol, K, k1, r_up);
      load_chunk<S, VEC>(utens, c0, ncol, K, k1, r_ut);
      load_chunk<S, VEC>(wcon, c0, ncol, K, k1, r_wl);
      load_chunk<S, VEC>(w_hi, c0, ncol, K, k1, r_wh);
    }
    const int k0 = n * KC;
    const int kn = std::min(KC, K - k0);
    for (int kk = 0; kk < kn; ++kk) {
      const int k = k0 + kk;
      const double u_next = u_lds[kk * kPad + lane];
      const double s_next = s_lds[kk * kPad + lane];
      const double base_next = base_lds[kk * kPad + lane];
      if (k == 1) {
        // level 0 (vadv_numpy.py:15-30)
        const double gcv = 0.25 * s_next;
        const double cs = gcv * kBetM;
        const double ck = gcv * kBetP;
        const double bcol = dtr - ck;
        const double corr = -cs * (u_next - u_cur);
        const double d = base_cur + corr;
        const double inv = 1.0 / bcol;
        c_prev = ck * inv;
        d_prev = d * inv;
      } else if (k >= 2) {
        // level m = k-1 (vadv_numpy.py:32-53)
        const double gav = -0.25 * s_cur;
        const double gcv = 0.25 * s_next;
        const double as_ = gav * kBetM;
        const double cs = gcv * kBetM;
        const double acol = gav * kBetP;
        const double ck = gcv * kBetP;
        const double bcol = dtr - acol - ck;
        const double corr = -as_ * (u_prev - u_cur) - cs * (u_next - u_cur);
        const double d = base_cur + corr;
        const double inv = 1.0 / (bcol - c_prev * acol);
        c_prev = ck * inv;
        d_prev = (d - d_prev * acol) * inv;
      }
      if (k >= 1 && active) {
        cp[static_cast<int64_t>(k - 1) * ncols + col] = c_prev;
        dp[static_cast<int64_t>(k - 1) * ncols + col] = d_prev;
      }
      u_prev = u_cur;
      u_cur = u_next;
      s_cur = s_next;
      base_cur = base_next;
    }
  }
  // level K-1 (vadv_numpy.py:55-68)
  {
    const double gav = -0.25 * s_cur;
    const double as_ = gav * kBetM;
    const double acol = gav * kBetP;
    const double bcol = dtr - acol;
    const double corr = -as_ * (u_prev - u_cur);
    const double d = base_cur + corr;
    const double inv = 1.0 / (bcol - c_prev * acol);
    d_prev = (d - d_prev * acol) * inv;
  }

  // ---- back substitution (vadv_numpy.py:70-78) -----------------------------
  // Chunk n-1's u_pos, c' and d' are fetched while chunk n is solved.
  double x = d_prev;
  double* up_lds = base_lds;
  double* out_lds = u_lds;
  S r_p[KC];
  double cpk[KC], dpk[KC];
  auto fetch = [&](int n) {
    const int k0 = n * KC;
    load_chunk<S, VEC>(u_pos, c0, ncol, K, k0, r_p);
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      const int k = k0 + kk;
      const bool ok = active && k < K - 1;
      cpk[kk] = ok ? cp[static_cast<int64_t>(k) * ncols + col] : 0.0;
      dpk[kk] = ok ? dp[static_cast<int64_t>(k) * ncols + col] : 0.0;
    }
  };
  fetch(nchunks - 1);
  for (int n = nchunks - 1; n >= 0; --n) {
    const int k0 = n * KC;
    const int kn = std::min(KC, K - k0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      const int e = (lane + kWave * (q / L::W)) * L::W + q % L::W;
      up_lds[slot(e)] = static_cast<double>(r_p[q]);
    }
    double c_here[KC], d_here[KC];
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      c_here[kk] = cpk[kk];
      d_here[kk] = dpk[kk];
    }
    __syncthreads();
    if (n > 0) fetch(n - 1);
#pragma unroll
    for (int kk = KC - 1; kk >= 0; --kk) {
      if (kk >= kn) continue;
      const int k = k0 + kk;
      if (k < K - 1) x = d_here[kk] - c_here[kk] * x;
      out_lds[kk * kPad + lane] = dtr * (x - up_lds[kk * kPad + lane]);
    }
    __syncthreads();
    // coalesced store of the chunk: the inverse of the load mapping
#pragma unroll
    for (int m = 0; m < L::LOADS; ++m) {
      const int e = (lane + kWave * m) * L::W;
      const int c = e / KC, kk = e % KC;
      if (c < ncol && kk < kn) {
        S* o = utens_stage + (c0 + c) * K + k0 + kk;
#pragma unroll
        for (int w = 0; w < L::W; ++w) o[w] = static_cast<S>(out_lds[slot(e + w)]);
      }
    }
  }
}

// ---- on-chip elimination (K <= 160) ---------------------------------------
// A workgroup of two waves owns 64 consecutive columns, lane = column, and
// keeps its columns' c'_k and d'_k in LDS ([level][lane], 16 bytes per level
// and column: K = 160 fills the 160 KiB of a CU, K <= 80 lets two workgroups
// share a CU, K <= 48 three), so the elimination never leaves the chip: HBM
// sees the five input streams once, u_pos a second time in the back
// substitution, and the output once.  The forward sweep is split between
// the waves along its two recurrences:
//   wave 0 (the pivots): from wcon alone, den_k = b_k - c'_{k-1} a_k, its
//     reciprocal inv_k and c'_k = c_k inv_k -- the serial chain through the
//     division -- writing c'_k and inv_k to LDS;
//   wave 1 (the right-hand sides): d_k from all five fields, then
//     d'_k = (d_k - d'_{k-1} a_k) inv_k with inv_k read from LDS, written
//     over inv_k; then the back substitution.
// Wave 0 runs one chunk of L levels ahead, a barrier between chunks.  Each
// lane loads its own column, 16 bytes at a time, 64 bytes per stream and
// chunk, two chunks ahead; every load is issued whether or not its chunk
// exists (past the last chunk it rereads the column's last vector), so the
// waits before a chunk is converted leave the next chunk's loads in flight,
// and middle chunks run in pairs of straight-line code with no branch
// between levels.  The arithmetic is vadv_wave's, operation for operation
// (both waves form a_k, c_k and the coefficients from the same sums).
// 1 / den with the bits of the division operator: the compiler's IEEE
// division sequence restated without its scaling steps (division_reciprocal,
// then r = fma(-den, y, 1), q = fma(r, y, y)) where that sequence does not
// rescale (den in range and 1/den below 2^765); a lane outside that range
// takes the division operator, behind a wave-uniform branch so that the
// common path does not also compute the full sequence.
__device__ __forceinline__ double pivot_reciprocal(double den) {
  const double y = division_reciprocal(den);
  const double r = __builtin_fma(-den, y, 1.0);
  double q = __builtin_fma(r, y, y);
  const bool ok = den_in_range(den) && fabs(y) < 0x1p765;
  if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
    double d = den;
    asm volatile("" : "+v"(d));  // keeps the division inside the branch
    if (!ok) q = 1.0 / d;
  }
  return q;
}

// The short form alone, keeping the smallest and largest |den| seen: where
// 2^-764 <= |den| < 2^878 for every den of a chunk (reciprocal_range_ok),
// the division sequence would not have rescaled (den's exponent within
// [100, 1900], 1/den below 2^765 and normal), so the short form has its
// bits; otherwise the caller redoes the chunk with the division operator.
__device__ __forceinline__ double short_reciprocal(double den, double& lo, double& hi) {
  const double y = division_reciprocal(den);
  const double r = __builtin_fma(-den, y, 1.0);
  lo = fmin(lo, fabs(den));
  hi = fmax(hi, fabs(den));
  return __builtin_fma(r, y, y);
}
__device__ __forceinline__ bool reciprocal_range_ok(double lo, double hi) {
  return lo >= 0x1p-764 && hi < 0x1p878;
}

// A 16-byte register quad (a vector type, so arrays of them stay in
// registers).
typedef uint32_t q32x4 __attribute__((ext_vector_type(4)));

template <class S>
struct PairShape {
  static constexpr int L = 64 / static_cast<int>(sizeof(S));  // levels per chunk
  static constexpr int V = 16 / static_cast<int>(sizeof(S));  // elements per 16-byte load
};

template <class S, int NLEV>
__global__ __launch_bounds__(2 * kWave) void vadv_pair(S* __restrict__ utens_stage,
                                                       const S* __restrict__ u_stage,
                                                       const S* __restrict__ wcon,
                                                       const S* __restrict__ u_pos,
                                                       const S* __restrict__ utens, double dtr,
                                                       int J, int K, int64_t ncols) {
  constexpr int L = PairShape<S>::L, V = PairShape<S>::V;
  static_assert(NLEV % L == 0, "whole chunks of LDS levels");
  __shared__ double cp[NLEV * kWave];
  __shared__ double dp[NLEV * kWave];  // inv_k from wave 0, then d'_k from wave 1
  const int lane = threadIdx.x % kWave;
  const int64_t c0 = static_cast<int64_t>(xcd_tile(blockIdx.x, gridDim.x)) * kWave;
  const int64_t col = c0 + lane;
  const bool active = col < ncols;
  // an inactive lane (past the last column) reads column c0 and stores nothing
  const int64_t cc = active ? col : c0;
  const int nchunks = (K + L - 1) / L;
  using F0 = std::integral_constant<int, 0>;
  using F2 = std::integral_constant<int, 2>;
  using Gy = std::true_type;
  using Gn = std::false_type;

  // Loads of chunk n of NS streams, clamped into the column (K % V == 0).
  auto loader = [&](const S* const* src, auto ns) {
    return [=](int n, S (&r)[decltype(ns)::value][L]) {
#pragma unroll
      for (int q = 0; q < decltype(ns)::value; ++q) {
#pragma unroll
        for (int m = 0; m < L / V; ++m) {
          const int k = min(max(n * L + m * V, 0), K - V);
          const uint4 raw = *reinterpret_cast<const uint4*>(src[q] + k);
          __builtin_memcpy(&r[q][m * V], &raw, 16);
        }
      }
    };
  };

  if (threadIdx.x < kWave) {
    // ======== wave 0: the pivot chain (a, b, c and the divisions of
    // vadv_numpy.py:15-68)
    const S* src[2] = {wcon + cc * K, wcon + (cc + J) * K};
    using N2 = std::integral_constant<int, 2>;
    auto load = loader(src, N2{});
    double sw[L];
    auto convert = [&](const S (&r)[2][L]) {
#pragma unroll
      for (int l = 0; l < L; ++l)
        sw[l] = static_cast<double>(r[1][l]) + static_cast<double>(r[0][l]);
    };
    double s_cur = 0, c_prev = 0;
    auto step = [&](int k, double s_next) {  // level k-1, k >= 2
      const double gav = -0.25 * s_cur;
      const double gcv = 0.25 * s_next;
      const double acol = gav * kBetP;
      const double ck = gcv * kBetP;
      const double bcol = dtr - acol - ck;
      const double inv = 1.0 / (bcol - c_prev * acol);
      c_prev = ck * inv;
      cp[(k - 1) * kWave + lane] = c_prev;
      dp[(k - 1) * kWave + lane] = inv;
      s_cur = s_next;
    };
    auto body = [&](auto first, auto guard, int k0) {
#pragma unroll
      for (int l = decltype(first)::value; l < L; ++l) {
        if constexpr (decltype(guard)::value) {
          if (k0 + l >= K) break;
        }
        step(k0 + l, sw[l]);
      }
    };
    auto last = [&]() {  // level K-1 (vadv_numpy.py:55-68)
      const double gav = -0.25 * s_cur;
      const double acol = gav * kBetP;
      const double bcol = dtr - acol;
      dp[(K - 1) * kWave + lane] = 1.0 / (bcol - c_prev * acol);
    };
    S ra[2][L], rb[2][L];
    load(0, ra);
    load(1, rb);
    convert(ra);
    load(2, ra);
    s_cur = sw[0];
    {  // level 0 (vadv_numpy.py:15-30)
      const double gcv = 0.25 * sw[1];
      const double ck = gcv * kBetP;
      const double bcol = dtr - ck;
      const double inv = 1.0 / bcol;
      c_prev = ck * inv;
      cp[lane] = c_prev;
      dp[lane] = inv;
      s_cur = sw[1];
    }
    if (nchunks == 1) {
      body(F2{}, Gy{}, 0);
      last();
    } else {
      body(F2{}, Gn{}, 0);
    }
    __syncthreads();
    int n = 1;
    for (; n + 1 <= nchunks - 2; n += 2) {
      convert(rb);
      load(n + 2, rb);
      body(F0{}, Gn{}, n * L);
      __syncthreads();
      convert(ra);
      load(n + 3, ra);
      body(F0{}, Gn{}, (n + 1) * L);
      __syncthreads();
    }
    if (n <= nchunks - 2) {  // one more middle chunk (n odd: in rb)
      convert(rb);
      load(n + 2, rb);
      body(F0{}, Gn{}, n * L);
      __syncthreads();
      ++n;
    }
    if (n <= nchunks - 1) {  // the last chunk
      if (n & 1) {
        convert(rb);
      } else {
        convert(ra);
      }
      body(F0{}, Gy{}, n * L);
      last();
      __syncthreads();
    }
    __syncthreads();  // wave 1's last chunk
    return;
  }

  // ======== wave 1: right-hand sides, d', back substitution
  const S* src[6] = {utens_stage + cc * K, u_stage + cc * K, u_pos + cc * K,
                     utens + cc * K,       wcon + cc * K,    wcon + (cc + J) * K};
  using N6 = std::integral_constant<int, 6>;
  auto load = loader(src, N6{});
  double u[L], sw[L], base[L];
  auto convert = [&](const S (&r)[6][L]) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      u[l] = static_cast<double>(r[1][l]);
      sw[l] = static_cast<double>(r[5][l]) + static_cast<double>(r[4][l]);
      base[l] = (dtr * static_cast<double>(r[2][l]) + static_cast<double>(r[3][l])) +
                static_cast<double>(r[0][l]);
    }
  };
  double u_prev = 0, u_cur = 0, s_cur = 0, base_cur = 0, d_prev = 0;
  auto step = [&](int k, double u_next, double s_next, double base_next) {  // level k-1
    const double gav = -0.25 * s_cur;
    const double gcv = 0.25 * s_next;
    const double as_ = gav * kBetM;
    const double cs = gcv * kBetM;
    const double acol = gav * kBetP;
    const double corr = -as_ * (u_prev - u_cur) - cs * (u_next - u_cur);
    const double d = base_cur + corr;
    double* slot = &dp[(k - 1) * kWave + lane];
    d_prev = (d - d_prev * acol) * *slot;
    *slot = d_prev;
    u_prev = u_cur;
    u_cur = u_next;
    s_cur = s_next;
    base_cur = base_next;
  };
  auto body = [&](auto first, auto guard, int k0) {
#pragma unroll
    for (int l = decltype(first)::value; l < L; ++l) {
      if constexpr (decltype(guard)::value) {
        if (k0 + l >= K) break;
      }
      step(k0 + l, u[l], sw[l], base[l]);
    }
  };
  auto last = [&]() {  // level K-1 (vadv_numpy.py:55-68)
    const double gav = -0.25 * s_cur;
    const double as_ = gav * kBetM;
    const double acol = gav * kBetP;
    const double corr = -as_ * (u_prev - u_cur);
    const double d = base_cur + corr;
    d_prev = (d - d_prev * acol) * dp[(K - 1) * kWave + lane];
  };
  S ra[6][L], rb[6][L];
  load(0, ra);
  load(1, rb);
  convert(ra);
  load(2, ra);
  __syncthreads();  // wave 0 has chunk 0's pivots
  u_cur = u[0];
  s_cur = sw[0];
  base_cur = base[0];
  {  // level 0 (vadv_numpy.py:15-30)
    const double gcv = 0.25 * sw[1];
    const double cs = gcv * kBetM;
    const double corr = -cs * (u[1] - u_cur);
    const double d = base_cur + corr;
    d_prev = d * dp[lane];
    dp[lane] = d_prev;
    u_prev = u_cur;
    u_cur = u[1];
    s_cur = sw[1];
    base_cur = base[1];
  }
  if (nchunks == 1) {
    body(F2{}, Gy{}, 0);
    last();
  } else {
    body(F2{}, Gn{}, 0);
  }
  __syncthreads();
  int n = 1;
  for (; n + 1 <= nchunks - 2; n += 2) {
    convert(rb);
    load(n + 2, rb);
    body(F0{}, Gn{}, n * L);
    __syncthreads();
    convert(ra);
    load(n + 3, ra);
    body(F0{}, Gn{}, (n + 1) * L);
    __syncthreads();
  }
  if (n <= nchunks - 2) {
    convert(rb);
    load(n + 2, rb);
    body(F0{}, Gn{}, n * L);
    __syncthreads();
    ++n;
  }
  if (n <= nchunks - 1) {
    if (n & 1) {
      convert(rb);
    } else {
      convert(ra);
    }
    body(F0{}, Gy{}, n * L);
    last();
    __syncthreads();
  }

  // ---- back substitution (vadv_numpy.py:70-78), top chunk first; u_pos of
  // the chunk below is in flight while a chunk is solved
  double x = d_prev;
  const S* up_src[1] = {src[2]};
  using N1 = std::integral_constant<int, 1>;
  auto load_up = loader(up_src, N1{});
  S* out = utens_stage + cc * K;
  double up[L];
  auto back = [&](auto top, int k0, const S (&r)[1][L]) {
#pragma unroll
    for (int l = 0; l < L; ++l) up[l] = static_cast<double>(r[0][l]);
    S o[L];
#pragma unroll
    for (int l = L - 1; l >= 0; --l) {
      const int k = k0 + l;
      if constexpr (decltype(top)::value) {
        if (k >= K) continue;
        if (k < K - 1) x = dp[k * kWave + lane] - cp[k * kWave + lane] * x;
      } else {
        x = dp[k * kWave + lane] - cp[k * kWave + lane] * x;
      }
      o[l] = static_cast<S>(dtr * (x - up[l]));
    }
    if (active) {
#pragma unroll
      for (int m = 0; m < L / V; ++m) {
        if (decltype(top)::value && k0 + m * V >= K) break;
        uint4 raw;
        __builtin_memcpy(&raw, &o[m * V], 16);
        *reinterpret_cast<uint4*>(out + k0 + m * V) = raw;
      }
    }
  };
  S pa[1][L], pb[1][L];
  const int nt = nchunks - 1;
  load_up(nt, pa);
  load_up(nt - 1, pb);
  back(Gy{}, nt * L, pa);
  int m = nt - 1;  // chunk m is in pb, m - 1 follows in pa
  for (; m >= 1; m -= 2) {
    load_up(m - 1, pa);
    back(Gn{}, m * L, pb);
    load_up(m - 2, pb);
    back(Gn{}, (m - 1) * L, pa);
  }
  if (m == 0) back(Gn{}, 0, pb);
}

// ---- c' in registers (80 < K <= 160) ---------------------------------------
// vadv_pair's split with wave 0 keeping its columns' c'_k in registers (a
// fully unrolled march over the NLEV levels: 2 VGPRs per level), so LDS
// holds only inv_k / d'_k: 80 KiB for 64 columns x 160 levels, two
// workgroups per CU where vadv_pair fits one.  Wave 0 also runs the back
// substitution (c' from its registers, d' from LDS, u_pos read again).  The
// pivot's reciprocal is the compiler's IEEE division sequence restated
// without its scaling steps (division_reciprocal, then r = fma(-den, y, 1),
// inv = fma(r, y, y)), used where that sequence does not rescale (den
// in range and 1/den below 2^765); elsewhere the division operator.  The
// arithmetic is vadv_pair's, operation for operation.
template <class S, int NLEV>
__global__ __launch_bounds__(2 * kWave) void vadv_reg(S* __restrict__ utens_stage,
                                                      const S* __restrict__ u_stage,
                                                      const S* __restrict__ wcon,
                                                      const S* __restrict__ u_pos,
                                                      const S* __restrict__ utens, double dtr,
                                                      int J, int K, int64_t ncols) {
  constexpr int L = PairShape<S>::L, V = PairShape<S>::V;
  constexpr int NCH = NLEV / L;  // chunks the unrolled march can hold
  static_assert(NLEV % L == 0, "whole chunks of levels");
  __shared__ double dp[NLEV * kWave];  // inv_k from wave 0, then d'_k from wave 1
  const int lane = threadIdx.x % kWave;
  const int64_t c0 = static_cast<int64_t>(xcd_tile(blockIdx.x, gridDim.x)) * kWave;
  const int64_t col = c0 + lane;
  const bool active = col < ncols;
  const int64_t cc = active ? col : c0;
  const int nchunks = (K + L - 1) / L;

  auto load_streams = [&](const S* const* src, auto ns, int n, S (&r)[decltype(ns)::value][L]) {
#pragma unroll
    for (int q = 0; q < decltype(ns)::value; ++q) {
#pragma unroll
      for (int m = 0; m < L / V; ++m) {
        const int k = min(max(n * L + m * V, 0), K - V);
        const uint4 raw = *reinterpret_cast<const uint4*>(src[q] + k);
        __builtin_memcpy(&r[q][m * V], &raw, 16);
      }
    }
  };
  // 1 / den: the division sequence's bits without its scaling steps where
  // they do not apply

  if (threadIdx.x < kWave) {
    // ======== wave 0: the pivot chain, c' kept in registers (every loop
    // over chunks and levels fully unrolled, guards instead of exits)
    using N2 = std::integral_constant<int, 2>;
    const S* src[2] = {wcon + cc * K, wcon + (cc + J) * K};
    double cpr[NLEV];
    S ra[2][L], rb[2][L];
    load_streams(src, N2{}, 0, ra);
    load_streams(src, N2{}, 1, rb);
    double s_cur = 0, c_prev = 0;
#pragma unroll
    for (int n = 0; n < NCH; ++n) {
      if (n < nchunks) {
        double sw[L];
        auto& cur = (n & 1) ? rb : ra;
#pragma unroll
        for (int l = 0; l < L; ++l)
          sw[l] = static_cast<double>(cur[1][l]) + static_cast<double>(cur[0][l]);
        load_streams(src, N2{}, n + 2, cur);  // two chunks ahead (clamped past the end)
#pragma unroll
        for (int l = 0; l < L; ++l) {
          const int k = n * L + l;
          if (k == 0) {
            s_cur = sw[0];
          } else if (k == 1) {  // level 0 (vadv_numpy.py:15-30)
            const double gcv = 0.25 * sw[1];
            const double ck = gcv * kBetP;
            const double bcol = dtr - ck;
            const double inv = pivot_reciprocal(bcol);
            c_prev = ck * inv;
            cpr[0] = c_prev;
            dp[lane] = inv;
            s_cur = sw[1];
          } else if (k < K) {  // level k-1 (vadv_numpy.py:32-53)
            const double gav = -0.25 * s_cur;
            const double gcv = 0.25 * sw[l];
            const double acol = gav * kBetP;
            const double ck = gcv * kBetP;
            const double bcol = dtr - acol - ck;
            const double inv = pivot_reciprocal(bcol - c_prev * acol);
            c_prev = ck * inv;
            cpr[k - 1] = c_prev;
            dp[(k - 1) * kWave + lane] = inv;
            s_cur = sw[l];
          }
        }
        if (n == nchunks - 1) {  // level K-1 (vadv_numpy.py:55-68)
          const double gav = -0.25 * s_cur;
          const double acol = gav * kBetP;
          const double bcol = dtr - acol;
          dp[(K - 1) * kWave + lane] = pivot_reciprocal(bcol - c_prev * acol);
        }
        __syncthreads();
      }
    }
    __syncthreads();  // wave 1's last chunk: every d'_k is in LDS

    // ---- back substitution (vadv_numpy.py:70-78), top chunk first; u_pos
    // of the chunk below is in flight while a chunk is solved
    using N1 = std::integral_constant<int, 1>;
    const S* up_src[1] = {u_pos + cc * K};
    S* out = utens_stage + cc * K;
    double x = dp[(K - 1) * kWave + lane];
    S pa[1][L], pb[1][L];  // chunk n's u_pos in pb when n is odd, else in pa
    if ((nchunks - 1) & 1)
      load_streams(up_src, N1{}, nchunks - 1, pb);
    else
      load_streams(up_src, N1{}, nchunks - 1, pa);
#pragma unroll
    for (int n = NCH - 1; n >= 0; --n) {
      if (n < nchunks) {
        auto& cur = (n & 1) ? pb : pa;
        auto& nxt = (n & 1) ? pa : pb;
        if (n > 0) load_streams(up_src, N1{}, n - 1, nxt);
        S o[L];
#pragma unroll
        for (int l = L - 1; l >= 0; --l) {
          const int k = n * L + l;
          if (k < K - 1) x = dp[k * kWave + lane] - cpr[k] * x;
          o[l] = static_cast<S>(dtr * (x - static_cast<double>(cur[0][l])));
        }
        if (active) {
#pragma unroll
          for (int m = 0; m < L / V; ++m) {
            if (n * L + m * V < K) {
              uint4 raw;
              __builtin_memcpy(&raw, &o[m * V], 16);
              *reinterpret_cast<uint4*>(out + n * L + m * V) = raw;
            }
          }
        }
      }
    }
    return;
  }

  // ======== wave 1: right-hand sides and d' (vadv_pair's wave 1 forward
  // pass), d'_{K-1} written to LDS as well
  const S* src[6] = {utens_stage + cc * K, u_stage + cc * K, u_pos + cc * K,
                     utens + cc * K,       wcon + cc * K,    wcon + (cc + J) * K};
  using N6 = std::integral_constant<int, 6>;
  double u_prev = 0, u_cur = 0, s_cur = 0, base_cur = 0, d_prev = 0;
  S ra[6][L], rb[6][L];
  load_streams(src, N6{}, 0, ra);
  load_streams(src, N6{}, 1, rb);
  for (int n = 0; n < nchunks; n += 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int nn = n + h;
      if (nn >= nchunks) break;
      auto& cur = h ? rb : ra;
      double u[L], sw[L], base[L];
#pragma unroll
      for (int l = 0; l < L; ++l) {
        u[l] = static_cast<double>(cur[1][l]);
        sw[l] = static_cast<double>(cur[5][l]) + static_cast<double>(cur[4][l]);
        base[l] = (dtr * static_cast<double>(cur[2][l]) + static_cast<double>(cur[3][l])) +
                  static_cast<double>(cur[0][l]);
      }
      load_streams(src, N6{}, nn + 2, cur);
      __syncthreads();  // wave 0 has chunk nn's pivots
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const int k = nn * L + l;
        if (k >= K) break;
        if (k == 0) {
          u_cur = u[0];
          s_cur = sw[0];
          base_cur = base[0];
          continue;
        }
        if (k == 1) {  // level 0 (vadv_numpy.py:15-30)
          const double gcv = 0.25 * sw[1];
          const double cs = gcv * kBetM;
          const double corr = -cs * (u[1] - u_cur);
          const double d = base_cur + corr;
          d_prev = d * dp[lane];
          dp[lane] = d_prev;
        } else {  // level k-1
          const double gav = -0.25 * s_cur;
          const double gcv = 0.25 * sw[l];
          const double as_ = gav * kBetM;
          const double cs = gcv * kBetM;
          const double acol = gav * kBetP;
          const double corr = -as_ * (u_prev - u_cur) - cs * (u[l] - u_cur);
          const double d = base_cur + corr;
          double* slot = &dp[(k - 1) * kWave + lane];
          d_prev = (d - d_prev * acol) * *slot;
          *slot = d_prev;
        }
        u_prev = u_cur;
        u_cur = u[l];
        s_cur = sw[l];
        base_cur = base[l];
      }
      if (nn == nchunks - 1) {  // level K-1 (vadv_numpy.py:55-68)
        const double gav = -0.25 * s_cur;
        const double as_ = gav * kBetM;
        const double acol = gav * kBetP;
        const double corr = -as_ * (u_prev - u_cur);
        const double d = base_cur + corr;
        double* slot = &dp[(K - 1) * kWave + lane];
        *slot = (d - d_prev * acol) * *slot;
      }
    }
  }
  __syncthreads();  // d' complete
}

// ---- coalesced loads, c' and d' in registers (80 < K <= 160) ---------------
// vadv_stage: the two-wave split of vadv_reg with every field chunk loaded
// coalesced -- a wave load covers 16 columns' 64-byte chunk segments, four
// lanes per segment, instead of one line per lane -- and turned to lane =
// column through LDS (a column's four 16-byte quads at XOR-swizzled slots,
// so both the coalesced writes and the per-column reads are free of bank
// conflicts).  Wave 0 keeps c'_k in registers, wave 1 keeps d'_k in
// registers, so LDS holds only the staged chunks and three small rings
// (inv_k and cs_k from wave 0 to wave 1 on the way down, c'_k on the way
// back up): 80 KiB (FP32) or 64 KiB (FP64) per 64 columns, two workgroups
// per CU.  Wave 1 needs no wcon: the reference's cs_k = gcv_k BET_M equals
// wave 0's ccol coefficient gcv_k BET_P (BET_M = BET_P), and as_k = acol_k
// = gav_k BET_M = -cs_{k-1} exactly (gav_k = -0.25 s_k and gcv_{k-1} =
// 0.25 s_k differ only in sign, and rounding is symmetric).  K is a whole
// number of chunks (L = 64 bytes of levels), so a chunk's levels run with no
// branch between them and its ring entries are read together; NCHK > 0
// fixes the chunk count at compile time (K = 160), which removes every
// branch between the intervals and lets the compiler count the loads in
// flight exactly (with a runtime count it waits for all of them).
//   forward, interval n (n = 0 .. nchunks, a barrier after each):
//     wave 0 runs the pivot chain of chunk n from T0, inv_k and cs_k into
//       the rings' buffer n % 2 (the chain's latency bounds the forward, so
//       wave 0 does nothing else);
//     wave 1 stages the four fields' chunk n (T1, two buffers) and wcon's
//       chunk n+1 (T0, two buffers), loads the fields' chunk n+1 and wcon's
//       n+2, and forms d'_k of chunk n-1 from T1 and the rings' buffer
//       (n-1) % 2;
//   backward, interval t (t = nchunks-1 .. -2):
//     wave 0 writes c'_k of chunk t into ring[t % 2], stages u_pos's chunk
//       t (loaded an interval ahead) and stores the output of chunk t+2
//       from LDS, coalesced;
//     wave 1 solves chunk t+1 (x_k = d'_k - c'_k x_{k+1}, out = dtr (x -
//       u_pos)) into an output buffer in LDS, lane = column.
// Chunk loops are unrolled (register arrays need constant indices); the
// arithmetic is vadv_reg's, operation for operation.
template <class S, int NCHK>
__global__ __launch_bounds__(2 * kWave) void vadv_stage(S* __restrict__ utens_stage,
                                                        const S* __restrict__ u_stage,
                                                        const S* __restrict__ wcon,
                                                        const S* __restrict__ u_pos,
                                                        const S* __restrict__ utens, double dtr,
                                                        int J, int K, int64_t ncols) {
  constexpr int NLEV = 160;
  constexpr int L = 64 / static_cast<int>(sizeof(S));  // levels per chunk
  constexpr int V = 16 / static_cast<int>(sizeof(S));  // levels per 16-byte quad
  constexpr int NCH = NLEV / L;
  // wcon rows i, i+1: [buffer][stream][slot]; a buffer also carries level
  // K-1's inv_k as doubles in the last forward interval (see fin)
  __shared__ union {
    q32x4 q[2][2][4 * kWave];
    double d[2][2 * 4 * kWave * 2];
  } t0u;
  auto& t0 = t0u.q;
  __shared__ q32x4 t1[2][4][4 * kWave];  // utens_stage, u_stage, u_pos, utens
  __shared__ double ring[2][L * kWave];     // inv_k down, c'_k back up
  __shared__ double ring_cs[2][L * kWave];  // cs_k down
  static_assert(sizeof(t0u) + sizeof(t1) + sizeof(ring) + sizeof(ring_cs) <= 80 * 1024,
                "two workgroups per CU");
  const int lane = threadIdx.x % kWave;
  const int64_t c0 = static_cast<int64_t>(xcd_tile(blockIdx.x, gridDim.x)) * kWave;
  const int nchunks = NCHK > 0 ? NCHK : K / L;
  // level K-1's inv_k, wave 0 to wave 1 in the last forward interval: the
  // t0 buffer no chunk is staged into then
  double* const fin = t0u.d[nchunks % 2];
  // coalesced role: load m of a stream's chunk gives lane g quad g % 4 of
  // column 16 m + g / 4 (columns past the last read the last column)
  const int lq = lane & 3, lcol = lane >> 2;
  const int sw_w = (lane >> 4) & 3;  // swizzle of the columns this lane writes
  const int sw_r = (lane >> 2) & 3;  // swizzle of column `lane`
  auto src_of = [&](const S* f, int m, int n) {
    const int64_t col = min(c0 + 16 * m + lcol, ncols - 1);
    return f + col * K + n * L + lq * V;
  };
  auto load4 = [&](const S* f, int n, q32x4 (&r)[4]) {
#pragma unroll
    for (int m = 0; m < 4; ++m) r[m] = *reinterpret_cast<const q32x4*>(src_of(f, m, n));
  };
  auto stage = [&](q32x4* buf, const q32x4 (&r)[4]) {
#pragma unroll
    for (int m = 0; m < 4; ++m) buf[(16 * m + lcol) * 4 + (lq ^ sw_w)] = r[m];
  };
  // quad j (V levels) of column `lane` of a staged chunk, as doubles
  auto conv = [&](const q32x4 q, double (&out)[V]) {
    S v[V];
    __builtin_memcpy(v, &q, 16);
#pragma unroll
    for (int e = 0; e < V; ++e) out[e] = static_cast<double>(v[e]);
  };
  auto quad = [&](const q32x4* buf, int j, double (&out)[V]) {
    conv(buf[lane * 4 + (j ^ sw_r)], out);
  };

  if (threadIdx.x < kWave) {
    // ======== wave 0: the pivot chain, c' in registers
    double cpr[NLEV];
    __syncthreads();  // wcon's chunk 0 staged
    double s_cur = 0, c_prev = 0;
#pragma unroll
    for (int n = 0; n <= NCH; ++n) {
      if (n <= nchunks) {
        if (n < nchunks) {
          // The chunk's levels with the short reciprocal; if any lane's
          // pivot left its range, the chunk again from the same state with
          // the division operator (bitwise the short form where it holds).
          double* rg = ring[n % 2];
          double* rc = ring_cs[n % 2];
          auto levels = [&](auto full) {
            double lo = INFINITY, hi = 0.0;
            auto recip = [&](double den) {
              if constexpr (decltype(full)::value)
                return 1.0 / den;
              else
                return short_reciprocal(den, lo, hi);
            };
#pragma unroll
            for (int j = 0; j < 4; ++j) {  // quad j: levels n L + j V ..
              double sw[V];
              {
                double a[V];
                quad(t0[n % 2][0], j, a);
                quad(t0[n % 2][1], j, sw);
#pragma unroll
                for (int e = 0; e < V; ++e) sw[e] = sw[e] + a[e];
              }
#pragma unroll
              for (int e = 0; e < V; ++e) {
                const int l = j * V + e;
                const int k = n * L + l;
                if (k == 0) {
                  s_cur = sw[e];
                } else if (k == 1) {  // level 0 (vadv_numpy.py:15-30)
                  const double gcv = 0.25 * sw[e];
                  const double ck = gcv * kBetP;
                  const double bcol = dtr - ck;
                  const double inv = recip(bcol);
                  c_prev = ck * inv;
                  cpr[0] = c_prev;
                  rg[l * kWave + lane] = inv;
                  rc[l * kWave + lane] = ck;
                  s_cur = sw[e];
                } else {  // level k-1 (vadv_numpy.py:32-53)
                  const double gav = -0.25 * s_cur;
                  const double gcv = 0.25 * sw[e];
                  const double acol = gav * kBetP;
                  const double ck = gcv * kBetP;
                  const double bcol = dtr - acol - ck;
                  const double inv = recip(bcol - c_prev * acol);
                  c_prev = ck * inv;
                  cpr[k - 1] = c_prev;
                  rg[l * kWave + lane] = inv;
                  rc[l * kWave + lane] = ck;
                  s_cur = sw[e];
                }
              }
            }
            if (n == nchunks - 1) {  // level K-1 (vadv_numpy.py:55-68)
              const double gav = -0.25 * s_cur;
              const double acol = gav * kBetP;
              const double bcol = dtr - acol;
              fin[lane] = recip(bcol - c_prev * acol);
            }
            return reciprocal_range_ok(lo, hi);
          };
          const double c_start = c_prev, s_start = s_cur;
          if (__builtin_amdgcn_ballot_w64(!levels(std::false_type{})) != 0) {
            c_prev = c_start;
            s_cur = s_start;
            levels(std::true_type{});
          }
        }
        __syncthreads();
      }
    }
    // ---- backward, interval t: c' of chunk t into ring[t % 2] (entries
    // 0 .. L-1 hold levels t L ..), u_pos's chunk t staged (loaded one
    // interval ahead), and chunk t+2's output (solved by wave 1 in interval
    // t+1) stored from obuf[t % 2] coalesced
    {
      q32x4 up4[4];
      load4(u_pos, nchunks - 1, up4);
#pragma unroll
      for (int t = NCH - 1; t >= -2; --t) {
        if (t <= nchunks - 1) {
          if (t >= 0) {
#pragma unroll
            for (int l = 0; l < L; ++l)
              if (t * L + l < NLEV - 1) ring[t % 2][l * kWave + lane] = cpr[t * L + l];
            stage(t1[t % 2][2], up4);
            if (t >= 1) load4(u_pos, t - 1, up4);
          }
          const int qs = t + 2;
          if (qs <= NCH - 1 && qs < nchunks) {
            const q32x4* ob = t0[0][qs % 2];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
              const int64_t col = c0 + 16 * m + lcol;
              if (col < ncols)
                *reinterpret_cast<q32x4*>(utens_stage + col * K + qs * L + lq * V) =
                    ob[(16 * m + lcol) * 4 + (lq ^ sw_w)];
            }
          }
          __syncthreads();
        }
      }
    }
    return;
  }

  // ======== wave 1: the fields, d' in registers, the back substitution
  double dpr[NLEV];
  q32x4 st[4][4];  // the next chunk's fields
  q32x4 wt[2][4];  // wcon rows i, i+1 of the chunk after
  const S* fields[4] = {utens_stage, u_stage, u_pos, utens};
  const S* wrow[2] = {wcon, wcon + static_cast<int64_t>(J) * K};
  {
    q32x4 w0[2][4];
#pragma unroll
    for (int r = 0; r < 2; ++r) load4(wrow[r], 0, w0[r]);
#pragma unroll
    for (int r = 0; r < 2; ++r) load4(wrow[r], 1, wt[r]);  // nchunks >= 2 (K > 80)
#pragma unroll
    for (int f = 0; f < 4; ++f) load4(fields[f], 0, st[f]);
#pragma unroll
    for (int r = 0; r < 2; ++r) stage(t0[0][r], w0[r]);
  }
  __syncthreads();  // wcon's chunk 0 staged
  double u_prev = 0, u_cur = 0, base_cur = 0, d_prev = 0, cs_prev = 0;
#pragma unroll
  for (int n = 0; n <= NCH; ++n) {
    if (n <= nchunks) {
      if (n < nchunks) {  // stage the fields' chunk n and wcon's n+1; load n+1, n+2
#pragma unroll
        for (int f = 0; f < 4; ++f) stage(t1[n % 2][f], st[f]);
        if (n + 1 < nchunks) {
#pragma unroll
          for (int r = 0; r < 2; ++r) stage(t0[(n + 1) % 2][r], wt[r]);
#pragma unroll
          for (int f = 0; f < 4; ++f) load4(fields[f], n + 1, st[f]);
        }
        if (n + 2 < nchunks) {
#pragma unroll
          for (int r = 0; r < 2; ++r) load4(wrow[r], n + 2, wt[r]);
        }
      }
      if (n >= 1) {  // d' of chunk p = n-1
        const int p = n - 1;
        const double* rg = ring[p % 2];
        const double* rc = ring_cs[p % 2];
#pragma unroll
        for (int j = 0; j < 4; ++j) {  // quad j
          double u[V], base[V], inv[V], cs[V];
#pragma unroll
          for (int e = 0; e < V; ++e) {
            inv[e] = rg[(j * V + e) * kWave + lane];
            cs[e] = rc[(j * V + e) * kWave + lane];
          }
          {
            double fs[V], fp[V], ft[V];
            quad(t1[p % 2][0], j, fs);
            quad(t1[p % 2][1], j, u);
            quad(t1[p % 2][2], j, fp);
            quad(t1[p % 2][3], j, ft);
#pragma unroll
            for (int e = 0; e < V; ++e) base[e] = (dtr * fp[e] + ft[e]) + fs[e];
          }
#pragma unroll
          for (int e = 0; e < V; ++e) {
            const int l = j * V + e;
            const int k = p * L + l;
            if (k == 0) {
              u_cur = u[e];
              base_cur = base[e];
              continue;
            }
            if (k == 1) {  // level 0 (vadv_numpy.py:15-30)
              const double corr = -cs[e] * (u[e] - u_cur);
              const double d = base_cur + corr;
              d_prev = d * inv[e];
              dpr[0] = d_prev;
            } else {  // level k-1
              const double as_ = -cs_prev, acol = as_;
              const double corr = -as_ * (u_prev - u_cur) - cs[e] * (u[e] - u_cur);
              const double d = base_cur + corr;
              d_prev = (d - d_prev * acol) * inv[e];
              dpr[k - 1] = d_prev;
            }
            cs_prev = cs[e];
            u_prev = u_cur;
            u_cur = u[e];
            base_cur = base[e];
          }
        }
        if (p == nchunks - 1) {  // level K-1 (vadv_numpy.py:55-68)
          const double as_ = -cs_prev, acol = as_;
          const double corr = -as_ * (u_prev - u_cur);
          const double d = base_cur + corr;
          dpr[p * L + L - 1] = (d - d_prev * acol) * fin[lane];
        }
      }
      __syncthreads();
    }
  }

  // ---- backward (vadv_numpy.py:70-78), interval t: chunk t+1 solved from
  // d' (registers), c' (ring) and u_pos (staged by wave 0), its output
  // written to obuf[(t+1) % 2] for wave 0 to store
  double x = 0;
#pragma unroll
  for (int t = NCH - 1; t >= -2; --t) {
    if (t <= nchunks - 1) {
      const int q = t + 1;  // the chunk solved in this interval
      if (q >= 0 && q <= NCH - 1 && q < nchunks) {
        q32x4* ob = t0[0][q % 2];
        const bool top = q == nchunks - 1;
        // the chunk's u_pos quads and c' values, all read before the chain
        // starts (one LDS round trip per interval instead of two per quad)
        q32x4 uq[4];
        double cpv[L];
#pragma unroll
        for (int j = 0; j < 4; ++j) uq[j] = t1[q % 2][2][lane * 4 + (j ^ sw_r)];
#pragma unroll
        for (int l = 0; l < L; ++l) cpv[l] = ring[q % 2][l * kWave + lane];
#pragma unroll
        for (int j = 3; j >= 0; --j) {
          double up[V];
          conv(uq[j], up);
          S o[V];
#pragma unroll
          for (int e = V - 1; e >= 0; --e) {
            const int l = j * V + e;
            const int k = q * L + l;
            if (l == L - 1)
              x = top ? dpr[k] : dpr[k] - cpv[l] * x;
            else
              x = dpr[k] - cpv[l] * x;
            o[e] = static_cast<S>(dtr * (x - up[e]));
          }
          q32x4 v;
          __builtin_memcpy(&v, o, 16);
          ob[lane * 4 + (j ^ sw_r)] = v;
        }
      }
      __syncthreads();
    }
  }
}

// Largest K the on-chip kernel holds, and its LDS levels for a given K.
constexpr int kLdsMaxK = 160;
inline int lds_levels(int64_t K) { return K <= 48 ? 48 : K <= 80 ? 80 : 160; }

// Kernel choice ("vadv.path"): 0 = by shape (vadv_stage where it applies,
// else vadv_pair, else vadv_wave), 1 = vadv_wave (device workspace), 2 =
// vadv_pair (c' and d' in LDS), 3 = vadv_reg (c' in registers), 4 =
// vadv_stage (coalesced staged loads, c' and d' in registers); a kernel
// that does not apply to the shape falls through to the next one.
std::atomic<int> g_vadv_path{0};

template <class S>
int run_vadv(S* utens_stage, const S* u_stage, const S* wcon, const S* u_pos, const S* utens,
             double dtr, int64_t I, int64_t J, int64_t K, hipStream_t stream) {
  const int64_t ncols = I * J;
  // 16-byte loads need every column and chunk start 16-byte aligned.
  constexpr int64_t W = 16 / sizeof(S);
  const bool vec = K % W == 0 && aligned16(utens_stage) && aligned16(u_stage) &&
                   aligned16(wcon) && aligned16(u_pos) && aligned16(utens);
  const dim3 grid(static_cast<unsigned>(cdiv(ncols, kWave)));
  const int path = g_vadv_path.load();
  if (vec && K > 80 && K <= kLdsMaxK && path == 3) {
    hipLaunchKernelGGL((vadv_reg<S, 160>), grid, dim3(2 * kWave), 0, stream, utens_stage, u_stage,
                       wcon, u_pos, utens, dtr, static_cast<int>(J), static_cast<int>(K), ncols);
    NPBX_LAUNCHED("vadv_reg");
    return NPBX_OK;
  }
  // vadv_stage: 80 < K <= 160, whole 64-byte chunks of levels per column
  if (vec && K > 80 && K <= kLdsMaxK && K % (64 / sizeof(S)) == 0 && (path == 4 || path == 0)) {
    constexpr int kFull = kLdsMaxK / (64 / static_cast<int>(sizeof(S)));
    if (K == kLdsMaxK)
      hipLaunchKernelGGL((vadv_stage<S, kFull>), grid, dim3(2 * kWave), 0, stream, utens_stage,
                         u_stage, wcon, u_pos, utens, dtr, static_cast<int>(J),
                         static_cast<int>(K), ncols);
    else
      hipLaunchKernelGGL((vadv_stage<S, 0>), grid, dim3(2 * kWave), 0, stream, utens_stage,
                         u_stage, wcon, u_pos, utens, dtr, static_cast<int>(J),
                         static_cast<int>(K), ncols);
    NPBX_LAUNCHED("vadv_stage");
    return NPBX_OK;
  }
  if (vec && K <= kLdsMaxK && path != 1) {
    const int Ji = static_cast<int>(J), Ki = static_cast<int>(K);
    switch (lds_levels(K)) {
      case 48:
        hipLaunchKernelGGL((vadv_pair<S, 48>), grid, dim3(2 * kWave), 0, stream, utens_stage, u_stage,
                           wcon, u_pos, utens, dtr, Ji, Ki, ncols);
        break;
      case 80:
        hipLaunchKernelGGL((vadv_pair<S, 80>), grid, dim3(2 * kWave), 0, stream, utens_stage, u_stage,
                           wcon, u_pos, utens, dtr, Ji, Ki, ncols);
        break;
      default:
        hipLaunchKernelGGL((vadv_pair<S, 160>), grid, dim3(2 * kWave), 0, stream, utens_stage, u_stage,
                           wcon, u_pos, utens, dtr, Ji, Ki, ncols);
    }
    NPBX_LAUNCHED("vadv_pair");
    return NPBX_OK;
  }
  void* ws = nullptr;
  NPBX_TRY(workspace(static_cast<size_t>(2 * ncols * K) * sizeof(double), &ws));
  double* cp = static_cast<double*>(ws);
  double* dp = cp + ncols * K;
  if (vec) {
    hipLaunchKernelGGL((vadv_wave<S, true>), grid, dim3(kWave), 0, stream, utens_stage, u_stage,
                       wcon, u_pos, utens, dtr, static_cast<int>(J), static_cast<int>(K), ncols,
                       cp, dp);
  } else {
    hipLaunchKernelGGL((vadv_wave<S, false>), grid, dim3(kWave), 0, stream, utens_stage, u_stage,
                       wcon, u_pos, utens, dtr, static_cast<int>(J), static_cast<int>(K), ncols,
                       cp, dp);
  }
  NPBX_LAUNCHED("vadv_wave");
  return NPBX_OK;
}

}  // namespace

int vadv_path() { return g_vadv_path.load(); }
void set_vadv_path(int path) { g_vadv_path.store(path); }

}  // namespace npbx

extern "C" int npbx_vadv(int field_dtype, void* utens_stage, const void* u_stage, const void* wcon,
                         const void* u_pos, const void* utens, double dtr_stage, int64_t I,
                         int64_t J, int64_t K, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(I >= 0 && J >= 0 && K >= 0, "vadv: negative extent");
  if (I == 0 || J == 0) return NPBX_OK;
  NPBX_REQUIRE(K >= 2, "vadv: K=%lld; the reference needs K >= 2", static_cast<long long>(K));
  NPBX_REQUIRE(I * J <= (int64_t(1) << 31) && K <= (1 << 20), "vadv: extent out of range");
  NPBX_REQUIRE(utens_stage && u_stage && wcon && u_pos && utens, "vadv: null array");
  hipStream_t s = as_stream(stream);
  switch (field_dtype) {
    case NPBX_F64:
      return run_vadv<double>(static_cast<double*>(utens_stage),
                              static_cast<const double*>(u_stage), static_cast<const double*>(wcon),
                              static_cast<const double*>(u_pos), static_cast<const double*>(utens),
                              dtr_stage, I, J, K, s);
    case NPBX_F32:
      return run_vadv<float>(static_cast<float*>(utens_stage), static_cast<const float*>(u_stage),
                             static_cast<const float*>(wcon), static_cast<const float*>(u_pos),
                             static_cast<const float*>(utens), dtr_stage, I, J, K, s);
    default:
      return fail(NPBX_EUNSUPPORTED, "vadv: field dtype %s not supported (f64, f32)",
                  dtype_name(field_dtype));
  }
}
