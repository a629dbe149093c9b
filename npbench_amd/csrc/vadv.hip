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

// Largest K the on-chip kernel holds, and its LDS levels for a given K.
constexpr int kLdsMaxK = 160;
inline int lds_levels(int64_t K) { return K <= 48 ? 48 : K <= 80 ? 80 : 160; }

// Kernel choice ("vadv.path"): 0 = by shape, 1 = the workspace kernel, 2 =
// the on-chip kernel where the shape allows it.
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
