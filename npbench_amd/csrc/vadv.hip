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

template <class S>
int run_vadv(S* utens_stage, const S* u_stage, const S* wcon, const S* u_pos, const S* utens,
             double dtr, int64_t I, int64_t J, int64_t K, hipStream_t stream) {
  const int64_t ncols = I * J;
  void* ws = nullptr;
  NPBX_TRY(workspace(static_cast<size_t>(2 * ncols * K) * sizeof(double), &ws));
  double* cp = static_cast<double*>(ws);
  double* dp = cp + ncols * K;
  // 16-byte loads need every column and chunk start 16-byte aligned.
  constexpr int64_t W = 16 / sizeof(S);
  const bool vec = K % W == 0 && aligned16(utens_stage) && aligned16(u_stage) &&
                   aligned16(wcon) && aligned16(u_pos) && aligned16(utens);
  const dim3 grid(static_cast<unsigned>(cdiv(ncols, kWave)));
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
