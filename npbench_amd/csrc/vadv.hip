// Implicit vertical advection (vadv_numpy.py:9-78): one tridiagonal system
// per (i, j) column of K levels, solved by the Thomas algorithm in FP64.
//
// Layout: utens_stage, u_stage, u_pos, utens (I, J, K); wcon (I+1, J, K); C
// order, so each column's K values are contiguous.  Field storage is FP64 or
// FP32; the coefficients, the forward sweep and the back substitution are
// FP64 in both cases, in the reference's order of operations (reciprocal of
// the pivot, then multiply), so the FP64 path is bitwise equal to NumPy.
//
// Kernel (one thread per column, 256 columns per workgroup): the forward
// sweep streams the five input fields once and parks c'_k and d'_k in a
// library workspace laid out [k][column] so that the writes and the reversed
// reads of the back substitution are coalesced across the wave.
#include "common.h"

namespace npbx {
namespace {

constexpr double kBetM = 0.5;  // vadv_numpy.py:4
constexpr double kBetP = 0.5;  // vadv_numpy.py:5
constexpr int kThreads = 256;

template <class S>
__device__ __forceinline__ double ld(const S* p) {
  return static_cast<double>(*p);
}

template <class S>
__global__ __launch_bounds__(kThreads) void vadv_thomas(S* __restrict__ utens_stage,
                                                        const S* __restrict__ u_stage,
                                                        const S* __restrict__ wcon,
                                                        const S* __restrict__ u_pos,
                                                        const S* __restrict__ utens, double dtr,
                                                        int J, int K, int64_t ncols,
                                                        double* __restrict__ cp,
                                                        double* __restrict__ dp) {
  const int64_t col = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (col >= ncols) return;
  const int64_t base = col * K;  // (i, j, 0) of the (I, J, K) fields
  // wcon[i, j, k] and wcon[i+1, j, k]: same column index, and one J*K plane on.
  const S* w_lo = wcon + base;
  const S* w_hi = wcon + base + static_cast<int64_t>(J) * K;
  auto wsum = [&](int k) { return ld(w_hi + k) + ld(w_lo + k); };
  auto rhs = [&](int k, double corr) {
    return ((dtr * ld(u_pos + base + k) + ld(utens + base + k)) + ld(utens_stage + base + k)) +
           corr;
  };

  // k = 0 (vadv_numpy.py:15-30)
  double s_next = wsum(1);
  double u_prev = ld(u_stage + base);
  double u_cur = ld(u_stage + base + 1);
  double gcv = 0.25 * s_next;
  double cs = gcv * kBetM;
  double ck = gcv * kBetP;
  double bcol = dtr - ck;
  double corr = -cs * (u_cur - u_prev);
  double d = rhs(0, corr);
  double inv = 1.0 / bcol;
  double c_prev = ck * inv;
  double d_prev = d * inv;
  cp[col] = c_prev;
  dp[col] = d_prev;

  // 0 < k < K-1 (vadv_numpy.py:32-53); u_prev/u_cur slide to k-1/k.
  double s_cur = s_next;
  for (int k = 1; k < K - 1; ++k) {
    s_next = wsum(k + 1);
    const double u_next = ld(u_stage + base + k + 1);
    const double gav = -0.25 * s_cur;
    gcv = 0.25 * s_next;
    const double as_ = gav * kBetM;
    cs = gcv * kBetM;
    const double acol = gav * kBetP;
    ck = gcv * kBetP;
    bcol = dtr - acol - ck;
    corr = -as_ * (u_prev - u_cur) - cs * (u_next - u_cur);
    d = rhs(k, corr);
    inv = 1.0 / (bcol - c_prev * acol);
    c_prev = ck * inv;
    d_prev = (d - d_prev * acol) * inv;
    cp[k * ncols + col] = c_prev;
    dp[k * ncols + col] = d_prev;
    u_prev = u_cur;
    u_cur = u_next;
    s_cur = s_next;
  }

  // k = K-1 (vadv_numpy.py:55-68)
  {
    const int k = K - 1;
    const double gav = -0.25 * s_cur;
    const double as_ = gav * kBetM;
    const double acol = gav * kBetP;
    bcol = dtr - acol;
    corr = -as_ * (u_prev - u_cur);
    d = rhs(k, corr);
    inv = 1.0 / (bcol - c_prev * acol);
    d_prev = (d - d_prev * acol) * inv;
  }

  // back substitution (vadv_numpy.py:70-78)
  double x = d_prev;
  utens_stage[base + K - 1] = static_cast<S>(dtr * (x - ld(u_pos + base + K - 1)));
  for (int k = K - 2; k >= 0; --k) {
    x = dp[k * ncols + col] - cp[k * ncols + col] * x;
    utens_stage[base + k] = static_cast<S>(dtr * (x - ld(u_pos + base + k)));
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
  hipLaunchKernelGGL(vadv_thomas<S>, dim3(static_cast<unsigned>(cdiv(ncols, kThreads))),
                     dim3(kThreads), 0, stream, utens_stage, u_stage, wcon, u_pos, utens, dtr,
                     static_cast<int>(J), static_cast<int>(K), ncols, cp, dp);
  NPBX_LAUNCHED("vadv_thomas");
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
