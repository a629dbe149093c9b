// Storage conversion between FP64, FP32 and BF16 (round-to-nearest-even).
// Streaming kernel: each thread converts 4 consecutive elements per step of
// a grid-stride loop; the tail is handled element-wise.
#include "common.h"

namespace npbx {
namespace {

// FP64 -> BF16 with a single rounding: first round to FP32 toward zero and
// set the FP32 lsb if anything was dropped ("round to odd"), which keeps the
// sticky information, then round that FP32 to BF16 to nearest-even.  FP32
// has 16 more significand bits than BF16, so the result equals direct RNE.
__device__ __forceinline__ uint16_t f64_to_bf16_bits(double x) {
  float t = __double2float_rz(x);
  if (static_cast<double>(t) != x && !isnan(x) && !isinf(t))
    t = __uint_as_float(__float_as_uint(t) | 1u);
  return f32_to_bf16_bits(t);
}

template <class S, class D>
struct Convert;

template <class T>
struct Convert<T, T> {
  static __device__ __forceinline__ T run(T v) { return v; }
};
template <>
struct Convert<double, float> {
  static __device__ __forceinline__ float run(double v) { return static_cast<float>(v); }
};
template <>
struct Convert<float, double> {
  static __device__ __forceinline__ double run(float v) { return static_cast<double>(v); }
};
template <>
struct Convert<double, bf16> {
  static __device__ __forceinline__ bf16 run(double v) { return bf16{f64_to_bf16_bits(v)}; }
};
template <>
struct Convert<float, bf16> {
  static __device__ __forceinline__ bf16 run(float v) { return bf16{f32_to_bf16_bits(v)}; }
};
template <>
struct Convert<bf16, float> {
  static __device__ __forceinline__ float run(bf16 v) { return Storage<bf16>::load(v); }
};
template <>
struct Convert<bf16, double> {
  static __device__ __forceinline__ double run(bf16 v) {
    return static_cast<double>(Storage<bf16>::load(v));
  }
};

template <class S, class D>
__global__ __launch_bounds__(256) void cast_kernel(const S* __restrict__ src, D* __restrict__ dst,
                                                   int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x * 4;
  int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  for (; i + 3 < n; i += stride) {
    S a = src[i], b = src[i + 1], c = src[i + 2], d = src[i + 3];
    dst[i] = Convert<S, D>::run(a);
    dst[i + 1] = Convert<S, D>::run(b);
    dst[i + 2] = Convert<S, D>::run(c);
    dst[i + 3] = Convert<S, D>::run(d);
  }
  for (; i < n; ++i) dst[i] = Convert<S, D>::run(src[i]);
}

template <class S, class D>
int launch_cast(const void* src, void* dst, int64_t n, hipStream_t stream) {
  int64_t blocks = cdiv(cdiv(n, 4), 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((cast_kernel<S, D>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                     static_cast<const S*>(src), static_cast<D*>(dst), n);
  NPBX_LAUNCHED("cast_kernel");
  return NPBX_OK;
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_cast(int src_dtype, const void* src, int dst_dtype, void* dst, int64_t count,
                         void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(count >= 0, "npbx_cast: negative count");
  if (count == 0) return NPBX_OK;
  NPBX_REQUIRE(src && dst, "npbx_cast: null pointer");
  NPBX_REQUIRE(dtype_size(src_dtype) && dtype_size(dst_dtype), "npbx_cast: bad dtype %d -> %d",
               src_dtype, dst_dtype);
  hipStream_t s = as_stream(stream);
  return dispatch_dtype(src_dtype, [&](auto s_tag) {
    using S = decltype(s_tag);
    return dispatch_dtype(dst_dtype, [&](auto d_tag) {
      using D = decltype(d_tag);
      return launch_cast<S, D>(src, dst, count, s);
    });
  });
}
