// Streaming copy used to calibrate bandwidth figures: the measured copy rate
// on the box is reported beside the 8 TB/s datasheet peak, and its HBM
// counter reading scales the FETCH_SIZE/WRITE_SIZE traffic of the stencils.
#include <algorithm>

#include "common.h"
#include "vecio.h"

namespace npbx {
namespace {

// Each thread moves four 16-byte vectors per iteration, all four loads issued
// before the stores, over a grid-stride loop.
__global__ __launch_bounds__(256) void stream_copy(const uint4* __restrict__ src,
                                                   uint4* __restrict__ dst, int64_t n16) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x * 4;
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x * 4 + threadIdx.x;
  for (; i + 3 * blockDim.x < n16; i += stride) {
    const uint4 a = src[i], b = src[i + blockDim.x], c = src[i + 2 * blockDim.x],
                d = src[i + 3 * blockDim.x];
    dst[i] = a;
    dst[i + blockDim.x] = b;
    dst[i + 2 * blockDim.x] = c;
    dst[i + 3 * blockDim.x] = d;
  }
  for (; i < n16; i += blockDim.x) dst[i] = src[i];
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_stream_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(bytes >= 0 && bytes % 16 == 0, "stream_copy: bytes must be a multiple of 16");
  if (bytes == 0) return NPBX_OK;
  NPBX_REQUIRE(dst && src && aligned16(dst) && aligned16(src),
               "stream_copy: null or misaligned pointer");
  const int64_t n16 = bytes / 16;
  // Up to 16 workgroups of 256 threads per CU (256 CUs), four vectors each.
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(cdiv(n16, 1024), 256 * 16));
  hipLaunchKernelGGL(stream_copy, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     as_stream(stream), static_cast<const uint4*>(src), static_cast<uint4*>(dst),
                     n16);
  NPBX_LAUNCHED("stream_copy");
  return NPBX_OK;
}
