// Streaming copy used to calibrate bandwidth figures: the measured copy rate
// on the box is reported beside the 8 TB/s datasheet peak, and its HBM
// counter reading scales the FETCH_SIZE/WRITE_SIZE traffic of the stencils.
#include <algorithm>

#include "common.h"
#include "vecio.h"

namespace npbx {
namespace {

__global__ __launch_bounds__(256) void stream_copy(const uint4* __restrict__ src,
                                                   uint4* __restrict__ dst, int64_t n16) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
       i += stride)
    dst[i] = src[i];
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
  // 8 workgroups of 256 threads per CU on 256 CUs keep enough requests in flight.
  const int64_t blocks = std::min<int64_t>(cdiv(n16, 256), 256 * 8);
  hipLaunchKernelGGL(stream_copy, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     as_stream(stream), static_cast<const uint4*>(src), static_cast<uint4*>(dst),
                     n16);
  NPBX_LAUNCHED("stream_copy");
  return NPBX_OK;
}
