// Shared declarations of libnpbx: error reporting, storage types and launch
// helpers.  All device code targets gfx950 (CDNA4, wave64) only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "npbx.h"

namespace npbx {

// ---- error reporting (thread-local last message, see npbx_last_error) ----
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t err, const char* what, const char* file, int line);
void clear_error();

#define NPBX_HIP(call)                                                     \
  do {                                                                     \
    hipError_t npbx_e_ = (call);                                           \
    if (npbx_e_ != hipSuccess)                                             \
      return ::npbx::hip_fail(npbx_e_, #call, __FILE__, __LINE__);         \
  } while (0)

#define NPBX_REQUIRE(cond, ...)                                            \
  do {                                                                     \
    if (!(cond)) return ::npbx::fail(NPBX_EINVAL, __VA_ARGS__);            \
  } while (0)

#define NPBX_TRY(expr)                                                     \
  do {                                                                     \
    int npbx_rc_ = (expr);                                                 \
    if (npbx_rc_ != NPBX_OK) return npbx_rc_;                              \
  } while (0)

// Checks the launch that was just enqueued.
int check_launch(const char* kernel, const char* file, int line);
#define NPBX_LAUNCHED(name) NPBX_TRY(::npbx::check_launch(name, __FILE__, __LINE__))

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- storage types -------------------------------------------------------
// bf16 is kept as its 16-bit pattern; arithmetic happens in float.
struct bf16 {
  uint16_t bits;
};

// Round-to-nearest-even float -> bf16 by the CDNA4 conversion instruction
// (v_cvt_pk_bf16_f32): bitwise the IEEE-754 definition (add half an ulp of
// the target, minus one when the kept lsb is 0, and truncate) for every
// non-NaN float, denormals included (exhaustive GPU test); NaN -> a quiet NaN.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
// Two conversions in one instruction: a in the low half, b in the high.
__device__ __forceinline__ uint32_t f32x2_to_bf16x2_bits(float a, float b) {
  const f32x2_t v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

template <class S>
struct Storage;

template <>
struct Storage<double> {
  using C = double;  // arithmetic type
  static __device__ __forceinline__ double load(double v) { return v; }
  static __device__ __forceinline__ double store(double v) { return v; }
};

template <>
struct Storage<float> {
  using C = float;
  static __device__ __forceinline__ float load(float v) { return v; }
  static __device__ __forceinline__ float store(float v) { return v; }
};

template <>
struct Storage<bf16> {
  using C = float;
  static __device__ __forceinline__ float load(bf16 v) {
    return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
  }
  static __device__ __forceinline__ bf16 store(float v) { return bf16{f32_to_bf16_bits(v)}; }
};

inline size_t dtype_size(int dtype) {
  switch (dtype) {
    case NPBX_F64: return 8;
    case NPBX_F32: return 4;
    case NPBX_BF16: return 2;
    default: return 0;
  }
}

inline const char* dtype_name(int dtype) {
  switch (dtype) {
    case NPBX_F64: return "f64";
    case NPBX_F32: return "f32";
    case NPBX_BF16: return "bf16";
    default: return "?";
  }
}

// Calls f(S{}) with S the storage type for dtype; EINVAL for unknown dtypes.
template <class F>
int dispatch_dtype(int dtype, F&& f) {
  switch (dtype) {
    case NPBX_F64: return f(double{});
    case NPBX_F32: return f(float{});
    case NPBX_BF16: return f(bf16{});
    default: return fail(NPBX_EINVAL, "unknown dtype %d", dtype);
  }
}

// Temporal-blocked sweeps of one grid (kind 0 = Jacobi 2-D, 1 = heat 3-D)
// through the tile kernel: depth T (>= 2) per HBM round trip.
int tiled_sweeps(int kind, int dtype, void* A, void* B, int n, int64_t first, int64_t nsweeps,
                 int T, int seg, hipStream_t stream);

// Depths of the rounds that advance nsweeps sweeps at depth <= T, the last
// two levels ending in A and B as in the reference (tiles.hip).
std::vector<int> sweep_rounds(int64_t nsweeps, int T);

// 2-D Jacobi temporal-blocked in LDS (jacobi2d_deep.hip): sweeps per launch
// ("jacobi.deep"; 0 or 1 = one launch per sweep), waves per workgroup
// ("jacobi.deep_waves") and the sweeps themselves.
int jacobi_deep_depth();
void set_jacobi_deep_depth(int d);
int jacobi_deep_waves();
void set_jacobi_deep_waves(int w);
int jacobi_deep_max();
int jacobi_deep_sweeps(int dtype, void* A, void* B, int n, int64_t first, int64_t nsweeps, int T,
                       hipStream_t stream);

// Plain-sweep shape chosen by the tuning knob "sweep.shape" (0 = 16 planes
// per workgroup with loads of 4 planes in flight, 1 = 16/8, 2 = 32/8,
// 3 = 8/4).
int sweep_shape();
// Fused hdiff steps per round ("hdiff.tblock"; 1 = one launch per step).
int hdiff_tblock();
void set_hdiff_tblock(int tb);
// hdiff marching shape ("hdiff.shape"; 0 = by storage type, 1 .. 13 =
// {columns per lane, rows per thread, bytes of k per lane}, hdiff.hip).
int hdiff_shape();
void set_hdiff_shape(int shape);
// Output rows per workgroup strip of the two-step fused hdiff
// ("hdiff.fuse_rows", 4 .. 256; 0 = by storage type).
int hdiff_fuse_rows();
void set_hdiff_fuse_rows(int rows);
int hdiff_shape_count();
// ADI launch path ("adi.path"; 0 = by size, 1 = every time step in one
// single-workgroup launch, 2 = one launch per half step).
int adi_path();
void set_adi_path(int path);
// vadv kernel choice ("vadv.path"; 0 = by shape, 1 = the workspace kernel,
// 2 = c' and d' in LDS, 3 = c' in registers, 4 = coalesced staged loads with
// c' and d' in registers; 2 .. 4 where the shape allows them).
int vadv_path();
void set_vadv_path(int path);
// Slot-blocks per seidel_2d launch ("seidel.blocks"; a slot-block is one
// wavefront slot over one 256-column block).
int seidel_blocks();
void set_seidel_blocks(int blocks);

// Resolves the current device's address of npbx_adi_noalloc's fixed scratch
// array (once per device; npbx_init calls it).
int adi_noalloc_setup();

// Per-device scratch owned by the library (grow-only, freed at exit).  Used
// by kernels that need storage beyond their inputs and outputs.
int workspace(size_t bytes, void** ptr);

}  // namespace npbx
