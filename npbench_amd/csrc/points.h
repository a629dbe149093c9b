// Point updates of the explicit stencils, shared by the plain sweeps and the
// tile kernels so that both produce bitwise the same values.
#pragma once

#include "common.h"

namespace npbx {

// heat (heat_3d_numpy.py:7-13).  The FP64 drop-in keeps the reference's
// operation order -- per axis 0.125*((x[+1] - 2c) + x[-1]), the axis terms
// summed i, j, k, then + c -- so it is bitwise NumPy.  FP32 arithmetic (FP32
// and BF16 storage) evaluates the same affine map as
// 0.125*(sum of the six neighbours) + 0.25*c with one FMA: 7 operations
// instead of 13, within an ulp or two of the reference order per sweep.  The
// six neighbours are summed plane i-1 first and plane i+1 last, so a kernel
// marching along i can form the first five-term partial sum as soon as plane
// i arrives (heat_partial) and finish the point when plane i+1 does
// (heat_finish); every FP32 kernel evaluates exactly this sequence.
__device__ __forceinline__ float heat_partial(float im, float jm, float jp, float km, float kp) {
  return (((im + jm) + jp) + km) + kp;
}

__device__ __forceinline__ float heat_finish(float partial, float ip, float c) {
  return __builtin_fmaf(0.125f, partial + ip, 0.25f * c);
}

template <class C>
__device__ __forceinline__ C heat_point(C c, C ip, C im, C jp, C jm, C kp, C km) {
  const C c2 = static_cast<C>(2.0) * c;
  const C q = static_cast<C>(0.125);
  const C ti = q * ((ip - c2) + im);
  const C tj = q * ((jp - c2) + jm);
  const C tk = q * ((kp - c2) + km);
  return ((ti + tj) + tk) + c;
}

template <>
__device__ __forceinline__ float heat_point<float>(float c, float ip, float im, float jp,
                                                   float jm, float kp, float km) {
  return heat_finish(heat_partial(im, jm, jp, km, kp), ip, c);
}

// Jacobi (jacobi_2d_numpy.py:7-10): ((((c + W) + E) + S) + N) * 0.2, already
// the minimum of five operations, the same for every precision.
template <class C>
__device__ __forceinline__ C jacobi_point(C c, C w, C e, C s, C n) {
  C acc = c + w;
  acc = acc + e;
  acc = acc + s;
  acc = acc + n;
  return static_cast<C>(0.2) * acc;
}

}  // namespace npbx
