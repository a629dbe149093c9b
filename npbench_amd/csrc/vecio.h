// 16-byte vector loads/stores of storage elements and the XCD-aware tile
// order shared by the sweep kernels.
#pragma once

#include "common.h"

namespace npbx {

// Elements of storage type S in one 16-byte vector.
template <class S>
struct Vec16 {
  static constexpr int N = 16 / sizeof(S);
};

// Loads V = Vec16<S>::N consecutive elements from a 16-byte aligned address
// into arithmetic-type registers (one global_load_dwordx4).
template <class S>
__device__ __forceinline__ void load16(const S* __restrict__ p,
                                       typename Storage<S>::C (&out)[Vec16<S>::N]) {
  constexpr int V = Vec16<S>::N;
  const uint4 raw = *reinterpret_cast<const uint4*>(p);
  S tmp[V];
  __builtin_memcpy(tmp, &raw, 16);
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = Storage<S>::load(tmp[v]);
}

template <class S>
__device__ __forceinline__ void store16(S* __restrict__ p,
                                        const typename Storage<S>::C (&in)[Vec16<S>::N]) {
  constexpr int V = Vec16<S>::N;
  S tmp[V];
#pragma unroll
  for (int v = 0; v < V; ++v) tmp[v] = Storage<S>::store(in[v]);
  uint4 raw;
  __builtin_memcpy(&raw, tmp, 16);
  *reinterpret_cast<uint4*>(p) = raw;
}

// V consecutive elements in one 8- or 16-byte access (V * sizeof(S)).
template <class S, int V>
__device__ __forceinline__ void loadv(const S* __restrict__ p, typename Storage<S>::C (&out)[V]) {
  static_assert(V * sizeof(S) == 16 || V * sizeof(S) == 8, "8- or 16-byte vectors");
  if constexpr (V * sizeof(S) == 16) {
    load16(p, out);
  } else {
    const uint2 raw = *reinterpret_cast<const uint2*>(p);
    S tmp[V];
    __builtin_memcpy(tmp, &raw, 8);
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = Storage<S>::load(tmp[v]);
  }
}

template <class S, int V>
__device__ __forceinline__ void storev(S* __restrict__ p,
                                       const typename Storage<S>::C (&in)[V]) {
  static_assert(V * sizeof(S) == 16 || V * sizeof(S) == 8, "8- or 16-byte vectors");
  if constexpr (V * sizeof(S) == 16) {
    store16(p, in);
  } else {
    S tmp[V];
#pragma unroll
    for (int v = 0; v < V; ++v) tmp[v] = Storage<S>::store(in[v]);
    uint2 raw;
    __builtin_memcpy(&raw, tmp, 8);
    *reinterpret_cast<uint2*>(p) = raw;
  }
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Workgroups of a 1-D grid are handed to the 8 XCDs round-robin (block b runs
// on XCD b % 8).  Renumber so that each XCD receives one contiguous range of
// tiles: neighbouring tiles, which share halo rows and planes, then meet in
// the same 4 MB L2.  A bijection on [0, nblocks).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int q = nblocks / kXcds, r = nblocks % kXcds;
  const int xcd = b % kXcds, local = b / kXcds;
  const int start = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return start + local;
}

}  // namespace npbx
