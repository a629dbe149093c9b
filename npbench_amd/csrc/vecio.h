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

// n storage elements of `in` packed into 32-bit words (BF16: a conversion
// instruction per pair).
template <class S, int N>
__device__ __forceinline__ void pack_words(const typename Storage<S>::C (&in)[N],
                                           uint32_t (&w)[N * sizeof(S) / 4]) {
  if constexpr (sizeof(S) == 2) {
#pragma unroll
    for (int i = 0; i < N / 2; ++i) w[i] = f32x2_to_bf16x2_bits(in[2 * i], in[2 * i + 1]);
  } else {
    S tmp[N];
#pragma unroll
    for (int v = 0; v < N; ++v) tmp[v] = Storage<S>::store(in[v]);
    __builtin_memcpy(w, tmp, N * sizeof(S));
  }
}

template <class S>
__device__ __forceinline__ void store16(S* __restrict__ p,
                                        const typename Storage<S>::C (&in)[Vec16<S>::N]) {
  uint32_t w[4];
  pack_words<S, Vec16<S>::N>(in, w);
  *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
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

// V consecutive storage elements as raw 32-bit words, loaded in one 8- or
// 16-byte access (loadraw) and converted to the arithmetic type later
// (unpackv): a prefetched BF16 vector holds half the registers of its
// converted form, and the conversion does not wait on the load early.
template <class S, int V>
struct RawVec {
  static constexpr int kWords = V * static_cast<int>(sizeof(S)) / 4;
  uint32_t w[kWords];
};

template <class S, int V>
__device__ __forceinline__ void loadraw(const S* __restrict__ p, RawVec<S, V>& r) {
  static_assert(V * sizeof(S) == 16 || V * sizeof(S) == 8 || V * sizeof(S) == 4,
                "4-, 8- or 16-byte vectors");
  if constexpr (V * sizeof(S) == 16) {
    const uint4 x = *reinterpret_cast<const uint4*>(p);
    r.w[0] = x.x, r.w[1] = x.y, r.w[2] = x.z, r.w[3] = x.w;
  } else if constexpr (V * sizeof(S) == 8) {
    const uint2 x = *reinterpret_cast<const uint2*>(p);
    r.w[0] = x.x, r.w[1] = x.y;
  } else {
    r.w[0] = *reinterpret_cast<const uint32_t*>(p);
  }
}

template <class S, int V>
__device__ __forceinline__ void unpackv(const RawVec<S, V>& r, typename Storage<S>::C (&out)[V]) {
  S tmp[V];
  __builtin_memcpy(tmp, r.w, sizeof(tmp));
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = Storage<S>::load(tmp[v]);
}

template <class S, int V>
__device__ __forceinline__ void storev(S* __restrict__ p,
                                       const typename Storage<S>::C (&in)[V]) {
  static_assert(V * sizeof(S) == 16 || V * sizeof(S) == 8 || V * sizeof(S) == 4,
                "4-, 8- or 16-byte vectors");
  if constexpr (V * sizeof(S) == 16) {
    store16(p, in);
  } else if constexpr (V * sizeof(S) == 8) {
    uint32_t w[2];
    pack_words<S, V>(in, w);
    *reinterpret_cast<uint2*>(p) = make_uint2(w[0], w[1]);
  } else {
    uint32_t w[1];
    pack_words<S, V>(in, w);
    *reinterpret_cast<uint32_t*>(p) = w[0];
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
