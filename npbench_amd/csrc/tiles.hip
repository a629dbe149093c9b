// Temporal-blocked tile kernels for the explicit stencils, and the work-tile
// queue that runs many grids of uneven size in one persistent launch.
//
// A tile advances a block of one grid by `depth` (<= T) sweeps in one HBM
// round trip: it reads level L0 once (tile plus a halo of T cells), computes
// the intermediate levels on chip, and writes only level L0+depth of its
// central region.  Arithmetic per point is exactly the plain sweep's
// (jacobi_2d_numpy.py:7-10, heat_3d_numpy.py:7-20 order), and every level
// kept on chip is rounded through the storage type as a plain sweep's store
// would round it, so results are bitwise independent of T and of the tiling.
//
// Workgroup: 256 threads = 4 waves; lane l owns V consecutive points along
// the contiguous axis (one 16-byte vector: 4 FP32, 2 FP64, 8 BF16), so a
// wave spans KT = 64*V points of a row and the row's left/right neighbours
// cross lanes by shuffle.  Wave w owns R consecutive rows; rows r-1 / r+1
// are in the lane's own registers except at a wave's first and last row,
// which are exchanged through LDS (two rows per wave per level).
//
// heat (3-D): the tile is 4R rows (j) x KT (k) and marches along i through
// an output segment of `seg` planes plus T planes of halo on each side.
// Level l keeps a three-plane register window (i-1, i, i+1); each step loads
// one new level-0 plane and pushes it through the T levels (2.5-D spatial
// plus temporal blocking); the plane for step P+1 is loaded during step P.
// Only level-0 planes are read from HBM and only level-T planes of the
// output segment are written.
//
// jacobi (2-D): the tile is 4R rows x KT columns held in registers; the T
// levels are computed in place with an LDS edge-row exchange per level.
//
// Fixed points: points on the global boundary keep the boundary value of the
// level's parity (A for even levels, B for odd), loaded from A or B.
//
// Queue: a launch carries up to 48 grid rounds by value (RoundBatch); a
// persistent grid of workgroups pulls tile indices from an atomic counter and
// the last workgroup to find the queue empty resets it, so consecutive
// launches and graph replays need no memset and no device-side tile list.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.h"
#include "points.h"
#include "tiles.h"
#include "vecio.h"

namespace npbx {
namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
// Heat tiles: R rows per wave (a tile is 4R rows) is a tuning choice (see
// npbx_set_tuning "heat.rows"); with_shape lists the instantiated (T, R).

// Rows per wave of a Jacobi tile (R*V <= 64 so a lane's points fit one
// 64-bit fixed-point mask).
constexpr int jacobi_rows(int v) { return v >= 8 ? 8 : 8; }

// Largest depth whose register windows fit: FP64 is also limited to the one
// vector (2 points) of halo along the contiguous axis.
template <class S>
struct MaxDepth {
  static constexpr int value = sizeof(S) == 4 ? 4 : 2;
};

// Boundary source of a level: A for even levels, B for odd.  Selecting
// between two pointer values (not indexing the descriptor) keeps the
// descriptor in registers.
template <class S>
__device__ __forceinline__ const S* bnd(const S* a, const S* b, int level_parity) {
  return (level_parity & 1) ? b : a;
}

// Contiguous-axis tile start and output range for a side of length n.
template <int KT, int V>
__device__ __forceinline__ void col_range(int n, int tiles_k, int tk, int& kl, int& klo, int& khi) {
  if (tiles_k == 1 && n <= KT) {
    kl = 0;
    klo = 1;
    khi = n - 1;
  } else {
    const int wk = KT - 2 * V;
    kl = tk * wk - V;
    klo = max(kl + V, 1);
    khi = min(kl + KT - V, n - 1);
  }
}

// ---------------------------------------------------------------- heat 3-D
template <class S, int T, int R>
__device__ __forceinline__ void heat_tile(const GridRound g, int tile,
                                          typename Storage<S>::C* lds) {
  using St = Storage<S>;
  using C = typename St::C;
  constexpr int V = Vec16<S>::N;
  constexpr int KT = 64 * V;
  constexpr int JT = kWaves * R;
  const int n = g.n, d = g.depth;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tk = tile % g.tiles_k;
  const int rest = tile / g.tiles_k;
  const int tj = rest % g.tiles_j;
  const int ti = rest / g.tiles_j;
  const int wj = JT - 2 * T;
  const int jl = 1 + tj * wj - T;
  int kl, klo, khi;
  col_range<KT, V>(n, g.tiles_k, tk, kl, klo, khi);
  const int jlo = max(jl + T, 1), jhi = min(jl + T + wj, n - 1);
  const int i0 = 1 + ti * g.seg, i1 = min(i0 + g.seg, n - 1);
  const int k0 = kl + lane * V;
  const bool vlive = k0 >= 0 && k0 + V <= n;
  const int64_t plane = static_cast<int64_t>(n) * n;
  const S* X = static_cast<const S*>(g.X);
  S* Y = static_cast<S*>(g.Y);
  const S* gA = static_cast<const S*>(g.A);
  const S* gB = static_cast<const S*>(g.B);

  bool row_live[R];
  uint32_t fixed = 0;  // bit r*V+v: global boundary in j or k
  int64_t rowoff[R];
  int kb = -1, jb = -1, nkb = 0, njb = 0;  // the lane's boundary column / row
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = jl + w * R + r;
    row_live[r] = vlive && j >= 0 && j < n;
    rowoff[r] = static_cast<int64_t>(j) * n + k0;
    if (row_live[r] && (j == 0 || j == n - 1)) {
      jb = r;
      ++njb;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = k0 + v;
      if (row_live[r] && (j == 0 || j == n - 1 || k == 0 || k == n - 1)) fixed |= 1u << (r * V + v);
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v)
    if (vlive && (k0 + v == 0 || k0 + v == n - 1)) {
      kb = v;
      ++nkb;
    }
  // Lanes with one boundary column and at most one boundary row (all but
  // grids of a few points) get their boundary values of each level from
  // registers loaded one step ahead; others load them when needed.
  const bool slow = nkb > 1 || njb > 1;

  // Whole-plane read of level parity p at plane q: boundary planes come
  // from A/B; interior planes from `src` with j/k boundary points patched.
  auto read_plane = [&](C (&dst)[R][V], int q, int p, const S* src) {
    const S* b = bnd<S>(gA, gB, p);
    const bool bplane = q == 0 || q == n - 1;
    const S* from = bplane ? b : src;
    const int64_t base = static_cast<int64_t>(q) * plane;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row_live[r]) load16(from + base + rowoff[r], dst[r]);
    if (fixed && !bplane) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (fixed & (1u << (r * V + v))) dst[r][v] = St::load(b[base + rowoff[r] + v]);
    }
  };

  // Boundary-column values that step P's level l (plane P-1-l, parity
  // parity0+l+1) writes into column kb, loaded one step ahead.  (A boundary
  // row is the same for the whole wave, so only edge waves load it per level.)
  C bk[T][R], bk_nx[T][R];
  auto fetch_bnd = [&](C (&ck)[T][R], int P) {
#pragma unroll
    for (int l = 0; l < T; ++l) {
      const int q = P - 1 - l;
      if (l >= d || q < 1 || q > n - 2 || slow || kb < 0) continue;
      const S* b = bnd<S>(gA, gB, g.parity0 + l + 1);
      const int64_t base = static_cast<int64_t>(q) * plane;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row_live[r]) ck[l][r] = St::load(b[base + rowoff[r] + kb]);
    }
  };

  C win[T][3][R][V];  // per level: [0] plane q-1, [1] plane q, [2] plane q+1
  C pre[R][V];        // level-0 plane P+1, loaded one step ahead
  const bool kb_lo = kb == 0, kb_hi = kb == V - 1;
  const int pbeg = max(i0 - d, 0), pend = i1 - 1 + d;
  if (pbeg <= n - 1) read_plane(pre, pbeg, g.parity0, X);
  fetch_bnd(bk_nx, pbeg);
  // One step per plane.  (Unrolling by three to rotate the window slots
  // instead of moving them made the kernel slower: the code no longer fits
  // the instruction cache.)
  for (int P = pbeg; P <= pend; ++P) {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int v = 0; v < V; ++v) win[0][2][r][v] = pre[r][v];
#pragma unroll
    for (int l = 0; l < T; ++l)
#pragma unroll
      for (int r = 0; r < R; ++r) bk[l][r] = bk_nx[l][r];
    // the next step's loads stay in flight across this step's barriers
    if (P + 1 <= n - 1 && P + 1 <= pend) read_plane(pre, P + 1, g.parity0, X);
    if (P + 1 <= pend) fetch_bnd(bk_nx, P + 1);
#pragma unroll
    for (int l = 0; l < T; ++l) {
      if (l >= d) break;
      const int q = P - 1 - l;  // centre plane of level l; level l+1 is computed there
      C* L = lds + l * (kWaves * 2 * KT);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        L[(w * 2 + 0) * KT + lane * V + v] = win[l][1][0][v];
        L[(w * 2 + 1) * KT + lane * V + v] = win[l][1][R - 1][v];
      }
      __syncthreads();
      C nw[R][V];
      if (q >= 1 && q <= n - 2) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          C jm[V], jp[V];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            jm[v] = r > 0 ? win[l][1][r - 1][v]
                          : (w > 0 ? L[((w - 1) * 2 + 1) * KT + lane * V + v] : win[l][1][r][v]);
            jp[v] = r < R - 1 ? win[l][1][r + 1][v]
                              : (w < kWaves - 1 ? L[((w + 1) * 2) * KT + lane * V + v]
                                                : win[l][1][r][v]);
          }
          const C left = __shfl_up(win[l][1][r][V - 1], 1, 64);
          const C right = __shfl_down(win[l][1][r][0], 1, 64);
#pragma unroll
          for (int v = 0; v < V; ++v) {
            const C km = v > 0 ? win[l][1][r][v - 1] : left;
            const C kp = v < V - 1 ? win[l][1][r][v + 1] : right;
            // Round through the storage type as the plain sweep's store does
            // (identity for FP32/FP64), so results do not depend on T.
            nw[r][v] = St::load(St::store(heat_point<C>(win[l][1][r][v], win[l][2][r][v],
                                                        win[l][0][r][v], jp[v], jm[v], kp, km)));
          }
        }
        if (slow) {
          const S* b = bnd<S>(gA, gB, g.parity0 + l + 1);
          const int64_t base = static_cast<int64_t>(q) * plane;
#pragma unroll
          for (int r = 0; r < R; ++r)
#pragma unroll
            for (int v = 0; v < V; ++v)
              if (fixed & (1u << (r * V + v))) nw[r][v] = St::load(b[base + rowoff[r] + v]);
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            if (kb_lo) nw[r][0] = bk[l][r];
            if (kb_hi) nw[r][V - 1] = bk[l][r];
          }
          if (jb >= 0) {
            const S* b = bnd<S>(gA, gB, g.parity0 + l + 1);
            const int64_t base = static_cast<int64_t>(q) * plane;
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (r == jb) load16(b + base + rowoff[r], nw[r]);
          }
        }
      } else if (q == 0 || q == n - 1) {
        read_plane(nw, q, g.parity0 + l + 1, X);
      }
      // slide level l's window up one plane
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          win[l][0][r][v] = win[l][1][r][v];
          win[l][1][r][v] = win[l][2][r][v];
        }
      if (l + 1 < d) {
        // l + 1 < d <= T; the clamp only keeps the (unrolled) index in range
        constexpr int kLast = T - 1;
        const int nxt = l + 1 < kLast ? l + 1 : kLast;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int v = 0; v < V; ++v) win[nxt][2][r][v] = nw[r][v];
      } else if (q >= i0 && q < i1) {
        const int64_t base = static_cast<int64_t>(q) * plane;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = jl + w * R + r;
          if (!row_live[r] || j < jlo || j >= jhi) continue;
          S* o = Y + base + rowoff[r];
          if (k0 >= klo && k0 + V <= khi) {
            store16(o, nw[r]);
          } else {
#pragma unroll
            for (int v = 0; v < V; ++v)
              if (k0 + v >= klo && k0 + v < khi) o[v] = St::store(nw[r][v]);
          }
        }
      }
    }
    if (d == 1) __syncthreads();  // LDS of level 0 is rewritten next step
  }
}

// ------------------------------------------------- heat 3-D, wave-private tile
// FP32 storage only.  Each wave owns JR rows (j) x 256 columns (k) and needs
// nothing from the other waves: its rows include a halo of T rows on each
// side, so the tile uses no LDS and no barrier.  A workgroup's four waves
// cover 4*(JR-2T) consecutive output rows; the halo rows two waves share are
// read by both (an L2 hit for the second).
//
// Marching along i, level l keeps per row a[l], its values at its newest
// plane q, and part[l], the five-term partial sums (heat_partial) of level
// l+1 at plane q.  When level l's plane q+1 arrives (from HBM for level 0,
// computed for l >= 1) level l+1 at plane q is finished (heat_finish) and
// arrives at level l+1.  Level l's rows shrink by one per side per level, so
// the wave holds JR, JR-2, ... rows.  k neighbours across lanes come from
// full-wave DPP shifts.  Level 0's plane P+1 is loaded while levels >= 1 of
// step P are computed.
//
// Global-boundary points of level L hold A's (L even) or B's (L odd) values.
// Level 0's boundary rows and planes are read from that array instead of X;
// its boundary columns, and those of intermediate levels, are patched from a
// lane-spread register (lane r < JR: row r's column 0; lane 32 + r: its
// column n-1) loaded one step ahead; an intermediate level's boundary row is
// loaded one step ahead too (synchronously in grids of a few rows).
constexpr int kPrivateRows = 10;  // heat.rows >= this selects the wave-private tile
constexpr bool wave_private(int rows) { return rows >= kPrivateRows; }

// Raw buffer loads from one plane (num_records = the plane's bytes; the
// descriptor's third word is the gfx9 raw-buffer format).  The 32-bit vector
// offset is the lane's, the scalar offset the wave's row.
struct Rsrc {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ Rsrc plane_rsrc(const float* base, int64_t elems) {
  return {__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), static_cast<short>(0),
                                            static_cast<int>(elems * 4), 0x00020000)};
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void buf_load16(Rsrc rs, uint32_t voff, uint32_t soff, float (&d)[4]) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs.r, voff, soff, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = __uint_as_float(v[i]);
}
__device__ __forceinline__ float buf_load1(Rsrc rs, uint32_t voff, uint32_t soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs.r, voff, soff, 0));
}

// Full-wave DPP shifts (wave_shr:1, wave_shl:1): lane l receives lane l-1's
// (from_left) or lane l+1's (from_right) value; lane 0 / lane 63 receive 0,
// which only ever reaches halo or global-boundary points.
__device__ __forceinline__ float from_left(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_value(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// One tile at compile-time depth D (1 <= D <= T).  Every step computes every
// level unconditionally: planes a level produces before its inputs are valid
// (while the march fills) or past the grid (loads clamp to the last plane)
// only ever feed other such planes or boundary planes, and only level D's
// planes inside the output segment are stored.
//
// The loop's loads are unconditional, so the compiler waits for exactly the
// loads a step consumes.  Stores are whole 16-byte vectors from the lanes
// whose vector holds output points (exec-masked global stores; buffer stores
// of this loop lost data in a few lanes when two waves shared a SIMD); a
// stored vector that holds a global-boundary column carries that column's
// current value of Y, rewritten unchanged.
// Level-0 planes and all patch values are loaded AH steps ahead: with AH = 2
// into two register sets that the twice-unrolled loop uses in turn, with
// AH = 1 into one set refilled as soon as level 0 is done with it (deeper
// tiles, whose longer steps cover the latency).  The register footprint
// (over 256 VGPRs) runs one wave per SIMD.
// EQ: A's and B's global boundaries are equal (the packed columns' flag), so
// a boundary point holds the same value at every level, the one X holds: the
// level-0 plane is X's as loaded, and a produced plane's boundary points are
// the previous level's at the same plane, already in registers.  No patch
// loads, no lane reads.
template <int T, int JR, int D, int AH, bool EQ>
__device__ __forceinline__ void heat_wave_tile(const GridRound g, int tile) {
  constexpr int V = 4;
  constexpr int KT = 64 * V;
  constexpr int WJ = JR - 2 * T;  // output rows of a wave
  static_assert(WJ >= 1 && JR <= 32, "a wave holds 2T halo rows plus its output rows, <= 32");
  static_assert(D >= 1 && D <= T, "depth");
  static_assert(AH == 1 || AH == 2, "look-ahead");
  const int n = g.n;
  // the wave index, known to the compiler as wave-uniform: row bases and row
  // conditions then live in scalar registers and branches
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tk = tile % g.tiles_k;
  const int rest = tile / g.tiles_k;
  const int tj = rest % g.tiles_j;
  const int ti = rest / g.tiles_j;
  const int jl = 1 + (tj * kWaves + w) * WJ - T;  // row of register row 0
  const int jlo = jl + T, jhi = min(jl + T + WJ, n - 1);
  if (jlo >= jhi) return;  // the wave lies past the last interior row
  int kl, klo, khi;
  col_range<KT, V>(n, g.tiles_k, tk, kl, klo, khi);
  const int i0 = 1 + ti * g.seg, i1 = min(i0 + g.seg, n - 1);
  const int k0 = kl + lane * V;
  const bool vlive = k0 >= 0 && k0 + V <= n;
  // Accesses go through buffer descriptors of one plane: a wave-uniform row
  // offset (scalar) plus the lane's 32-bit byte offset within the row (lanes
  // outside the grid read the row's first vector).  Planes and rows are
  // clamped into the grid.  Lanes whose vector holds output points store it.
  const uint32_t lob = vlive ? static_cast<uint32_t>(k0) * 4u : 0u;
  const bool sto = vlive && k0 + V > klo && k0 < khi;
  const int64_t plane = static_cast<int64_t>(n) * n;
  const float* X = static_cast<const float*>(g.X);
  float* Y = static_cast<float*>(g.Y);
  const float* gA = static_cast<const float*>(g.A);
  const float* gB = static_cast<const float*>(g.B);
  const int p0 = g.parity0;
  auto bsrc = [&](int level) { return ((p0 + level) & 1) ? gB : gA; };
  auto row_bnd = [&](int r) { return jl + r == 0 || jl + r == n - 1; };
  auto plane_of = [&](const float* base, int q) {
    return plane_rsrc(base + static_cast<int64_t>(min(max(q, 0), n - 1)) * plane, plane);
  };
  auto rowoff = [&](int r) { return static_cast<uint32_t>(min(max(jl + r, 0), n - 1) * n) * 4u; };
  auto ld = [&](Rsrc rs, int r, float (&dst)[V]) { buf_load16(rs, lob, rowoff(r), dst); };

  // Boundary row among the rows intermediate levels compute (1 .. JR-2).
  int jb = -1, njb = 0;
#pragma unroll
  for (int r = 1; r < JR - 1; ++r)
    if (row_bnd(r)) {
      jb = r;
      ++njb;
    }
  const bool slow = njb > 1;  // grids of a few rows: load boundary rows when needed
  const int jbr = jb >= 0 ? jb : 0;  // the row the boundary-row prefetch reads

  // Boundary columns: lane-spread values (lane r < JR: row r's column 0;
  // lane 32 + r: its column n-1), patched into the lanes that hold them.
  // They come from the packed copy when there is one (the wave's values of a
  // plane are 2 x JR consecutive floats), else from the arrays themselves
  // (one cache line per value).
  const int kr = lane & 31, side = lane >= 32 ? 1 : 0;
  const int krow = min(max(jl + kr, 0), n - 1);
  const bool is_k0 = vlive && k0 == 0, is_kn = vlive && k0 + V == n;
  const float* cols = static_cast<const float*>(g.cols);
  auto load_cols = [&](const float* base, int q) {
    const int64_t qq = min(max(q, 0), n - 1);
    const int64_t arr = base == gB ? 1 : 0;
    const float* p = cols ? cols + ((arr * n + qq) * n + krow) * 2 + side
                          : base + qq * plane + static_cast<int64_t>(krow) * n + (side ? n - 1 : 0);
    return *p;
  };
  auto patch_k = [&](float (&x)[JR][V], float kb, int rlo, int rhi) {
#pragma unroll
    for (int r = 0; r < JR; ++r) {
      if (r < rlo || r >= rhi) continue;
      const float vlo = lane_value(kb, r), vhi = lane_value(kb, 32 + r);
      if (is_k0) x[r][0] = vlo;
      if (is_kn) x[r][V - 1] = vhi;
    }
  };
  // EQ: boundary columns and rows of a produced plane from level l's plane
  auto keep_k = [&](float (&x)[JR][V], const float (&src)[JR][V], int rlo, int rhi) {
#pragma unroll
    for (int r = rlo; r < rhi; ++r) {
      if (is_k0) x[r][0] = src[r][0];
      if (is_kn) x[r][V - 1] = src[r][V - 1];
    }
  };
  auto keep_rows = [&](float (&x)[JR][V], const float (&src)[JR][V], int rlo, int rhi) {
#pragma unroll
    for (int r = rlo; r < rhi; ++r)
      if (row_bnd(r))
#pragma unroll
        for (int v = 0; v < V; ++v) x[r][v] = src[r][v];
  };

  // What step S consumes: level-0 plane S and its boundary columns; level
  // m's boundary columns and row at plane S - m (1 <= m < D); Y's boundary
  // columns at the stored plane S - D.
  struct Ahead {
    float pre[JR][V];
    float kq[T];
    float jrow[T][V];
    float ky;
  };
  auto fetch = [&](int S, Ahead& f) {
    if constexpr (EQ) {
      const Rsrc rx = plane_of(X, S);
#pragma unroll
      for (int r = 0; r < JR; ++r) ld(rx, r, f.pre[r]);
      return;
    }
    const float* b = bsrc(0);
    const bool bplane = S <= 0 || S >= n - 1;
    const Rsrc rx = plane_of(bplane ? b : X, S), rb = plane_of(b, S);
#pragma unroll
    for (int r = 0; r < JR; ++r) ld(row_bnd(r) ? rb : rx, r, f.pre[r]);
    f.kq[0] = load_cols(b, S);
#pragma unroll
    for (int m = 1; m < D; ++m) {
      f.kq[m] = load_cols(bsrc(m), S - m);
      // without a boundary row this reads a line the level-0 loads just read
      ld(jb >= 0 ? plane_of(bsrc(m), S - m) : rx, jbr, f.jrow[m]);
    }
    f.ky = load_cols(Y, S - D);
  };

  float a[T][JR][V];  // level l at its newest plane (rows l+1 .. JR-l-2)
  float part[T][JR][V];
#pragma unroll
  for (int l = 0; l < T; ++l)
#pragma unroll
    for (int r = 0; r < JR; ++r)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        a[l][r][v] = 0.0f;
        part[l][r][v] = 0.0f;
      }

  // One step: consumes f (loaded AH steps ago) and refills it for step P+AH
  // as soon as level 0 is done with it.  Level 0 computes straight from f.
  auto step = [&](int P, Ahead& f) {
    if constexpr (!EQ) patch_k(f.pre, f.kq[0], 0, JR);
    // patch values used after the refill below (levels 2 and up, and Y's)
    const float ky = f.ky;
    float kq[T], jrow[T][V];
#pragma unroll
    for (int m = 1; m < T; ++m) {
      kq[m] = f.kq[m];
#pragma unroll
      for (int v = 0; v < V; ++v) jrow[m][v] = f.jrow[m][v];
    }
    float lv[JR][V];  // the arriving plane of level 1 and up
#pragma unroll
    for (int l = 0; l < D; ++l) {
      float (&in)[JR][V] = l == 0 ? f.pre : lv;  // the arriving plane of level l
      const int m = l + 1;  // the level produced, at plane q
      const int q = P - m;
      float o[JR][V];
#pragma unroll
      for (int r = m; r < JR - m; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v) o[r][v] = heat_finish(part[l][r][v], in[r][v], a[l][r][v]);
      if (m == D) {
        // level l's plane q (a[l] before the update below) carries the
        // boundary columns Y holds
        if constexpr (EQ)
          keep_k(o, a[l], T, JR - T);
        else
          patch_k(o, ky, T, JR - T);
        const bool inseg = q >= i0 && q < i1;
        float* yq = Y + static_cast<int64_t>(inseg ? q : i0) * plane;
#pragma unroll
        for (int r = T; r < JR - T; ++r)
          if (inseg && jl + r < jhi && sto)
            store16(reinterpret_cast<float*>(reinterpret_cast<char*>(yq) + rowoff(r) + lob), o[r]);
      } else if constexpr (EQ) {
        if (q == 0 || q == n - 1) {
#pragma unroll
          for (int r = m; r < JR - m; ++r)
#pragma unroll
            for (int v = 0; v < V; ++v) o[r][v] = a[l][r][v];
        } else {
          keep_k(o, a[l], m, JR - m);
          keep_rows(o, a[l], m, JR - m);
        }
      } else if (q == 0 || q == n - 1) {
        // a boundary plane of an intermediate level (first and last tiles
        // along i).  The wait for every load on this rare path keeps the
        // compiler from waiting for these loads on the common one.
        const Rsrc rb = plane_of(bsrc(m), q);
#pragma unroll
        for (int r = m; r < JR - m; ++r) ld(rb, r, o[r]);
        __builtin_amdgcn_s_waitcnt(0);
      } else {
        patch_k(o, kq[m < T ? m : T - 1], m, JR - m);
        if (slow) {
          const Rsrc rb = plane_of(bsrc(m), q);
#pragma unroll
          for (int r = m; r < JR - m; ++r)
            if (row_bnd(r)) ld(rb, r, o[r]);
          __builtin_amdgcn_s_waitcnt(0);
        } else {
#pragma unroll
          for (int r = m; r < JR - m; ++r)
            if (r == jb)
#pragma unroll
              for (int v = 0; v < V; ++v) o[r][v] = jrow[m < T ? m : T - 1][v];
        }
      }
      // level l's arrival (plane P - l) becomes its newest plane
#pragma unroll
      for (int r = m; r < JR - m; ++r) {
        const float left = from_left(in[r][V - 1]);
        const float right = from_right(in[r][0]);
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const float km = v > 0 ? in[r][v - 1] : left;
          const float kp = v < V - 1 ? in[r][v + 1] : right;
          part[l][r][v] = heat_partial(a[l][r][v], in[r - 1][v], in[r + 1][v], km, kp);
          a[l][r][v] = in[r][v];
        }
      }
      if (l == 0) fetch(P + AH, f);
      if (m < D) {
#pragma unroll
        for (int r = m; r < JR - m; ++r)
#pragma unroll
          for (int v = 0; v < V; ++v) lv[r][v] = o[r][v];
      }
    }
  };

  const int pbeg = max(i0 - D, 0), pfin = i1 - 1 + D;
  if constexpr (AH == 2) {
    Ahead fa, fb;
    fetch(pbeg, fa);
    fetch(pbeg + 1, fb);
    for (int P = pbeg; P <= pfin; P += 2) {
      step(P, fa);
      if (P + 1 > pfin) break;
      step(P + 1, fb);
    }
  } else {
    Ahead fa;
    fetch(pbeg, fa);
    for (int P = pbeg; P <= pfin; ++P) step(P, fa);
  }
}

// Dispatches a round's runtime depth to the compile-time one.
template <int T, int JR, bool EQ>
__device__ __forceinline__ void heat_wave_depth(const GridRound g, int tile) {
  constexpr int AH = T >= 3 ? 1 : 2;
  if constexpr (T >= 2) {
    if (g.depth == 2) return heat_wave_tile<T, JR, 2, AH, EQ>(g, tile);
  }
  if constexpr (T >= 3) {
    if (g.depth == 3) return heat_wave_tile<T, JR, 3, AH, EQ>(g, tile);
  }
  heat_wave_tile<T, JR, 1, AH, EQ>(g, tile);
}

// Grids whose packed columns record equal boundaries take the tile that
// needs no patches.
template <int T, int JR>
__device__ __forceinline__ void heat_wave_round(const GridRound g, int tile) {
  const int* eqflag = g.cols ? reinterpret_cast<const int*>(static_cast<const float*>(g.cols) +
                                                             4 * static_cast<int64_t>(g.n) * g.n)
                             : nullptr;
  if (eqflag && *eqflag) return heat_wave_depth<T, JR, true>(g, tile);
  heat_wave_depth<T, JR, false>(g, tile);
}

// ---------------------------------------------------------------- jacobi 2-D
template <class S, int T>
__device__ __forceinline__ void jacobi_tile(const GridRound g, int tile,
                                            typename Storage<S>::C* lds) {
  using St = Storage<S>;
  using C = typename St::C;
  constexpr int V = Vec16<S>::N;
  constexpr int R = jacobi_rows(V);
  constexpr int KT = 64 * V;
  constexpr int JT = kWaves * R;
  const int n = g.n, d = g.depth;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tk = tile % g.tiles_k;
  const int tj = tile / g.tiles_k;
  const int wj = JT - 2 * T;
  const int il = 1 + tj * wj - T;  // first row of the tile
  int kl, klo, khi;
  col_range<KT, V>(n, g.tiles_k, tk, kl, klo, khi);
  const int ilo = max(il + T, 1), ihi = min(il + T + wj, n - 1);
  const int k0 = kl + lane * V;
  const bool vlive = k0 >= 0 && k0 + V <= n;
  const S* X = static_cast<const S*>(g.X);
  S* Y = static_cast<S*>(g.Y);
  const S* gA = static_cast<const S*>(g.A);
  const S* gB = static_cast<const S*>(g.B);

  bool row_live[R];
  uint64_t fixed = 0;  // bit r*V+v (R*V <= 64)
  uint64_t brow = 0;   // rows that are boundary rows
  int kb = -1, jb = -1, nkb = 0, njb = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = il + w * R + r;
    row_live[r] = vlive && i >= 0 && i < n;
    if (row_live[r] && (i == 0 || i == n - 1)) {
      brow |= 1ull << r;
      jb = r;
      ++njb;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = k0 + v;
      if (row_live[r] && (i == 0 || i == n - 1 || k == 0 || k == n - 1))
        fixed |= 1ull << (r * V + v);
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v)
    if (vlive && (k0 + v == 0 || k0 + v == n - 1)) {
      kb = v;
      ++nkb;
    }
  const bool slow = nkb > 1 || njb > 1;
  auto off = [&](int r) { return static_cast<int64_t>(il + w * R + r) * n + k0; };
  // Boundary column / row values of both parities, loaded once per tile.
  C kbv[2][R], jbv[2][V];
  if (!slow) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const S* b = p ? gB : gA;
      if (kb >= 0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (row_live[r]) kbv[p][r] = St::load(b[off(r) + kb]);
      }
      if (jb >= 0) load16(b + off(jb), jbv[p]);
    }
  }

  C cur[R][V];
  {
    const S* b = bnd<S>(gA, gB, g.parity0);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row_live[r]) load16(((brow >> r) & 1) ? b + off(r) : X + off(r), cur[r]);
    if (fixed) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v)
          if ((fixed >> (r * V + v)) & 1) cur[r][v] = St::load(b[off(r) + v]);
    }
  }
  for (int l = 0; l < d; ++l) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      lds[(w * 2 + 0) * KT + lane * V + v] = cur[0][v];
      lds[(w * 2 + 1) * KT + lane * V + v] = cur[R - 1][v];
    }
    __syncthreads();
    // Rows are updated in place top to bottom: row r's north neighbour is
    // the saved old row r-1, its south neighbour cur[r+1] is still old.
    const S* b = bnd<S>(gA, gB, g.parity0 + l + 1);
    C prev[V];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C north[V], south[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        north[v] = r > 0 ? prev[v]
                         : (w > 0 ? lds[((w - 1) * 2 + 1) * KT + lane * V + v] : cur[r][v]);
        south[v] = r < R - 1 ? cur[r + 1][v]
                             : (w < kWaves - 1 ? lds[((w + 1) * 2) * KT + lane * V + v] : cur[r][v]);
      }
      const C left = __shfl_up(cur[r][V - 1], 1, 64);
      const C right = __shfl_down(cur[r][0], 1, 64);
      C nw[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const C west = v > 0 ? cur[r][v - 1] : left;
        const C east = v < V - 1 ? cur[r][v + 1] : right;
        // rounded through the storage type, as the plain sweep's store does
        nw[v] = St::load(St::store(jacobi_point<C>(cur[r][v], west, east, south[v], north[v])));
        if (slow) {
          if ((fixed >> (r * V + v)) & 1) nw[v] = St::load(b[off(r) + v]);
        } else {
          const int p = (g.parity0 + l + 1) & 1;
          if (v == kb) nw[v] = p ? kbv[1][r] : kbv[0][r];
          if (r == jb) nw[v] = p ? jbv[1][v] : jbv[0][v];
        }
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        prev[v] = cur[r][v];
        cur[r][v] = nw[v];
      }
    }
    __syncthreads();  // every wave has read the edge rows before they change
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = il + w * R + r;
    if (!row_live[r] || i < ilo || i >= ihi) continue;
    S* o = Y + off(r);
    if (k0 >= klo && k0 + V <= khi) {
      store16(o, cur[r]);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (k0 + v >= klo && k0 + v < khi) o[v] = St::store(cur[r][v]);
    }
  }
}

template <class S, int T, int RH>
struct TileLds {
  using C = typename Storage<S>::C;
  static constexpr int V = Vec16<S>::N;
  static constexpr int KT = 64 * V;
  static constexpr int HEAT = wave_private(RH) ? 0 : T * kWaves * 2 * KT;
  static constexpr int JACOBI = kWaves * 2 * KT;  // one exchange buffer
  static constexpr int ELEMS = HEAT > JACOBI ? HEAT : JACOBI;
};

// Tile t of a launch belongs to the round g with cum[g] <= t < cum[g+1].
__device__ __forceinline__ int round_of(const RoundBatch& b, int t) {
  int lo = 0, hi = b.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.cum[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <class S, int T, int RH>
__device__ __forceinline__ void run_tile(const RoundBatch& b, int t,
                                         typename Storage<S>::C* lds) {
  const int g = round_of(b, t);
  const GridRound r = b.round[g];
  if (r.kind == kHeat3D) {
    if constexpr (wave_private(RH))
      heat_wave_round<T, RH>(r, t - b.cum[g]);
    else
      heat_tile<S, T, RH>(r, t - b.cum[g], lds);
  } else {
    jacobi_tile<S, T>(r, t - b.cum[g], lds);
  }
}

// queue == nullptr: one workgroup per tile (tile = blockIdx.x).  Otherwise a
// persistent workgroup pulls tile indices from `queue` until they run out.
template <class S, int T, int RH>
__global__ __launch_bounds__(kThreads) void stencil_tiles(const RoundBatch batch,
                                                          QueueState* queue) {
  using C = typename Storage<S>::C;
  __shared__ C lds[TileLds<S, T, RH>::ELEMS];
  const int ntiles = batch.cum[batch.count];
  if (queue == nullptr) {
    run_tile<S, T, RH>(batch, blockIdx.x, lds);
    return;
  }
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(&queue->next, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= ntiles) break;
    run_tile<S, T, RH>(batch, t, lds);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue->done, 1) == static_cast<int>(gridDim.x) - 1) {
      atomicExch(&queue->next, 0);
      atomicExch(&queue->done, 0);
    }
  }
}

// ------------------------------------------------------------- host planning
// Tuning knobs (npbx_set_tuning): heat rows per wave and default heat
// planes per tile.  Read when a launch is planned.
std::atomic<int> g_heat_rows{4};
std::atomic<int> g_heat_seg{32};
std::atomic<int> g_sweep_shape{0};

// Packed boundary columns of a heat grid: cols[((a*n + q)*n + j)*2 + s] =
// (a ? B : A)[q][j][s ? n-1 : 0].  The FP32 wave-private tile reads its
// boundary patches from here: 2 x rows consecutive floats per wave and plane
// instead of one cache line per value.  The word after the columns holds 1
// when A's and B's global boundaries (all six faces) are bitwise equal, as
// when B = A.copy(); the tile then takes every boundary value from its own
// registers (the value of any level there is the same).
__global__ void pack_flag_kernel(int* __restrict__ flag) { *flag = 1; }

__global__ __launch_bounds__(256) void pack_cols_kernel(const float* __restrict__ A,
                                                        const float* __restrict__ B,
                                                        float* __restrict__ cols, int n) {
  const int64_t nn = static_cast<int64_t>(n) * n;
  int* flag = reinterpret_cast<int*>(cols + 4 * nn);
  unsigned same = 1;  // bitwise and of every comparison: no branch between the loads
  auto eq = [](float x, float y) { return __float_as_uint(x) == __float_as_uint(y) ? 1u : 0u; };
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  // the two boundary columns of every row, packed (and compared)
  for (int64_t t = t0; t < nn; t += stride) {
    const int64_t row = t * n;  // start of row j of plane q, t = q*n + j
    const float a0 = A[row], a1 = A[row + n - 1], b0 = B[row], b1 = B[row + n - 1];
    cols[t * 2] = a0;
    cols[t * 2 + 1] = a1;
    cols[(nn + t) * 2] = b0;
    cols[(nn + t) * 2 + 1] = b1;
    same = same & eq(a0, b0) & eq(a1, b1);
  }
  // the rest of the boundary faces, element by element along k (coalesced):
  // planes 0 and n-1 whole, rows 0 and n-1 of the planes between
  const int64_t planes = 2 * nn, rows = 2 * static_cast<int64_t>(n - 2) * n;
  for (int64_t e = t0; e < planes + rows; e += stride) {
    int64_t at;
    if (e < planes) {
      at = (e < nn ? 0 : (static_cast<int64_t>(n) - 1) * nn) + e % nn;
    } else {
      const int64_t r = e - planes, q = 1 + r / (2 * static_cast<int64_t>(n));
      const int64_t side = (r / n) % 2, k = r % n;
      at = q * nn + (side ? static_cast<int64_t>(n - 1) * n : 0) + k;
    }
    same = same & eq(A[at], B[at]);
  }
  if (!same) *flag = 0;  // every writer stores the same value
}

int64_t cols_elems(int64_t n) { return 4 * n * n + 1; }

int pack_cols(const void* A, const void* B, int n, void* cols, hipStream_t stream) {
  const int64_t nn = static_cast<int64_t>(n) * n;
  int* flag = reinterpret_cast<int*>(static_cast<float*>(cols) + 4 * nn);
  hipLaunchKernelGGL(pack_flag_kernel, dim3(1), dim3(1), 0, stream, flag);
  NPBX_LAUNCHED("pack_flag");
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(cdiv(4 * nn, 256), 4096));
  hipLaunchKernelGGL(pack_cols_kernel, dim3(blocks), dim3(256), 0, stream,
                     static_cast<const float*>(A), static_cast<const float*>(B),
                     static_cast<float*>(cols), n);
  NPBX_LAUNCHED("pack_cols");
  return NPBX_OK;
}

// Whether a grid's rounds run the wave-private tile (and so use packed columns).
bool uses_cols(int kind, int dtype, int rows) {
  return kind == kHeat3D && dtype == NPBX_F32 && wave_private(rows);
}

int max_depth(int dtype) {
  switch (dtype) {
    case NPBX_F64: return MaxDepth<double>::value;
    case NPBX_F32: return MaxDepth<float>::value;
    case NPBX_BF16: return MaxDepth<bf16>::value;
    default: return 0;
  }
}

int vec_elems(int dtype) { return static_cast<int>(16 / dtype_size(dtype)); }

// Fills the tiling fields of a GridRound for side n and max depth T.
void tile_grid(GridRound& g, int kind, int n, int dtype, int T, int seg) {
  const int V = vec_elems(dtype);
  g.rows = g_heat_rows.load();
  const int KT = 64 * V;
  g.tiles_k = n <= KT ? 1 : static_cast<int>(cdiv(n, KT - 2 * V));
  if (kind == kHeat3D) {
    // a shared tile is 4*rows rows with one halo; a wave-private one is four
    // waves of rows-2T output rows each
    const int out_rows = wave_private(g.rows) ? kWaves * (g.rows - 2 * T) : kWaves * g.rows - 2 * T;
    g.tiles_j = static_cast<int>(cdiv(n - 2, out_rows));
    g.seg = seg;
  } else {
    g.tiles_j = static_cast<int>(cdiv(n - 2, kWaves * jacobi_rows(V) - 2 * T));
    g.seg = 0;
  }
}

int64_t tiles_per_round(const GridRound& g) {
  int64_t t = int64_t(g.tiles_k) * g.tiles_j;
  if (g.kind == kHeat3D) t *= cdiv(g.n - 2, g.seg);
  return t;
}

// Depths of the rounds that advance nsweeps sweeps so that, as in the
// reference, level first+nsweeps ends in its home array and level
// first+nsweeps-1 in the other: rounds of depth <= T cover nsweeps-1 sweeps,
// their count has the parity of nsweeps-1 (each round flips the array the
// state lives in), and a final depth-1 round performs the last sweep.
std::vector<int> round_depths(int64_t nsweeps, int T) {
  std::vector<int> out;
  if (nsweeps <= 0) return out;
  const int64_t m = nsweeps - 1;
  int64_t R = cdiv(m, T);
  if ((R & 1) != (m & 1)) ++R;
  for (int64_t r = 0; r < R; ++r) out.push_back(static_cast<int>(m / R + (r < m % R ? 1 : 0)));
  out.push_back(1);
  return out;
}

template <class S, int T, int RH>
int launch_tiles(const RoundBatch& b, QueueState* queue, int nblocks, hipStream_t stream) {
  hipLaunchKernelGGL((stencil_tiles<S, T, RH>), dim3(static_cast<unsigned>(nblocks)),
                     dim3(kThreads), 0, stream, b, queue);
  NPBX_LAUNCHED("stencil_tiles");
  return NPBX_OK;
}

// Calls f.template operator()<T, RH>() for an instantiated (T, RH) pair.
template <class S, class F>
int with_shape(int T, int RH, F&& f) {
  if constexpr (MaxDepth<S>::value >= 2) {
    if (T == 2 && RH == 2) return f.template operator()<2, 2>();
    if (T == 2 && RH == 4) return f.template operator()<2, 4>();
  }
  if constexpr (sizeof(S) >= 4) {  // BF16 windows with 6 or 8 rows spill
    if (T == 2 && RH == 6) return f.template operator()<2, 6>();
  }
  if constexpr (sizeof(S) == 4) {  // and FP64 ones with 8
    if (T == 2 && RH == 8) return f.template operator()<2, 8>();
  }
  if constexpr (MaxDepth<S>::value >= 3) {
    if (T == 3 && RH == 4) return f.template operator()<3, 4>();
    if (T == 3 && RH == 6) return f.template operator()<3, 6>();
  }
  if constexpr (MaxDepth<S>::value >= 4) {
    if (T == 4 && RH == 4) return f.template operator()<4, 4>();
  }
  if constexpr (std::is_same_v<S, float>) {  // wave-private heat tiles
    if (T == 2 && RH == 10) return f.template operator()<2, 10>();
    if (T == 2 && RH == 12) return f.template operator()<2, 12>();
    if (T == 3 && RH == 12) return f.template operator()<3, 12>();
    if (T == 3 && RH == 14) return f.template operator()<3, 14>();
  }
  return fail(NPBX_EUNSUPPORTED, "tile depth T=%d with %d heat rows per wave is not built for "
              "this dtype", T, RH);
}

int launch_any(int dtype, int T, int RH, const RoundBatch& b, QueueState* queue, int nblocks,
               hipStream_t stream) {
  return dispatch_dtype(dtype, [&](auto tag) {
    using S = decltype(tag);
    return with_shape<S>(T, RH, [&]<int TT, int RR>() {
      return launch_tiles<S, TT, RR>(b, queue, nblocks, stream);
    });
  });
}

// Workgroups that fit on the device at once for the persistent grid.
int resident_blocks(int dtype, int T, int RH) {
  int dev = 0, cus = 0, per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  dispatch_dtype(dtype, [&](auto tag) {
    using S = decltype(tag);
    return with_shape<S>(T, RH, [&]<int TT, int RR>() {
      int b = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, stencil_tiles<S, TT, RR>, kThreads,
                                                       0) == hipSuccess &&
          b > 0)
        per_cu = b;
      return NPBX_OK;
    });
  });
  return cus * per_cu;
}

}  // namespace

int sweep_shape() { return g_sweep_shape.load(); }

std::vector<int> sweep_rounds(int64_t nsweeps, int T) { return round_depths(nsweeps, T); }

// Single grid, temporal-blocked: rounds of depth <= T through the tile
// kernel (one workgroup per tile), then the final depth-1 round.
int tiled_sweeps(int kind, int dtype, void* A, void* B, int n, int64_t first, int64_t nsweeps,
                 int T, int seg, hipStream_t stream) {
  const int tmax = max_depth(dtype);
  NPBX_REQUIRE(T >= 2 && T <= tmax, "tblock=%d outside [2, %d] for %s", T, tmax,
               dtype_name(dtype));
  NPBX_REQUIRE(n % vec_elems(dtype) == 0 && aligned16(A) && aligned16(B),
               "temporal blocking needs rows of whole 16-byte vectors");
  GridRound g{};
  g.A = A;
  g.B = B;
  g.n = n;
  g.kind = kind;
  tile_grid(g, kind, n, dtype, T, seg > 0 ? seg : g_heat_seg.load());
  if (uses_cols(kind, dtype, g.rows)) {
    void* ws = nullptr;
    NPBX_TRY(workspace(static_cast<size_t>(cols_elems(n)) * sizeof(float), &ws));
    NPBX_TRY(pack_cols(A, B, n, ws, stream));
    g.cols = ws;
  }
  const int64_t ntiles = tiles_per_round(g);
  NPBX_REQUIRE(ntiles < (int64_t(1) << 31), "too many tiles");
  int64_t level = first;
  // Round 1 reads the array that holds level `first` in the reference; each
  // round then writes the other array and the next round reads it.
  void* X = (first & 1) ? B : A;
  RoundBatch b{};
  b.count = 1;
  b.cum[0] = 0;
  b.cum[1] = static_cast<int32_t>(ntiles);
  for (int depth : round_depths(nsweeps, T)) {
    g.X = X;
    g.Y = X == A ? B : A;
    g.parity0 = static_cast<int>(level & 1);
    g.depth = depth;
    b.round[0] = g;
    NPBX_TRY(launch_any(dtype, T, g.rows, b, nullptr, static_cast<int>(ntiles), stream));
    X = g.Y;
    level += depth;
  }
  return NPBX_OK;
}

}  // namespace npbx

// ------------------------------------------------------------------- C ABI
using namespace npbx;

namespace {

// Validates one grid of a batch and fills the fixed part of its GridRound.
int check_grid(const char* who, int i, int32_t kind, int32_t dtype, int64_t n, void* A, void* B,
               int want_dtype, GridRound& g) {
  NPBX_REQUIRE(dtype == want_dtype, "%s: grid %d dtype differs (one dtype per call)", who, i);
  NPBX_REQUIRE(kind == NPBX_KIND_JACOBI2D || kind == NPBX_KIND_HEAT3D,
               "%s: grid %d has unknown kind %d", who, i, kind);
  const int V = vec_elems(dtype);
  NPBX_REQUIRE(n >= 3 && n % V == 0,
               "%s: grid %d side %lld must be >= 3 and a multiple of %d", who, i,
               static_cast<long long>(n), V);
  NPBX_REQUIRE(n <= (kind == NPBX_KIND_HEAT3D ? 2048 : 46340), "%s: grid %d too large", who, i);
  NPBX_REQUIRE(A && B && A != B && aligned16(A) && aligned16(B),
               "%s: grid %d needs distinct 16-byte aligned A and B", who, i);
  g = GridRound{};
  g.A = A;
  g.B = B;
  g.n = static_cast<int32_t>(n);
  g.kind = kind == NPBX_KIND_HEAT3D ? kHeat3D : kJacobi2D;
  return NPBX_OK;
}

// Packs grid rounds into launches of at most kMaxBatch rounds.
struct Packer {
  std::vector<RoundBatch> out;
  int64_t tiles = 0;
  void add(const GridRound& g) {
    if (out.empty() || out.back().count == kMaxBatch) {
      out.emplace_back();
      RoundBatch& b = out.back();
      std::memset(&b, 0, sizeof(b));
    }
    RoundBatch& b = out.back();
    const int64_t nt = tiles_per_round(g);
    b.round[b.count] = g;
    b.cum[b.count + 1] = b.cum[b.count] + static_cast<int32_t>(nt);
    ++b.count;
    tiles += nt;
  }
};

int64_t grid_lups(const GridRound& g, int64_t sweeps) {
  const int64_t inner = g.n - 2;
  return sweeps * (g.kind == kHeat3D ? inner * inner * inner : inner * inner);
}

struct ColsPack {
  const void* A;
  const void* B;
  int n;
  int64_t off;  // elements into Plan::d_cols
};

struct Plan {
  int dtype = 0;
  int T = 0;
  int rows = 4;
  int nblocks = 0;
  std::vector<RoundBatch> launches;
  std::vector<ColsPack> packs;  // grids whose boundary columns each launch packs
  float* d_cols = nullptr;
  QueueState* d_queue = nullptr;
  int64_t ntiles = 0, lups = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int plan_free(Plan* p) {
  if (!p) return NPBX_OK;
  // Release paths ignore errors: the plan is gone either way.
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_queue) (void)hipFree(p->d_queue);
  if (p->d_cols) (void)hipFree(p->d_cols);
  delete p;
  return NPBX_OK;
}

int plan_enqueue(Plan* p, hipStream_t s) {
  // A and B may have been rewritten since the last launch: pack their
  // boundary columns afresh (one small kernel per grid, before the rounds).
  for (const ColsPack& c : p->packs) NPBX_TRY(pack_cols(c.A, c.B, c.n, p->d_cols + c.off, s));
  for (const RoundBatch& b : p->launches) {
    const int nt = b.cum[b.count];
    if (nt == 0) continue;
    NPBX_TRY(launch_any(p->dtype, p->T, p->rows, b, p->d_queue, std::min(nt, p->nblocks), s));
  }
  return NPBX_OK;
}

int new_queue(QueueState** q) {
  NPBX_HIP(hipMalloc(reinterpret_cast<void**>(q), sizeof(QueueState)));
  NPBX_HIP(hipMemset(*q, 0, sizeof(QueueState)));
  return NPBX_OK;
}

}  // namespace

extern "C" int npbx_plan_create(const npbx_grid_spec* specs, int32_t count, int32_t tblock,
                                int32_t seg, void** plan) {
  NPBX_REQUIRE(plan, "plan_create: null output pointer");
  *plan = nullptr;
  NPBX_REQUIRE(specs && count > 0, "plan_create: no grids");
  const int dtype = specs[0].dtype;
  NPBX_REQUIRE(dtype_size(dtype) != 0, "plan_create: bad dtype %d", dtype);
  const int tmax = max_depth(dtype);
  NPBX_REQUIRE(tblock >= 2 && tblock <= tmax, "plan_create: tblock=%d outside [2, %d] for %s",
               tblock, tmax, dtype_name(dtype));
  if (seg <= 0) seg = g_heat_seg.load();
  std::vector<GridRound> base(count);
  std::vector<std::vector<int>> depths(count);
  size_t nrounds = 0;
  int64_t lups = 0;
  for (int i = 0; i < count; ++i) {
    const npbx_grid_spec& s = specs[i];
    NPBX_TRY(check_grid("plan_create", i, s.kind, s.dtype, s.n, s.A, s.B, dtype, base[i]));
    NPBX_REQUIRE(s.first >= 0 && s.nsweeps >= 0, "plan_create: grid %d negative sweeps", i);
    tile_grid(base[i], base[i].kind, base[i].n, dtype, tblock, seg);
    depths[i] = round_depths(s.nsweeps, tblock);
    nrounds = std::max(nrounds, depths[i].size());
    lups += grid_lups(base[i], s.nsweeps);
  }
  // Launch group r runs round r of every grid that has one.
  auto p = new Plan();
  p->dtype = dtype;
  p->T = tblock;
  p->lups = lups;
  int64_t cols_total = 0;
  for (int i = 0; i < count; ++i)
    if (uses_cols(base[i].kind, dtype, base[i].rows)) {
      p->packs.push_back({specs[i].A, specs[i].B, base[i].n, cols_total});
      cols_total += (cols_elems(base[i].n) + 63) / 64 * 64;  // 256-byte aligned
    }
  if (cols_total > 0) {
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->d_cols),
                                   static_cast<size_t>(cols_total) * sizeof(float));
    if (e != hipSuccess) {
      plan_free(p);
      return fail(NPBX_ENOMEM, "plan_create: %s", hipGetErrorString(e));
    }
    size_t k = 0;
    for (int i = 0; i < count; ++i)
      if (uses_cols(base[i].kind, dtype, base[i].rows)) base[i].cols = p->d_cols + p->packs[k++].off;
  }
  std::vector<int64_t> level(count);
  std::vector<void*> X(count);
  for (int i = 0; i < count; ++i) {
    level[i] = specs[i].first;
    X[i] = (level[i] & 1) ? specs[i].B : specs[i].A;
  }
  for (size_t r = 0; r < nrounds; ++r) {
    Packer pk;
    for (int i = 0; i < count; ++i) {
      if (r >= depths[i].size()) continue;
      GridRound g = base[i];
      g.X = X[i];
      g.Y = X[i] == specs[i].A ? specs[i].B : specs[i].A;
      g.parity0 = static_cast<int32_t>(level[i] & 1);
      g.depth = depths[i][r];
      pk.add(g);
      X[i] = g.Y;
      level[i] += g.depth;
    }
    for (auto& b : pk.out) p->launches.push_back(b);
    p->ntiles += pk.tiles;
  }
  p->rows = base[0].rows;
  p->nblocks = resident_blocks(dtype, tblock, p->rows);
  const int rc = new_queue(&p->d_queue);
  if (rc != NPBX_OK) {
    plan_free(p);
    return rc;
  }
  *plan = p;
  return NPBX_OK;
}

extern "C" int npbx_plan_launch(void* plan, void* stream) {
  NPBX_REQUIRE(plan, "plan_launch: null plan");
  Plan* p = static_cast<Plan*>(plan);
  hipStream_t s = as_stream(stream);
  if (p->exec) {
    NPBX_HIP(hipGraphLaunch(p->exec, s));
    return NPBX_OK;
  }
  return plan_enqueue(p, s);
}

extern "C" int npbx_plan_capture(void* plan, void* stream) {
  NPBX_REQUIRE(plan, "plan_capture: null plan");
  Plan* p = static_cast<Plan*>(plan);
  NPBX_REQUIRE(stream, "plan_capture: capture needs a non-default stream");
  hipStream_t s = as_stream(stream);
  if (p->exec) return NPBX_OK;
  NPBX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = plan_enqueue(p, s);
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &graph);
  NPBX_TRY(rc);
  NPBX_HIP(end);
  p->graph = graph;
  NPBX_HIP(hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0));
  return NPBX_OK;
}

extern "C" int npbx_plan_info(void* plan, int64_t* launches, int64_t* tiles, int64_t* lups) {
  NPBX_REQUIRE(plan, "plan_info: null plan");
  const Plan* p = static_cast<const Plan*>(plan);
  if (launches) *launches = static_cast<int64_t>(p->launches.size());
  if (tiles) *tiles = p->ntiles;
  if (lups) *lups = p->lups;
  return NPBX_OK;
}

extern "C" int npbx_plan_destroy(void* plan) {
  if (!plan) return NPBX_OK;
  // hipFree waits for work that may still use the queue state.
  return plan_free(static_cast<Plan*>(plan));
}

// ---- serving: one round per grid per launch ---------------------------------
extern "C" int npbx_round_depths(int64_t nsweeps, int32_t tblock, int32_t* depths,
                                 int32_t capacity, int32_t* count) {
  NPBX_REQUIRE(count && nsweeps >= 0 && tblock >= 2, "round_depths: bad arguments");
  const std::vector<int> d = round_depths(nsweeps, tblock);
  *count = static_cast<int32_t>(d.size());
  NPBX_REQUIRE(static_cast<int64_t>(d.size()) <= capacity && (depths || d.empty()),
               "round_depths: %zu rounds do not fit %d", d.size(), capacity);
  for (size_t i = 0; i < d.size(); ++i) depths[i] = d[i];
  return NPBX_OK;
}

extern "C" int npbx_cols_elems(int64_t n, int64_t* elems) {
  NPBX_REQUIRE(elems && n >= 0, "cols_elems: bad arguments");
  *elems = cols_elems(n);
  return NPBX_OK;
}

extern "C" int npbx_pack_cols(int dtype, const void* A, const void* B, int64_t n, void* cols,
                              void* stream) {
  NPBX_REQUIRE(dtype == NPBX_F32, "pack_cols: packed columns are FP32 only (got %s)",
               dtype_name(dtype));
  NPBX_REQUIRE(n >= 3 && n <= 2048, "pack_cols: n=%lld out of range", static_cast<long long>(n));
  NPBX_REQUIRE(A && B && cols, "pack_cols: null array");
  return pack_cols(A, B, static_cast<int>(n), cols, as_stream(stream));
}

extern "C" int npbx_queue_create(void** queue) {
  NPBX_REQUIRE(queue, "queue_create: null output pointer");
  QueueState* q = nullptr;
  NPBX_TRY(new_queue(&q));
  *queue = q;
  return NPBX_OK;
}

extern "C" int npbx_queue_destroy(void* queue) {
  if (queue) NPBX_HIP(hipFree(queue));
  return NPBX_OK;
}

extern "C" int npbx_rounds_launch(const npbx_round_spec* specs, int32_t count, int32_t tblock,
                                  int32_t seg, void* queue, void* stream) {
  NPBX_REQUIRE(specs && count > 0 && queue, "rounds_launch: no rounds or no queue");
  const int dtype = specs[0].dtype;
  NPBX_REQUIRE(dtype_size(dtype) != 0, "rounds_launch: bad dtype %d", dtype);
  const int tmax = max_depth(dtype);
  NPBX_REQUIRE(tblock >= 2 && tblock <= tmax, "rounds_launch: tblock=%d outside [2, %d] for %s",
               tblock, tmax, dtype_name(dtype));
  if (seg <= 0) seg = g_heat_seg.load();
  Packer pk;
  for (int i = 0; i < count; ++i) {
    const npbx_round_spec& s = specs[i];
    GridRound g;
    NPBX_TRY(check_grid("rounds_launch", i, s.kind, s.dtype, s.n, s.A, s.B, dtype, g));
    NPBX_REQUIRE((s.X == s.A || s.X == s.B) && (s.Y == s.A || s.Y == s.B) && s.X != s.Y,
                 "rounds_launch: grid %d must read one of A, B and write the other", i);
    NPBX_REQUIRE(s.depth >= 1 && s.depth <= tblock && s.level0 >= 0,
                 "rounds_launch: grid %d depth %d outside [1, %d]", i, s.depth, tblock);
    tile_grid(g, g.kind, g.n, dtype, tblock, seg);
    if (uses_cols(g.kind, dtype, g.rows)) g.cols = s.cols;
    g.X = s.X;
    g.Y = s.Y;
    g.parity0 = static_cast<int32_t>(s.level0 & 1);
    g.depth = s.depth;
    pk.add(g);
  }
  const int rows = g_heat_rows.load();
  const int nblocks = resident_blocks(dtype, tblock, rows);
  for (const RoundBatch& b : pk.out) {
    const int nt = b.cum[b.count];
    NPBX_TRY(launch_any(dtype, tblock, rows, b, static_cast<QueueState*>(queue),
                        std::min(nt, nblocks), as_stream(stream)));
  }
  return NPBX_OK;
}

// ---- tuning ------------------------------------------------------------------
extern "C" int npbx_set_tuning(const char* key, int64_t value) {
  NPBX_REQUIRE(key, "set_tuning: null key");
  if (std::strcmp(key, "heat.rows") == 0) {
    NPBX_REQUIRE(value == 2 || value == 4 || value == 6 || value == 8 || value == 10 ||
                     value == 12 || value == 14,
                 "set_tuning: heat.rows must be 2, 4, 6, 8 (shared tile) or 10, 12, 14 "
                 "(wave-private FP32 tile)");
    g_heat_rows.store(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "heat.seg") == 0) {
    NPBX_REQUIRE(value >= 1 && value <= 4096, "set_tuning: heat.seg must be in [1, 4096]");
    g_heat_seg.store(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "sweep.shape") == 0) {
    NPBX_REQUIRE(value >= 0 && value <= 3, "set_tuning: sweep.shape must be in [0, 3]");
    g_sweep_shape.store(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "hdiff.tblock") == 0) {
    NPBX_REQUIRE(value == 1 || value == 2 || value == 4,
                 "set_tuning: hdiff.tblock must be 1, 2 or 4");
    set_hdiff_tblock(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "adi.path") == 0) {
    NPBX_REQUIRE(value >= 0 && value <= 3, "set_tuning: adi.path must be 0, 1, 2 or 3");
    set_adi_path(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "jacobi.deep") == 0) {
    NPBX_REQUIRE(value >= 0 && value <= jacobi_deep_max(),
                 "set_tuning: jacobi.deep must be in [0, %d]", jacobi_deep_max());
    set_jacobi_deep_depth(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "hdiff.fuse_rows") == 0) {
    NPBX_REQUIRE(value == 0 || (value >= 4 && value <= 256),
                 "set_tuning: hdiff.fuse_rows must be 0 or in [4, 256]");
    set_hdiff_fuse_rows(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "hdiff.shape") == 0) {
    NPBX_REQUIRE(value >= 0 && value < hdiff_shape_count(),
                 "set_tuning: hdiff.shape must be in [0, %d]", hdiff_shape_count() - 1);
    set_hdiff_shape(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "jacobi.deep_waves") == 0) {
    NPBX_REQUIRE(value == 0 || value == 4 || value == 8 || value == 16,
                 "set_tuning: jacobi.deep_waves must be 0 (by shape), 4, 8 or 16");
    set_jacobi_deep_waves(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "vadv.path") == 0) {
    NPBX_REQUIRE(value >= 0 && value <= 4, "set_tuning: vadv.path must be in [0, 4]");
    set_vadv_path(static_cast<int>(value));
    return NPBX_OK;
  }
  if (std::strcmp(key, "seidel.blocks") == 0) {
    NPBX_REQUIRE(value >= 1 && value <= (1 << 20),
                 "set_tuning: seidel.blocks must be in [1, 1048576]");
    set_seidel_blocks(static_cast<int>(value));
    return NPBX_OK;
  }
  return fail(NPBX_EINVAL,
              "set_tuning: unknown key '%s' (heat.rows, heat.seg, sweep.shape, hdiff.tblock, "
              "hdiff.shape, hdiff.fuse_rows, adi.path, vadv.path, seidel.blocks, jacobi.deep, jacobi.deep_waves)",
              key);
}

extern "C" int npbx_get_tuning(const char* key, int64_t* value) {
  NPBX_REQUIRE(key && value, "get_tuning: null argument");
  if (std::strcmp(key, "heat.rows") == 0) {
    *value = g_heat_rows.load();
    return NPBX_OK;
  }
  if (std::strcmp(key, "heat.seg") == 0) {
    *value = g_heat_seg.load();
    return NPBX_OK;
  }
  if (std::strcmp(key, "sweep.shape") == 0) {
    *value = g_sweep_shape.load();
    return NPBX_OK;
  }
  if (std::strcmp(key, "adi.path") == 0) {
    *value = adi_path();
    return NPBX_OK;
  }
  if (std::strcmp(key, "hdiff.tblock") == 0) {
    *value = hdiff_tblock();
    return NPBX_OK;
  }
  if (std::strcmp(key, "jacobi.deep") == 0) {
    *value = jacobi_deep_depth();
    return NPBX_OK;
  }
  if (std::strcmp(key, "hdiff.fuse_rows") == 0) {
    *value = hdiff_fuse_rows();
    return NPBX_OK;
  }
  if (std::strcmp(key, "hdiff.shape") == 0) {
    *value = hdiff_shape();
    return NPBX_OK;
  }
  if (std::strcmp(key, "jacobi.deep_waves") == 0) {
    *value = jacobi_deep_waves();
    return NPBX_OK;
  }
  if (std::strcmp(key, "vadv.path") == 0) {
    *value = vadv_path();
    return NPBX_OK;
  }
  if (std::strcmp(key, "seidel.blocks") == 0) {
    *value = seidel_blocks();
    return NPBX_OK;
  }
  return fail(NPBX_EINVAL, "get_tuning: unknown key '%s'", key);
}
