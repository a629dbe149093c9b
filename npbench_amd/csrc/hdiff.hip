// Fourth-order horizontal diffusion with flux limiter (hdiff_numpy.py:5-29),
// fused into one pass: lap -> flx/fly with limiter -> out, no temporaries in
// HBM.
//
// Layout: in (I+4, J+4, K), out and coeff (I, J, K), C order, k contiguous.
// The stencil acts in (i, j) only; every k is independent.  Output point
// (x, y, k) sits over in-field point (x+2, y+2, k).
//
// Marching kernel (hdiff_march): a lane owns W adjacent output columns and
// one vector of V consecutive k (4, 8 or 16 bytes by shape)
// (lanes enumerate (column group, k-vector) pairs, so any K maps densely
// onto waves and loads stay contiguous), and walks down a strip of output
// rows x, keeping the in-field rows it still needs in registers and
// carrying the Laplacians and the i-flux of the face below from one row to
// the next; the next row's loads are issued before the current row is
// computed.  The 2-cell halo of the input makes every access in bounds.
// Same operations in the same order as the reference, so FP64 is bitwise
// NumPy.
//
// Fallback (hdiff_kernel, K not a multiple of V or unaligned arrays): one
// output point per thread, the Laplacians rebuilt from 13 cached loads.
// The limiter keeps the reference test "r * grad > 0 -> 0".
//
// Iterated form (nsteps > 1, build-defined): step n reads the previous
// step's output as the interior of the input field, with the 2-cell halo of
// the original in_field held fixed.  Steps ping-pong through a library
// workspace holding one (I+4, J+4, K) field; in_field is never written.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "common.h"
#include "vecio.h"

namespace npbx {
namespace {

constexpr int kLanes = 64;
constexpr int kCols = 4;

template <class C>
__device__ __forceinline__ C limit(C r, C grad) {
  return (r * grad > static_cast<C>(0)) ? static_cast<C>(0) : r;
}

// One output point.  `f` points at in-field (x+2, y+2, k); si / sj are the
// element strides of the in field along i and j.
template <class S>
__device__ __forceinline__ typename Storage<S>::C hdiff_point(const S* __restrict__ f, int64_t si,
                                                              int64_t sj,
                                                              typename Storage<S>::C cf) {
  using T = Storage<S>;
  using C = typename T::C;
  auto at = [&](int di, int dj) { return T::load(f[di * si + dj * sj]); };
  const C c = at(0, 0);
  const C ip = at(1, 0), im = at(-1, 0), jp = at(0, 1), jm = at(0, -1);
  // Laplacian centred at in-field offset (di, dj):
  // 4*c - ((((i+1) + (i-1)) + (j+1)) + (j-1))
  auto lap = [&](C centre, int di, int dj) {
    C nb = at(di + 1, dj) + at(di - 1, dj);
    nb = nb + at(di, dj + 1);
    nb = nb + at(di, dj - 1);
    return static_cast<C>(4.0) * centre - nb;
  };
  const C l0 = lap(c, 0, 0);
  const C l_ip = lap(ip, 1, 0), l_im = lap(im, -1, 0);
  const C l_jp = lap(jp, 0, 1), l_jm = lap(jm, 0, -1);
  // flx across faces (x+2 | x+3) and (x+1 | x+2); fly across (y+2 | y+3)
  // and (y+1 | y+2), hdiff_numpy.py:11-25.
  const C flx_hi = limit<C>(l_ip - l0, ip - c);
  const C flx_lo = limit<C>(l0 - l_im, c - im);
  const C fly_hi = limit<C>(l_jp - l0, jp - c);
  const C fly_lo = limit<C>(l0 - l_jm, c - jm);
  C d = flx_hi - flx_lo;
  d = d + fly_hi;
  d = d - fly_lo;
  return c - cf * d;
}

template <class S>
__global__ __launch_bounds__(kLanes* kCols) void hdiff_kernel(const S* __restrict__ in,
                                                             S* __restrict__ out,
                                                             const S* __restrict__ coeff, int I,
                                                             int J, int K, int64_t in_si,
                                                             int64_t in_sj, S* __restrict__ halo_dst) {
  using T = Storage<S>;
  const int k = blockIdx.x * kLanes + threadIdx.x;
  const int y = blockIdx.y * kCols + threadIdx.y;
  const int x = blockIdx.z;
  if (k >= K || y >= J) return;
  const S* f = in + (x + 2) * in_si + (y + 2) * in_sj + k;
  const int64_t o = (static_cast<int64_t>(x) * J + y) * K + k;
  const auto v = T::store(hdiff_point<S>(f, in_si, in_sj, T::load(coeff[o])));
  if (halo_dst) {
    // Intermediate step: write into the interior of the next input field.
    halo_dst[(x + 2) * in_si + (y + 2) * in_sj + k] = v;
  } else {
    out[o] = v;
  }
}

// Copies the 2-cell halo ring (everything outside the interior) of src into
// dst; both are (I+4, J+4, K).  Only the ring's elements are visited: the
// columns (p, q) of rows p < 2 and p >= I+2 (whole rows) and of columns
// q < 2 and q >= J+2 of the rows between, each column's K elements
// contiguous.
template <class S>
__global__ __launch_bounds__(256) void hdiff_copy_halo(const S* __restrict__ src,
                                                       S* __restrict__ dst, int I, int J, int K) {
  const int64_t row_cols = 4 * static_cast<int64_t>(J + 4);  // rows 0, 1, I+2, I+3
  const int64_t side_cols = 4 * static_cast<int64_t>(I);     // 4 columns of rows 2 .. I+1
  const int64_t total = (row_cols + side_cols) * K;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t c = e / K, k = e % K;
    int64_t p, q;
    if (c < row_cols) {
      const int64_t r = c / (J + 4);
      p = r < 2 ? r : I + r;  // 0, 1, I+2, I+3
      q = c % (J + 4);
    } else {
      const int64_t s = c - row_cols;
      p = 2 + s / 4;
      const int64_t w = s % 4;
      q = w < 2 ? w : J + w;  // 0, 1, J+2, J+3
    }
    const int64_t at = (p * (J + 4) + q) * K + k;
    dst[at] = src[at];
  }
}

// hdiff_march<S, W, STRIP, VB>: a lane owns W adjacent output columns y0
// .. y0+W-1 at one k-vector of VB bytes (V = VB / sizeof(S) elements) and
// marches down STRIP output rows.  Per row it
// loads the W own columns' row p+2, the side columns' (y0-1, y0+W) row p+1
// and the far columns' (y0-2, y0+W+1) row p -- W + 4 vectors for W points --
// and computes W Laplacians one row down (carried to the next row), the two
// side columns' Laplacians, W i-fluxes (the face below is carried) and W + 1
// j-fluxes: the j-flux between two own columns is one value, the Laplacian
// of an own column is the one its neighbour needs.  Every Laplacian and
// flux has the reference's operands in the reference's order, whoever
// computes it, so any W gives bitwise the same output.  The last column
// group is shifted left to end at column J-1 (J >= W) and stores only its
// own columns, so every column is stored once and every load is in the
// field.  Prefetched vectors are held as raw storage words until the row
// that uses them.
template <class S, int W, int STRIP, int VB>
__global__ __launch_bounds__(256) void hdiff_march(const S* __restrict__ in, S* __restrict__ out,
                                                   const S* __restrict__ coeff, int I, int J,
                                                   int K, int xblocks, S* __restrict__ halo_dst) {
  using T = Storage<S>;
  using C = typename T::C;
  constexpr int V = VB / static_cast<int>(sizeof(S));
  static_assert(V >= 1, "a vector holds at least one element");
  using Raw = RawVec<S, V>;
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int xb = b % xblocks, strip = b / xblocks;
  const int kvs = K / V;
  const int groups = (J + W - 1) / W;
  const int t = xb * 256 + threadIdx.x;
  if (t >= groups * kvs) return;
  const int yn = (t / kvs) * W;     // the group's first column
  const int y0 = min(yn, J - W);    // first column computed
  const int k0 = (t % kvs) * V;
  const int x0 = strip * STRIP, x1 = min(x0 + STRIP, I);
  const int64_t si = static_cast<int64_t>(J + 4) * K;  // in-field row stride
  // in-field column y0 + c, c = 0 .. W+3: far left, side left, the W own
  // columns, side right, far right
  const S* __restrict__ base = in + static_cast<int64_t>(y0) * K + k0;
  auto ld = [&](int p, int c, Raw& r) { loadraw<S, V>(base + p * si + c * K, r); };
  auto lap = [](C c, C ip, C im, C jp, C jm) {
    C nb = ip + im;
    nb = nb + jp;
    nb = nb + jm;
    return static_cast<C>(4.0) * c - nb;
  };

  // Row slots: the own columns' rows p, p+1, p+2 and the side columns' (0
  // left, 1 right) rows p-1, p, p+1 sit in three slots each, which the
  // unrolled march rotates instead of moving registers; far[f] is far
  // column f at row p.
  using Own = C[W][V];
  using Side = C[2][V];
  Own oa, ob, oc;
  Side sa, sb, sc;
  C far[2][V];
  C L0[W][V], fxl[W][V];  // carried: lap(p) and the i-flux across (p-1 | p)
  const int p0 = x0 + 2;
  {
    Raw r[W][4], rs[2][2];
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
      for (int i = 0; i < 4; ++i) ld(p0 - 2 + i, w + 2, r[w][i]);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ld(p0 - 1 + i, 1, rs[0][i]);
      ld(p0 - 1 + i, W + 2, rs[1][i]);
    }
    C om2[W][V], om1[W][V];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      unpackv<S, V>(r[w][0], om2[w]);
      unpackv<S, V>(r[w][1], om1[w]);
      unpackv<S, V>(r[w][2], oa[w]);
      unpackv<S, V>(r[w][3], ob[w]);
    }
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      unpackv<S, V>(rs[s2][0], sa[s2]);
      unpackv<S, V>(rs[s2][1], sb[s2]);
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const C c = oa[w][v];
        const C rp = w + 1 < W ? oa[w + 1][v] : sb[1][v];   // (p, q+1)
        const C lp = w > 0 ? oa[w - 1][v] : sb[0][v];       // (p, q-1)
        const C rm = w + 1 < W ? om1[w + 1][v] : sa[1][v];  // (p-1, q+1)
        const C lm = w > 0 ? om1[w - 1][v] : sa[0][v];      // (p-1, q-1)
        L0[w][v] = lap(c, ob[w][v], om1[w][v], rp, lp);
        const C Lim = lap(om1[w][v], c, om2[w][v], rm, lm);
        fxl[w][v] = limit<C>(L0[w][v] - Lim, c - om1[w][v]);
      }
    }
  }
  // the loads that complete row p0's window (and start each next row)
  Raw nown[W], nsd[2], nfar[2], ncf[W];
#pragma unroll
  for (int w = 0; w < W; ++w) ld(p0 + 2, w + 2, nown[w]);
  ld(p0 + 1, 1, nsd[0]);
  ld(p0 + 1, W + 2, nsd[1]);
  ld(p0, 0, nfar[0]);
  ld(p0, W + 3, nfar[1]);
  const S* __restrict__ cfp = coeff + (static_cast<int64_t>(x0) * J + y0) * K + k0;
#pragma unroll
  for (int w = 0; w < W; ++w) loadraw<S, V>(cfp + w * K, ncf[w]);

  // One output row x (in-field row p = x + 2): o0/o1/o2 hold the own rows
  // p, p+1, p+2 (o2 filled here), s0/s1/s2 the side rows p-1, p, p+1 (s2
  // filled here).
  auto row = [&](int x, Own& o0, Own& o1, Own& o2, Side& s0, Side& s1, Side& s2) {
    const int p = x + 2;
    C cf[W][V];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      unpackv<S, V>(nown[w], o2[w]);
      unpackv<S, V>(ncf[w], cf[w]);
    }
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      unpackv<S, V>(nsd[f], s2[f]);
      unpackv<S, V>(nfar[f], far[f]);
    }
    {  // next row's loads, in flight while this row is computed; after the
       // field's last row they reload rows of the field (rows clamped), so
       // they need no branch
      const int pn = min(p + 1, I + 1);
#pragma unroll
      for (int w = 0; w < W; ++w) ld(pn + 2, w + 2, nown[w]);
      ld(pn + 1, 1, nsd[0]);
      ld(pn + 1, W + 2, nsd[1]);
      ld(pn, 0, nfar[0]);
      ld(pn, W + 3, nfar[1]);
      const S* nc = cfp + static_cast<int64_t>(pn - 2 - x0) * J * K;
#pragma unroll
      for (int w = 0; w < W; ++w) loadraw<S, V>(nc + w * K, ncf[w]);
    }
    C res[W][V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      // Laplacians at row p of columns -1 .. W (own: carried), row p's values
      C lc[W + 2], cv[W + 2];
      lc[0] = lap(s1[0][v], s2[0][v], s0[0][v], o0[0][v], far[0][v]);
      lc[W + 1] = lap(s1[1][v], s2[1][v], s0[1][v], far[1][v], o0[W - 1][v]);
      cv[0] = s1[0][v];
      cv[W + 1] = s1[1][v];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        lc[w + 1] = L0[w][v];
        cv[w + 1] = o0[w][v];
      }
      // j-fluxes across the faces (c-1 | c), c = 0 .. W (own column w sits
      // at c = w + 1)
      C fy[W + 1];
#pragma unroll
      for (int c = 0; c <= W; ++c) fy[c] = limit<C>(lc[c + 1] - lc[c], cv[c + 1] - cv[c]);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const C c = o0[w][v];
        const C rp1 = w + 1 < W ? o1[w + 1][v] : s2[1][v];  // (p+1, q+1)
        const C lp1 = w > 0 ? o1[w - 1][v] : s2[0][v];      // (p+1, q-1)
        const C Lip = lap(o1[w][v], o2[w][v], c, rp1, lp1);
        const C flx_hi = limit<C>(Lip - L0[w][v], o1[w][v] - c);
        C dv = flx_hi - fxl[w][v];
        dv = dv + fy[w + 1];
        dv = dv - fy[w];
        res[w][v] = c - cf[w][v] * dv;
        L0[w][v] = Lip;
        fxl[w][v] = flx_hi;
      }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (y0 + w < yn) continue;  // stored by the group to the left
      if (halo_dst)
        storev<S, V>(halo_dst + p * si + static_cast<int64_t>(y0 + w + 2) * K + k0, res[w]);
      else
        storev<S, V>(out + (static_cast<int64_t>(x) * J + y0 + w) * K + k0, res[w]);
    }
  };
  // Fully unrolled: the slots rotate and the carried values change
  // registers at compile time, with no loop-back copies.
#pragma unroll
  for (int r = 0; r < STRIP; ++r) {
    if (x0 + r >= x1) break;
    if (r % 3 == 0)
      row(x0 + r, oa, ob, oc, sa, sb, sc);
    else if (r % 3 == 1)
      row(x0 + r, ob, oc, oa, sb, sc, sa);
    else
      row(x0 + r, oc, oa, ob, sc, sa, sb);
  }
}

// Marching shapes (tuning knob "hdiff.shape"): {W columns per lane, STRIP
// rows per thread, VB bytes of k per lane}; a shape whose vector is
// narrower than one element (4 bytes of FP64) runs the storage type's
// default.
struct MarchShape {
  int w, strip, vb;
};
constexpr MarchShape kMarchShapes[] = {{0, 0, 0},   {1, 12, 16}, {1, 6, 16}, {2, 6, 16},
                                       {2, 12, 16}, {1, 18, 16}, {4, 6, 8},  {2, 6, 8},
                                       {4, 12, 8},  {4, 6, 4},   {4, 12, 4}, {2, 12, 8},
                                       {2, 12, 4},  {3, 6, 8}};
constexpr int kMarchShapeCount = sizeof(kMarchShapes) / sizeof(kMarchShapes[0]);
std::atomic<int> g_hdiff_shape{0};

// Default shape per storage type (knob value 0), the fastest at 256 x 256 x
// 160 on the MI355X: FP64 and FP32 2 columns x 6 rows with 8-byte vectors,
// BF16 4 x 6 with 4-byte vectors (two elements: the four columns' window
// then fits in 110 VGPRs, four waves per SIMD).
template <class S>
constexpr int default_march_shape() { return sizeof(S) == 2 ? 9 : 7; }

// ---------------------------------------------------------------- fused steps
// hdiff_tb<S, TB>: up to TB iterated steps in one HBM round trip.  A
// workgroup owns a 32 x 32 block of output points (i, j) for one k-vector
// (V = 16 bytes of k) and stages the input field over the block plus a halo
// of 2 cells per fused step in LDS; each step recomputes the region shrunk
// by 2 on every side into the other LDS buffer (field-halo points keep their
// fixed values; intermediate values are rounded through the storage type as
// the iterated form's stores round them), and the last step writes only the
// block.  Per point the arithmetic is hdiff_point's, so results are bitwise
// the iterated form's.  Consecutive workgroups are the k-vectors of one
// block (XCD-contiguous), so the 128-byte lines a workgroup reads 16 bytes
// of are read by its neighbours from the same L2.  It cuts HBM bytes per
// step (32x32 blocks, two steps: about 0.6x) but each wave load touches one
// line per lane, and measured slower than one launch per step.
constexpr int kTbB = 32;  // output block side (i and j)

template <class S, int TB>
struct TbShape {
  static constexpr int R = kTbB + 4 * TB;  // side of the staged region
  static constexpr int V = Vec16<S>::N;
  static constexpr int ELEMS = R * R * V;  // one LDS field buffer, arithmetic type
};

template <class S, int TB>
__global__ __launch_bounds__(256) void hdiff_tb(const S* __restrict__ src,
                                                const S* __restrict__ coeff, S* __restrict__ out,
                                                S* __restrict__ next, int I, int J, int K, int steps,
                                                int tiles_y, int nblocks) {
  using St = Storage<S>;
  using C = typename St::C;
  using Sh = TbShape<S, TB>;
  constexpr int V = Sh::V, R = Sh::R;
  __shared__ __align__(16) C fa[Sh::ELEMS];
  __shared__ __align__(16) C fb[Sh::ELEMS];
  __shared__ __align__(16) C cfs[Sh::ELEMS];
  const int kvs = K / V;
  const int b = xcd_tile(blockIdx.x, nblocks);
  const int kc = b % kvs, tile = b / kvs;
  const int X0 = (tile / tiles_y) * kTbB, Y0 = (tile % tiles_y) * kTbB;
  const int pa = X0 + 2 - 2 * TB, qa = Y0 + 2 - 2 * TB;  // field point of buffer (0, 0)
  const int64_t si = static_cast<int64_t>(J + 4) * K, sj = K;
  const int k0 = kc * V;
  auto in_field = [&](int p, int q) { return p >= 0 && p < I + 4 && q >= 0 && q < J + 4; };
  auto interior = [&](int p, int q) { return p >= 2 && p < I + 2 && q >= 2 && q < J + 2; };

  for (int e = threadIdx.x; e < R * R; e += 256) {
    const int i = e / R, j = e % R, p = pa + i, q = qa + j;
    C v[V];
    if (in_field(p, q)) {
      load16(src + p * si + q * sj + k0, v);
    } else {
#pragma unroll
      for (int u = 0; u < V; ++u) v[u] = C(0);
    }
#pragma unroll
    for (int u = 0; u < V; ++u) fa[e * V + u] = v[u];
    if (interior(p, q) && i >= 2 && i < R - 2 && j >= 2 && j < R - 2) {
      load16(coeff + ((static_cast<int64_t>(p) - 2) * J + (q - 2)) * K + k0, v);
#pragma unroll
      for (int u = 0; u < V; ++u) cfs[e * V + u] = v[u];
    }
  }
  __syncthreads();

  // One step from buffer A into buffer B over region [lo, R-lo)^2 (t < steps)
  // or into global memory over the output block (t == steps).
  auto step = [&](const C* __restrict__ A, C* __restrict__ B, int t) {
    const bool last = t == steps;
    const int lo = last ? 2 * TB : 2 * t, n = last ? kTbB : R - 4 * t;
    for (int e = threadIdx.x; e < n * n; e += 256) {
      const int i = lo + e / n, j = lo + e % n, p = pa + i, q = qa + j;
      if (!in_field(p, q)) continue;
      const int at = i * R + j;
      C res[V];
      if (!interior(p, q)) {  // the field's halo: fixed
        if (last) continue;
#pragma unroll
        for (int u = 0; u < V; ++u) res[u] = A[at * V + u];
      } else {
#pragma unroll
        for (int u = 0; u < V; ++u) {
          auto F = [&](int idx) { return A[idx * V + u]; };
          auto lap = [&](int ci) {
            C nb = F(ci + R) + F(ci - R);
            nb = nb + F(ci + 1);
            nb = nb + F(ci - 1);
            return static_cast<C>(4.0) * F(ci) - nb;
          };
          const C c = F(at), ip = F(at + R), im = F(at - R), jp = F(at + 1), jm = F(at - 1);
          const C l0 = lap(at);
          const C flx_hi = limit<C>(lap(at + R) - l0, ip - c);
          const C flx_lo = limit<C>(l0 - lap(at - R), c - im);
          const C fly_hi = limit<C>(lap(at + 1) - l0, jp - c);
          const C fly_lo = limit<C>(l0 - lap(at - 1), c - jm);
          C d = flx_hi - flx_lo;
          d = d + fly_hi;
          d = d - fly_lo;
          res[u] = St::load(St::store(c - cfs[at * V + u] * d));
        }
      }
      if (!last) {
#pragma unroll
        for (int u = 0; u < V; ++u) B[at * V + u] = res[u];
      } else if (p - 2 < I && q - 2 < J) {
        if (next)
          store16(next + p * si + q * sj + k0, res);
        else
          store16(out + ((static_cast<int64_t>(p) - 2) * J + (q - 2)) * K + k0, res);
      }
    }
  };
  for (int t = 1; t <= steps; ++t) {
    if (t & 1)
      step(fa, fb, t);
    else
      step(fb, fa, t);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- two steps
// hdiff_fuse2w<S, VB, W, BAND, KV>: two iterated steps in one pass, without
// staging whole blocks.  A workgroup owns BAND in-field columns y0 ..
// y0+BAND-1 (y0 = band (BAND-4)) at KV consecutive k-vectors of VB bytes;
// a lane owns W adjacent columns at one k-vector (lane = (column group,
// k-vector), k fastest, so a wave load covers KV x VB contiguous bytes of
// each of its columns) and marches down a strip of rows, running
// hdiff_march's W-column row twice over: the first step at in-field row p1
// from the input, the second at row p1 - 2 from first-step values -- its
// own columns' from its registers, the side and far columns either side of
// the group from neighbouring lanes through a four-row LDS ring (written
// before and read after one barrier per row).  A W-column row shares the
// Laplacian and the j-flux between the lane's own columns.  First-step
// values outside the field's interior are the input's (the halo the
// iterated form holds fixed), interior ones are rounded through the storage
// type as the iterated form's stores round them, and every Laplacian and
// flux has the reference's operands in the reference's order, so the output
// is bitwise two launches of hdiff_march.  The band's outer two columns on
// each side give first-step values only (BAND - 4 columns of output), a
// strip of S output rows computes S + 4 first-step rows, and the three-row
// windows rotate through fixed register slots (rows in threes) instead of
// being copied.
template <class S, int VB, int W, int BAND, int KV>
__global__ __launch_bounds__(BAND / W * KV) void hdiff_fuse2w(const S* __restrict__ src,
                                                              const S* __restrict__ coeff,
                                                              S* __restrict__ out,
                                                              S* __restrict__ next, int I, int J,
                                                              int K, int bands, int kgroups,
                                                              int strip) {
  using T = Storage<S>;
  using C = typename T::C;
  constexpr int V = VB / static_cast<int>(sizeof(S));
  static_assert(BAND % W == 0, "whole column groups per band");
  using Raw = RawVec<S, V>;
  __shared__ C ring[4][BAND * KV * V];  // first-step rows: [row % 4][column][k-vector][v]
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int band = b % bands, rest = b / bands;
  const int kg = rest % kgroups, sidx = rest / kgroups;
  const int cg = threadIdx.x / KV, kv = threadIdx.x % KV;
  const int c0 = cg * W;                  // the group's first column in the band
  const int y0 = band * (BAND - 4);
  const int qa = y0 + c0;                 // in-field column of own column 0
  const int kraw = (kg * KV + kv) * V;
  const bool kok = kraw + V <= K;
  const int k0 = kok ? kraw : K - V;  // lanes past K compute a copy and store nothing
  const int x0 = sidx * strip, x1 = min(x0 + strip, I);  // second-step output rows
  const int64_t si = static_cast<int64_t>(J + 4) * K;
  auto colo = [&](int qq) { return static_cast<int64_t>(min(max(qq, 0), J + 3)) * K + k0; };
  // in-field columns qa-2 .. qa+W+1 (far left, side left, own, side right,
  // far right), clamped (clamped columns feed only values that are replaced
  // or not stored)
  int64_t cq[W + 4];
#pragma unroll
  for (int c = 0; c < W + 4; ++c) cq[c] = colo(qa - 2 + c);
  auto ld = [&](int p, int c, Raw& r) {
    loadraw<S, V>(src + static_cast<int64_t>(min(max(p, 0), I + 3)) * si + cq[c], r);
  };
  auto ldcf = [&](int x, int w, Raw& r) {
    loadraw<S, V>(coeff + (static_cast<int64_t>(min(max(x, 0), I - 1)) * J +
                           min(max(qa + w - 2, 0), J - 1)) * K + k0, r);
  };
  auto lap = [](C c, C ip, C im, C jp, C jm) {
    C nb = ip + im;
    nb = nb + jp;
    nb = nb + jm;
    return static_cast<C>(4.0) * c - nb;
  };
  bool col_in[W];
#pragma unroll
  for (int w = 0; w < W; ++w) col_in[w] = qa + w >= 2 && qa + w < J + 2;
  using Own = C[3][W][V];   // own columns' rows p, p+1, p+2 at slots rot, rot+1, rot+2 (mod 3)
  using Side = C[3][2][V];  // side columns' (0 left, 1 right) rows p-1, p, p+1, same slots
  // one W-column row at rotation rot (hdiff_march's row): far[f] = row p of
  // the far columns; L0, fx carried
  auto mrow = [&](auto rot, const Own& own, const Side& sd, const C (&far)[2][V],
                  const C (&cf)[W][V], C (&L0)[W][V], C (&fx)[W][V], C (&res)[W][V]) {
    constexpr int a0 = decltype(rot)::value % 3, a1 = (a0 + 1) % 3, a2 = (a0 + 2) % 3;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      C lc[W + 2], cv[W + 2];
      lc[0] = lap(sd[a1][0][v], sd[a2][0][v], sd[a0][0][v], own[a0][0][v], far[0][v]);
      lc[W + 1] = lap(sd[a1][1][v], sd[a2][1][v], sd[a0][1][v], far[1][v], own[a0][W - 1][v]);
      cv[0] = sd[a1][0][v];
      cv[W + 1] = sd[a1][1][v];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        lc[w + 1] = L0[w][v];
        cv[w + 1] = own[a0][w][v];
      }
      C fy[W + 1];
#pragma unroll
      for (int c = 0; c <= W; ++c) fy[c] = limit<C>(lc[c + 1] - lc[c], cv[c + 1] - cv[c]);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const C c = own[a0][w][v];
        const C rp1 = w + 1 < W ? own[a1][w + 1][v] : sd[a2][1][v];  // (p+1, q+1)
        const C lp1 = w > 0 ? own[a1][w - 1][v] : sd[a2][0][v];      // (p+1, q-1)
        const C Lip = lap(own[a1][w][v], own[a2][w][v], c, rp1, lp1);
        const C flx_hi = limit<C>(Lip - L0[w][v], own[a1][w][v] - c);
        C dv = flx_hi - fx[w][v];
        dv = dv + fy[w + 1];
        dv = dv - fy[w];
        res[w][v] = c - cf[w][v] * dv;
        L0[w][v] = Lip;
        fx[w][v] = flx_hi;
      }
    }
  };
  // march start at row P (hdiff_march's): own rows P-2 .. P+1 in om2, om1,
  // own[0], own[1] (slots 0, 1); side rows P-1, P in sd[0], sd[1]
  auto minit = [&](const C (&om2)[W][V], const C (&om1)[W][V], const Own& own, const Side& sd,
                   C (&L0)[W][V], C (&fx)[W][V]) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const C c = own[0][w][v];
        const C rp = w + 1 < W ? own[0][w + 1][v] : sd[1][1][v];  // (P, q+1)
        const C lp = w > 0 ? own[0][w - 1][v] : sd[1][0][v];      // (P, q-1)
        const C rm = w + 1 < W ? om1[w + 1][v] : sd[0][1][v];     // (P-1, q+1)
        const C lm = w > 0 ? om1[w - 1][v] : sd[0][0][v];         // (P-1, q-1)
        L0[w][v] = lap(c, own[1][w][v], om1[w][v], rp, lp);
        const C Lim = lap(om1[w][v], c, om2[w][v], rm, lm);
        fx[w][v] = limit<C>(L0[w][v] - Lim, c - om1[w][v]);
      }
    }
  };
  auto ring_at = [&](int row, int c) {  // row's slot, band column c (clamped into the band)
    const int cc = min(max(c, 0), BAND - 1);
    return &ring[row & 3][(cc * KV + kv) * V];
  };
  auto ring_ld = [&](int row, int c, C (&v)[V]) {
    const C* p = ring_at(row, c);
#pragma unroll
    for (int e = 0; e < V; ++e) v[e] = p[e];
  };

  // ---- first step: state at row x0
  Own o1;
  Side s1;
  C far1[2][V], L01[W][V], fx1[W][V];
  {
    Raw r[W][4], rs[2][2];
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
      for (int i = 0; i < 4; ++i) ld(x0 - 2 + i, w + 2, r[w][i]);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ld(x0 - 1 + i, 1, rs[0][i]);
      ld(x0 - 1 + i, W + 2, rs[1][i]);
    }
    C om2[W][V], om1[W][V];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      unpackv<S, V>(r[w][0], om2[w]);
      unpackv<S, V>(r[w][1], om1[w]);
      unpackv<S, V>(r[w][2], o1[0][w]);
      unpackv<S, V>(r[w][3], o1[1][w]);
    }
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      unpackv<S, V>(rs[f][0], s1[0][f]);
      unpackv<S, V>(rs[f][1], s1[1][f]);
    }
    minit(om2, om1, o1, s1, L01, fx1);
  }
  Raw nown[W], nsd[2], nfar[2], ncf[W];
#pragma unroll
  for (int w = 0; w < W; ++w) ld(x0 + 2, w + 2, nown[w]);
  ld(x0 + 1, 1, nsd[0]);
  ld(x0 + 1, W + 2, nsd[1]);
  ld(x0, 0, nfar[0]);
  ld(x0, W + 3, nfar[1]);
#pragma unroll
  for (int w = 0; w < W; ++w) ldcf(x0 - 2, w, ncf[w]);
  // coefficients of the last three first-step rows: row p1's in slot
  // rot % 3 of its step
  C cfh[3][W][V];
  // one first-step row p1 at rotation rot: its values at the own columns,
  // into the ring
  auto step1 = [&](auto rot, int p1, C (&sv)[W][V]) {
    constexpr int a0 = decltype(rot)::value % 3, a2 = (a0 + 2) % 3;
    C (&cf)[W][V] = cfh[a0];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      unpackv<S, V>(nown[w], o1[a2][w]);
      unpackv<S, V>(ncf[w], cf[w]);
    }
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      unpackv<S, V>(nsd[f], s1[a2][f]);
      unpackv<S, V>(nfar[f], far1[f]);
    }
    // next row's loads (rows clamped into the field)
#pragma unroll
    for (int w = 0; w < W; ++w) ld(p1 + 3, w + 2, nown[w]);
    ld(p1 + 2, 1, nsd[0]);
    ld(p1 + 2, W + 2, nsd[1]);
    ld(p1 + 1, 0, nfar[0]);
    ld(p1 + 1, W + 3, nfar[1]);
#pragma unroll
    for (int w = 0; w < W; ++w) ldcf(p1 - 1, w, ncf[w]);
    C res[W][V];
    mrow(rot, o1, s1, far1, cf, L01, fx1, res);
    const bool rin = p1 >= 2 && p1 < I + 2;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const bool in = rin && col_in[w];
#pragma unroll
      for (int v = 0; v < V; ++v) sv[w][v] = in ? T::load(T::store(res[w][v])) : o1[a0][w][v];
      C* wp = ring_at(p1, c0 + w);
#pragma unroll
      for (int e = 0; e < V; ++e) wp[e] = sv[w][e];
    }
  };

  // ---- warm-up: first-step rows x0 .. x0+3, then the second step's state
  // at row P = x0+2
  using R0 = std::integral_constant<int, 0>;
  using R1 = std::integral_constant<int, 1>;
  using R2 = std::integral_constant<int, 2>;
  C h[4][W][V];
  step1(R0{}, x0, h[0]);
  __syncthreads();
  step1(R1{}, x0 + 1, h[1]);
  __syncthreads();
  step1(R2{}, x0 + 2, h[2]);
  __syncthreads();
  step1(R0{}, x0 + 3, h[3]);
  __syncthreads();
  Own o2;
  Side s2;
  C far2[2][V], L02[W][V], fx2[W][V];
  ring_ld(x0 + 1, c0 - 1, s2[0][0]);
  ring_ld(x0 + 2, c0 - 1, s2[1][0]);
  ring_ld(x0 + 1, c0 + W, s2[0][1]);
  ring_ld(x0 + 2, c0 + W, s2[1][1]);
#pragma unroll
  for (int w = 0; w < W; ++w)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      o2[0][w][v] = h[2][w][v];
      o2[1][w][v] = h[3][w][v];
    }
  minit(h[0], h[1], o2, s2, L02, fx2);
  bool store_col[W];
#pragma unroll
  for (int w = 0; w < W; ++w)
    store_col[w] = kok && c0 + w >= 2 && c0 + w < BAND - 2 && qa + w < J + 2;
  // ---- main: first-step row p1, second-step row p2 = p1 - 2, in threes
  // (the windows' period; the first step's rotation is one more than the
  // second's after four warm-up rows).  The guards are uniform across the
  // workgroup.
  const int nrows = x1 - x0;
  auto row = [&](auto rot2, int p1) {
    using Rot1 = std::integral_constant<int, (decltype(rot2)::value + 1) % 3>;
    constexpr int a2 = (decltype(rot2)::value + 2) % 3;
    C sv[W][V];
    step1(Rot1{}, p1, sv);
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
      for (int v = 0; v < V; ++v) o2[a2][w][v] = sv[w][v];
    ring_ld(p1 - 1, c0 - 1, s2[a2][0]);
    ring_ld(p1 - 1, c0 + W, s2[a2][1]);
    ring_ld(p1 - 2, c0 - 2, far2[0]);
    ring_ld(p1 - 2, c0 + W + 1, far2[1]);
    C res[W][V];
    // coefficients of output row p2 - 2 = p1 - 4: the first step's at row
    // p1 - 2, two first-step rotations back
    mrow(rot2, o2, s2, far2, cfh[(Rot1::value + 1) % 3], L02, fx2, res);
    const int p2 = p1 - 2;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (!store_col[w]) continue;
      if (next)
        storev<S, V>(next + static_cast<int64_t>(p2) * si + cq[w + 2], res[w]);
      else
        storev<S, V>(out + (static_cast<int64_t>(p2 - 2) * J + (qa + w - 2)) * K + k0, res[w]);
    }
  };
  for (int base = 0; base < nrows; base += 3) {
    row(R0{}, x0 + 4 + base);
    if (base + 1 < nrows) row(R1{}, x0 + 5 + base);
    if (base + 2 < nrows) row(R2{}, x0 + 6 + base);
  }
}

// Fused steps per round (tuning knob "hdiff.tblock"; 1 = one launch per step,
// the default: at 256x256x160 on the MI355X the fused kernel is slower, its
// lanes each reading 16 bytes of a different 640-byte k-column).
std::atomic<int> g_hdiff_tblock{1};

template <class S>
constexpr int hdiff_tb_max() { return sizeof(S) == 2 ? 2 : 4; }  // BF16 x4 exceeds LDS

template <class S>
int run_hdiff(const S* in, S* out, const S* coeff, int I, int J, int K, int64_t nsteps,
              hipStream_t stream) {
  int shape = g_hdiff_shape.load();
  if (shape <= 0 || shape >= kMarchShapeCount || kMarchShapes[shape].vb < int(sizeof(S)))
    shape = default_march_shape<S>();
  if (kMarchShapes[shape].w > J) shape = 1;  // a group needs J >= W columns
  const int W = kMarchShapes[shape].w, strip = kMarchShapes[shape].strip;
  const int V = kMarchShapes[shape].vb / static_cast<int>(sizeof(S));
  const bool aligned = aligned16(in) && aligned16(out) && aligned16(coeff);
  const bool vec = K % V == 0 && aligned;
  const int64_t lanes = cdiv(J, W) * (K / V);
  const int xblocks = static_cast<int>(cdiv(lanes, 256));
  const int64_t mblocks = int64_t(xblocks) * cdiv(I, strip);
  NPBX_REQUIRE(!vec || (lanes < (int64_t(1) << 31) && mblocks < (int64_t(1) << 31)),
               "hdiff: grid too large");
  auto launch = [&](const S* src, S* next) -> int {
    if (vec) {
      const dim3 grid(static_cast<unsigned>(mblocks));
      auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, stream, src, out, coeff, I, J, K, xblocks,
                           next);
      };
      constexpr int b4 = sizeof(S) <= 4 ? 4 : 8;  // FP64 has no 4-byte vector
      switch (shape) {
        case 1: go(hdiff_march<S, 1, 12, 16>); break;
        case 2: go(hdiff_march<S, 1, 6, 16>); break;
        case 3: go(hdiff_march<S, 2, 6, 16>); break;
        case 4: go(hdiff_march<S, 2, 12, 16>); break;
        case 5: go(hdiff_march<S, 1, 18, 16>); break;
        case 6: go(hdiff_march<S, 4, 6, 8>); break;
        case 7: go(hdiff_march<S, 2, 6, 8>); break;
        case 8: go(hdiff_march<S, 4, 12, 8>); break;
        case 9: go(hdiff_march<S, 4, 6, b4>); break;
        case 10: go(hdiff_march<S, 4, 12, b4>); break;
        case 11: go(hdiff_march<S, 2, 12, 8>); break;
        case 12: go(hdiff_march<S, 2, 12, b4>); break;
        default: go(hdiff_march<S, 3, 6, 8>); break;
      }
      NPBX_LAUNCHED("hdiff_march");
    } else {
      const int64_t si = static_cast<int64_t>(J + 4) * K, sj = K;
      const dim3 block(kLanes, kCols);
      const dim3 grid(static_cast<unsigned>(cdiv(K, kLanes)),
                      static_cast<unsigned>(cdiv(J, kCols)), static_cast<unsigned>(I));
      hipLaunchKernelGGL(hdiff_kernel<S>, grid, block, 0, stream, src, out, coeff, I, J, K, si,
                         sj, next);
      NPBX_LAUNCHED("hdiff_kernel");
    }
    return NPBX_OK;
  };
  if (nsteps == 1) return launch(in, nullptr);
  const int tb = std::min(hdiff_tblock(), hdiff_tb_max<S>());
  // Two work fields ping-pong; each carries the original halo.
  const size_t field = static_cast<size_t>(I + 4) * (J + 4) * K;
  void* ws = nullptr;
  NPBX_TRY(workspace(2 * field * sizeof(S), &ws));
  S* w[2] = {static_cast<S*>(ws), static_cast<S*>(ws) + field};
  const int64_t ring = (4 * static_cast<int64_t>(J + 4) + 4 * static_cast<int64_t>(I)) * K;
  const unsigned copy_blocks = static_cast<unsigned>(std::min<int64_t>(cdiv(ring, 256), 4096));
  for (int b = 0; b < 2; ++b) {
    hipLaunchKernelGGL(hdiff_copy_halo<S>, dim3(copy_blocks), dim3(256), 0, stream, in, w[b], I, J,
                       K);
    NPBX_LAUNCHED("hdiff_copy_halo");
  }
  // Two steps in one pass (hdiff_fuse2w), strips of hdiff.fuse_rows output
  // rows (0: 16 for BF16, 32 otherwise).  BF16: 64 columns (60 of output)
  // per workgroup, two per lane, 8 k-vectors of 8 bytes; FP32 and FP64: 32
  // columns (28 of output), one per lane (two columns of 16-byte vectors
  // run one wave per SIMD), 8 k-vectors of 16 bytes.  The shapes measured
  // fastest at 256 x 256 x 160.
  constexpr int kFuseBand = sizeof(S) == 2 ? 64 : 32;
  constexpr int kFuseW = sizeof(S) == 2 ? 2 : 1, kFuseKV = 8;
  constexpr int kFuseVB = sizeof(S) == 2 ? 8 : 16;
  constexpr int kFuseV = kFuseVB / static_cast<int>(sizeof(S));
  auto fuse2 = [&](const S* from, S* to) -> int {
    const int bands = static_cast<int>(cdiv(J, kFuseBand - 4));
    const int kgroups = static_cast<int>(cdiv(K / kFuseV, kFuseKV));
    const int strip = hdiff_fuse_rows() > 0 ? hdiff_fuse_rows() : (sizeof(S) == 2 ? 16 : 32);
    const int64_t nb = int64_t(bands) * kgroups * cdiv(I, strip);
    NPBX_REQUIRE(nb < (int64_t(1) << 31), "hdiff: grid too large");
    hipLaunchKernelGGL((hdiff_fuse2w<S, kFuseVB, kFuseW, kFuseBand, kFuseKV>),
                       dim3(static_cast<unsigned>(nb)), dim3(kFuseBand / kFuseW * kFuseKV), 0,
                       stream, from, coeff, out, to, I, J, K, bands, kgroups, strip);
    NPBX_LAUNCHED("hdiff_fuse2w");
    return NPBX_OK;
  };
  const S* src = in;
  if (tb >= 2 && aligned && K % Vec16<S>::N == 0) {  // whole 16-byte vectors per column
    constexpr int V = Vec16<S>::N;
    const int tiles_x = static_cast<int>(cdiv(I, kTbB)), tiles_y = static_cast<int>(cdiv(J, kTbB));
    const int64_t nb = int64_t(tiles_x) * tiles_y * (K / V);
    NPBX_REQUIRE(nb < (int64_t(1) << 31), "hdiff: grid too large");
    int64_t done = 0, round = 0;
    while (done < nsteps) {
      const int steps = static_cast<int>(std::min<int64_t>(tb, nsteps - done));
      const bool last = done + steps == nsteps;
      S* next = last ? nullptr : w[round % 2];
      auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, stream, src,
                           coeff, out, next, I, J, K, steps, tiles_y, static_cast<int>(nb));
      };
      if (tb >= 4) {
        if constexpr (hdiff_tb_max<S>() >= 4) go(hdiff_tb<S, 4>);
        NPBX_LAUNCHED("hdiff_tb");
      } else if (steps == 2) {
        NPBX_TRY(fuse2(src, next));
      } else {
        NPBX_TRY(launch(src, next));
      }
      src = next;
      done += steps;
      ++round;
    }
    return NPBX_OK;
  }
  for (int64_t n = 0; n < nsteps; ++n) {
    const bool last = n + 1 == nsteps;
    S* next = w[n % 2];
    NPBX_TRY(launch(src, last ? nullptr : next));
    src = next;
  }
  return NPBX_OK;
}

}  // namespace

int hdiff_tblock() { return g_hdiff_tblock.load(); }
void set_hdiff_tblock(int tb) { g_hdiff_tblock.store(tb); }
std::atomic<int> g_hdiff_fuse_rows{0};
int hdiff_fuse_rows() { return g_hdiff_fuse_rows.load(); }
void set_hdiff_fuse_rows(int rows) { g_hdiff_fuse_rows.store(rows); }
int hdiff_shape() { return g_hdiff_shape.load(); }
void set_hdiff_shape(int shape) { g_hdiff_shape.store(shape); }
int hdiff_shape_count() { return kMarchShapeCount; }

}  // namespace npbx

extern "C" int npbx_hdiff(int dtype, const void* in_field, void* out_field, const void* coeff,
                          int64_t I, int64_t J, int64_t K, int64_t nsteps, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(I >= 0 && J >= 0 && K >= 0, "hdiff: negative extent");
  NPBX_REQUIRE(I <= 65535 && J <= (1 << 24) && K <= (1 << 24), "hdiff: extent out of range");
  NPBX_REQUIRE(nsteps >= 1, "hdiff: nsteps must be >= 1");
  if (I == 0 || J == 0 || K == 0) return NPBX_OK;
  NPBX_REQUIRE(in_field && out_field && coeff, "hdiff: null array");
  NPBX_REQUIRE(in_field != out_field && coeff != out_field, "hdiff: out_field aliases an input");
  hipStream_t s = as_stream(stream);
  return dispatch_dtype(dtype, [&](auto tag) {
    using S = decltype(tag);
    return run_hdiff<S>(static_cast<const S*>(in_field), static_cast<S*>(out_field),
                        static_cast<const S*>(coeff), static_cast<int>(I), static_cast<int>(J),
                        static_cast<int>(K), nsteps, s);
  });
}
