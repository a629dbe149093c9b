// Fourth-order horizontal diffusion with flux limiter (hdiff_numpy.py:5-29),
// fused into one pass: lap -> flx/fly with limiter -> out, no temporaries in
// HBM.
//
// Layout: in (I+4, J+4, K), out and coeff (I, J, K), C order, k contiguous.
// The stencil acts in (i, j) only; every k is independent.  Output point
// (x, y, k) sits over in-field point (x+2, y+2, k).
//
// Marching kernel (hdiff_march): a lane owns one output column y and one
// vector of V consecutive k (16 bytes; 8 for BF16, see march_vec) (lanes
// enumerate (y, k-vector) pairs,
// so any K maps densely onto waves and loads stay contiguous), and walks
// down a strip of march_strip<S>() output rows x.  It keeps in registers the in-field
// rows p-2..p+2 of its column q, rows p-1..p+1 of columns q+-1 and row p of
// q+-2 (p = x+2, q = y+2), and carries lap(p-1), lap(p) and the i-flux of
// the face below from one row to the next, so each output point loads five
// new vectors and computes three Laplacians and three fluxes.  The next
// row's loads are issued before the current row is computed.  The 2-cell
// halo of the input makes every access in bounds.  Same operations in the
// same order as the reference, so FP64 is bitwise NumPy.
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

// Output rows per thread: 16, and 8 for BF16, whose loads carry half the
// bytes and which needs the doubled number of waves to hide their latency
// (FP32 measured slower with 8).
template <class S>
constexpr int march_strip() { return sizeof(S) == 2 ? 8 : 16; }

// k-elements per lane of the marching kernel: one 16-byte vector, except
// BF16, which takes 4 (8 bytes) so that it runs as many waves with the same
// register footprint as FP32 (8 would halve the waves and double the window).
template <class S>
constexpr int march_vec() { return sizeof(S) == 2 ? 4 : Vec16<S>::N; }

template <class S>
__global__ __launch_bounds__(256) void hdiff_march(const S* __restrict__ in, S* __restrict__ out,
                                                   const S* __restrict__ coeff, int I, int J,
                                                   int K, int xblocks, S* __restrict__ halo_dst) {
  using T = Storage<S>;
  using C = typename T::C;
  constexpr int V = march_vec<S>();
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int xb = b % xblocks, strip = b / xblocks;
  const int kvs = K / V;
  const int t = xb * 256 + threadIdx.x;
  if (t >= J * kvs) return;
  const int y = t / kvs, k0 = (t % kvs) * V;
  constexpr int kStrip = march_strip<S>();
  const int x0 = strip * kStrip, x1 = min(x0 + kStrip, I);
  const int64_t si = static_cast<int64_t>(J + 4) * K;  // in-field row stride
  const int64_t sj = K;
  const S* f = in + (y + 2) * sj + k0;                // column q = y+2, row 0
  auto ld = [&](int p, int dq, C (&v)[V]) { loadv<S, V>(f + p * si + dq * sj, v); };

  // window at row p = x0+2: c0[0..4] = rows p-2..p+2 (col q), cm/cp[0..2] =
  // rows p-1..p+1 (cols q-1 / q+1), cmm / cpp = row p (cols q-2 / q+2)
  C c0[5][V], cm[3][V], cp[3][V], cmm[V], cpp[V];
  const int p0 = x0 + 2;
#pragma unroll
  for (int r = 0; r < 4; ++r) ld(p0 - 2 + r, 0, c0[r]);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    ld(p0 - 1 + r, -1, cm[r]);
    ld(p0 - 1 + r, 1, cp[r]);
  }
  // the loads that complete the window of row p0 (and start each next row)
  C n0[V], nm[V], np[V], nmm[V], npp[V], cf[V];
  ld(p0 + 2, 0, n0);
  ld(p0 + 1, -1, nm);
  ld(p0 + 1, 1, np);
  ld(p0, -2, nmm);
  ld(p0, 2, npp);
  const S* cfp = coeff + (static_cast<int64_t>(x0) * J + y) * K + k0;
  loadv<S, V>(cfp, cf);

  auto lap = [](C c, C ip, C im, C jp, C jm) {
    C nb = ip + im;
    nb = nb + jp;
    nb = nb + jm;
    return static_cast<C>(4.0) * c - nb;
  };
  C L0[V], flx_lo[V];  // carried: lap(p) and the flux across (p-1 | p)
  for (int x = x0; x < x1; ++x) {
    const int p = x + 2;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      c0[4][v] = n0[v];
      cm[2][v] = nm[v];
      cp[2][v] = np[v];
      cmm[v] = nmm[v];
      cpp[v] = npp[v];
    }
    C cfx[V];
#pragma unroll
    for (int v = 0; v < V; ++v) cfx[v] = cf[v];
    if (x + 1 < x1) {  // next row's loads, in flight while this row is computed
      ld(p + 3, 0, n0);
      ld(p + 2, -1, nm);
      ld(p + 2, 1, np);
      ld(p + 1, -2, nmm);
      ld(p + 1, 2, npp);
      loadv<S, V>(cfp + static_cast<int64_t>(x + 1 - x0) * J * K, cf);
    }
    C res[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const C c = c0[2][v];
      if (x == x0) {
        L0[v] = lap(c, c0[3][v], c0[1][v], cp[1][v], cm[1][v]);
        const C Lim = lap(c0[1][v], c, c0[0][v], cp[0][v], cm[0][v]);
        flx_lo[v] = limit<C>(L0[v] - Lim, c - c0[1][v]);
      }
      const C Lip = lap(c0[3][v], c0[4][v], c, cp[2][v], cm[2][v]);
      const C Ljp = lap(cp[1][v], cp[2][v], cp[0][v], cpp[v], c);
      const C Ljm = lap(cm[1][v], cm[2][v], cm[0][v], c, cmm[v]);
      const C flx_hi = limit<C>(Lip - L0[v], c0[3][v] - c);
      const C fly_hi = limit<C>(Ljp - L0[v], cp[1][v] - c);
      const C fly_lo = limit<C>(L0[v] - Ljm, c - cm[1][v]);
      C dv = flx_hi - flx_lo[v];
      dv = dv + fly_hi;
      dv = dv - fly_lo;
      res[v] = c - cfx[v] * dv;
      // carry to row p+1
      L0[v] = Lip;
      flx_lo[v] = flx_hi;
    }
    C o[V];
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = res[v];
    if (halo_dst)
      storev<S, V>(halo_dst + p * si + (y + 2) * sj + k0, o);
    else
      storev<S, V>(out + (static_cast<int64_t>(x) * J + y) * K + k0, o);
    // slide the window down one row
#pragma unroll
    for (int v = 0; v < V; ++v) {
#pragma unroll
      for (int r = 0; r < 4; ++r) c0[r][v] = c0[r + 1][v];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        cm[r][v] = cm[r + 1][v];
        cp[r][v] = cp[r + 1][v];
      }
    }
  }
}

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

// Fused steps per round (tuning knob "hdiff.tblock"; 1 = one launch per step,
// the default: at 256x256x160 on the MI355X the fused kernel is slower, its
// lanes each reading 16 bytes of a different 640-byte k-column).
std::atomic<int> g_hdiff_tblock{1};

template <class S>
constexpr int hdiff_tb_max() { return sizeof(S) == 2 ? 2 : 4; }  // BF16 x4 exceeds LDS

template <class S>
int run_hdiff(const S* in, S* out, const S* coeff, int I, int J, int K, int64_t nsteps,
              hipStream_t stream) {
  constexpr int V = march_vec<S>();
  const bool vec = K % V == 0 && aligned16(in) && aligned16(out) && aligned16(coeff);
  const int xblocks = static_cast<int>(cdiv(static_cast<int64_t>(J) * (K / V), 256));
  const int64_t mblocks = int64_t(xblocks) * cdiv(I, march_strip<S>());
  NPBX_REQUIRE(!vec || mblocks < (int64_t(1) << 31), "hdiff: grid too large");
  auto launch = [&](const S* src, S* next) -> int {
    if (vec) {
      hipLaunchKernelGGL(hdiff_march<S>, dim3(static_cast<unsigned>(mblocks)), dim3(256), 0,
                         stream, src, out, coeff, I, J, K, xblocks, next);
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
  const S* src = in;
  if (tb >= 2 && vec) {
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
      } else {
        go(hdiff_tb<S, 2>);
      }
      NPBX_LAUNCHED("hdiff_tb");
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
