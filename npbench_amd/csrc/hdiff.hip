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
// dst; both are (I+4, J+4, K).
template <class S>
__global__ __launch_bounds__(256) void hdiff_copy_halo(const S* __restrict__ src,
                                                       S* __restrict__ dst, int I, int J, int K) {
  const int64_t total = static_cast<int64_t>(I + 4) * (J + 4) * K;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t col = e / K;
    const int p = static_cast<int>(col / (J + 4));
    const int q = static_cast<int>(col % (J + 4));
    if (p < 2 || p >= I + 2 || q < 2 || q >= J + 2) dst[e] = src[e];
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
  // Two work fields ping-pong; each carries the original halo.
  const size_t field = static_cast<size_t>(I + 4) * (J + 4) * K;
  void* ws = nullptr;
  NPBX_TRY(workspace(2 * field * sizeof(S), &ws));
  S* w[2] = {static_cast<S*>(ws), static_cast<S*>(ws) + field};
  const unsigned copy_blocks = static_cast<unsigned>(std::min<int64_t>(cdiv(field, 256), 4096));
  for (int b = 0; b < 2; ++b) {
    hipLaunchKernelGGL(hdiff_copy_halo<S>, dim3(copy_blocks), dim3(256), 0, stream, in, w[b], I, J,
                       K);
    NPBX_LAUNCHED("hdiff_copy_halo");
  }
  const S* src = in;
  for (int64_t n = 0; n < nsteps; ++n) {
    const bool last = n + 1 == nsteps;
    S* next = w[n % 2];
    NPBX_TRY(launch(src, last ? nullptr : next));
    src = next;
  }
  return NPBX_OK;
}

}  // namespace
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
