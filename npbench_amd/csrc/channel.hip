// Channel flow of npbench's channel_flow
// (npbench/benchmarks/channel_flow/channel_flow_numpy.py:13-152), FP64: the
// cavity-flow time step with periodic columns (x-neighbours wrap around),
// walls at rows 0 and ny-1, a body force F, and time steps repeated while
// the relative change of sum(u), (sum(u) - sum(un)) / sum(u), exceeds 0.001.
//
// Like cavity_flow, one workgroup runs every phase with a barrier between
// phases.  The reference computes columns 0 and nx-1 with the interior's
// expressions and wrapped neighbours (:24-37, :53-62, :104-141), so every
// column of rows 1..ny-2 is one interior point here.  After each pressure
// iteration rows ny-1 and 0 copy rows ny-2 and 1 (:65-66): the threads that
// compute those rows write the copies.  Grids of up to kLdsPoints points keep
// both pressure buffers in LDS and b in registers (channel_flow_lds).
//
// The reference loops without a bound.  A launch here runs at most
// max_steps time steps and records how many it ran and whether the
// criterion stopped it; the host launches again until it has, so no launch
// runs unboundedly.  The sums add each thread's points in order, then the
// lanes of a wave, then the waves in order: a fixed order, but not NumPy's
// pairwise blocks, so a step whose udiff lies within rounding of 0.001 could
// be decided differently; the step count is checked against the oracle.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "division.h"

namespace npbx {
namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kLdsPoints = 10240;              // two FP64 pressure buffers: 160 KiB
constexpr int kMaxPts = kLdsPoints / kThreads;  // interior points per thread, LDS grids
constexpr int kChunkP = 6;
constexpr int kChunkV = 1;  // velocity points per chunk (14 loads each, and the sums)

struct ChannelArgs {
  double* u;
  double* v;
  double* p;
  double* un;
  double* vn;
  double* pn;
  double* b;
  double* red;   // kWaves partial sums
  double* stat;  // [0] steps run, [1] 1 when the criterion stopped the loop
  int nx, ny, nit, max_steps;
  double rho, nu, F;
  double inv_dt, two_dx, two_dy, dx2, dy2, den, cb, dt, dx, dy, cpu, cpv, cdx2, cdy2;
  double fdt;  // F * dt
};

struct Divisors {
  Divisor den, two_dx, two_dy, dx, dy;
  double lo, hi;
};

__device__ __forceinline__ Divisors prepare_divisors(const ChannelArgs& g) {
  double lo = 0.0, hi = INFINITY;
  Divisors dv;
  dv.den = prepare_divisor(g.den, lo, hi);
  dv.two_dx = prepare_divisor(g.two_dx, lo, hi);
  dv.two_dy = prepare_divisor(g.two_dy, lo, hi);
  dv.dx = prepare_divisor(g.dx, lo, hi);
  dv.dy = prepare_divisor(g.dy, lo, hi);
  dv.lo = uniform_double(lo);
  dv.hi = uniform_double(hi);
  return dv;
}

// A point of rows 1..ny-2, packed in one int: its flat index x (< 2^26)
// shifted left by 4 over four flags -- kTop: row 1 (row 0 copies it),
// kBottom: row ny-2 (row ny-1 copies it), kWest / kEast: column 0 / nx-1,
// whose x-neighbour wraps to the other end of the row.
enum : int { kTop = 1, kBottom = 2, kWest = 4, kEast = 8 };
struct Pt {
  int pk;
  __device__ __forceinline__ int x() const { return pk >> 4; }
  __device__ __forceinline__ int fl() const { return pk & 15; }
  __device__ __forceinline__ int e(int nx) const { return x() + ((pk & kEast) ? 1 - nx : 1); }
  __device__ __forceinline__ int w(int nx) const { return x() + ((pk & kWest) ? nx - 1 : -1); }
};

struct Walk {
  int i0, j0, si, sj, w, h, nx;
  __device__ Walk(int t, int nx_, int ny_) : nx(nx_) {
    w = nx_;
    h = ny_ - 2;
    i0 = t / w;
    j0 = t - i0 * w;
    si = kThreads / w;
    sj = kThreads - si * w;
  }
  __device__ __forceinline__ Pt point(int i, int j, bool ok) const {
    if (!ok) {  // a safe point to read, never written
      i = 0;
      j = 1;
    }
    const int fl = (ok && i == 0 ? kTop : 0) | (ok && i == h - 1 ? kBottom : 0) |
                   (j == 0 ? kWest : 0) | (j == nx - 1 ? kEast : 0);
    return {(((i + 1) * nx + j) << 4) | fl};
  }
  template <int C, class F>
  __device__ __forceinline__ void each(F&& f) const {
    int i = i0, j = j0;
    while (i < h) {
      Pt q[C];
      bool ok[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        ok[c] = i < h;
        q[c] = point(i, j, ok[c]);
        i += si;
        j += sj;
        if (j >= w) {
          j -= w;
          ++i;
        }
      }
      f(q, ok);
    }
  }
};

// build_up_b (:13-39) at a point.
template <class Div>
__device__ __forceinline__ double source_at(const ChannelArgs& g, const Divisors& dv, const Pt& q,
                                            Div& div) {
  const int nx = g.nx, x = q.x(), e = q.e(nx), w = q.w(nx);
  const double* u = g.u;
  const double* v = g.v;
  const double dudx = div(u[e] - u[w], dv.two_dx);
  const double dvdy = div(v[x + nx] - v[x - nx], dv.two_dy);
  const double t1 = g.inv_dt * (dudx + dvdy);
  const double t2 = dudx * dudx;
  const double t3 = 2.0 * div(div(u[x + nx] - u[x - nx], dv.two_dy) * (v[e] - v[w]), dv.two_dx);
  const double t4 = dvdy * dvdy;
  return g.rho * (((t1 - t2) - t3) - t4);
}

// One pressure iterate (:47-62) at a point.
template <class P, class Div>
__device__ __forceinline__ double pressure_at(const ChannelArgs& g, const Divisors& dv,
                                              const P* src, const Pt& q, double bx, Div& div) {
  const int nx = g.nx, x = q.x();
  return div((src[q.e(nx)] + src[q.w(nx)]) * g.dy2 + (src[x + nx] + src[x - nx]) * g.dx2, dv.den) -
         g.cb * bx;
}

template <class P>
__device__ __forceinline__ void store_pressure(P* dst, const Pt& q, int nx, double val) {
  const int x = q.x();
  dst[x] = val;
  if (q.pk & kTop) dst[x - nx] = val;     // p[0, :] = p[1, :]
  if (q.pk & kBottom) dst[x + nx] = val;  // p[-1, :] = p[-2, :]
}

// u (:80-90, :104-121) and v (:92-101, :124-141) at a point.
template <class P, class Div>
__device__ __forceinline__ void velocity_at(const ChannelArgs& g, const Divisors& dv, const P* p,
                                            const Pt& q, double& nu, double& nv, Div& div) {
  const int nx = g.nx, x = q.x(), e = q.e(nx), w = q.w(nx);
  const double* un = g.un;
  const double* vn = g.vn;
  const double uc = un[x], vc = vn[x];
  const double udx = div(uc * g.dt, dv.dx);
  const double vdy = div(vc * g.dt, dv.dy);
  const double x1 = udx * (uc - un[w]);
  const double x2 = vdy * (uc - un[x - nx]);
  const double x3 = g.cpu * (p[e] - p[w]);
  const double x4 = g.nu * (g.cdx2 * (un[e] - 2.0 * uc + un[w]) +
                            g.cdy2 * (un[x + nx] - 2.0 * uc + un[x - nx]));
  nu = (((uc - x1 - x2) - x3) + x4) + g.fdt;
  const double y1 = udx * (vc - vn[w]);
  const double y2 = vdy * (vc - vn[x - nx]);
  const double y3 = g.cpv * (p[x + nx] - p[x - nx]);
  const double y4 = g.nu * (g.cdx2 * (vn[e] - 2.0 * vc + vn[w]) +
                            g.cdy2 * (vn[x + nx] - 2.0 * vc + vn[x - nx]));
  nv = ((vc - y1 - y2) - y3) + y4;
}

// Sums each thread's value over the workgroup in a fixed order: the lanes of
// a wave by a butterfly, then the wave totals in wave order.  Ends with a
// barrier; every thread returns the same total.
__device__ __forceinline__ double block_sum(double s, double* red, int t) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) tot += red[k];
  __syncthreads();  // red may be written again by the next sum
  return tot;
}

// un = u.copy(); vn = v.copy() (:74-75), with sum(un) on the way.
__device__ __forceinline__ double copy_velocities(const ChannelArgs& g, int t, int np) {
  constexpr int U = 4;
  double s = 0.0;
  for (int k0 = t; k0 < np; k0 += U * kThreads) {
    double uu[U], vv[U];
#pragma unroll
    for (int c = 0; c < U; ++c) {
      const int k = min(k0 + c * kThreads, np - 1);
      uu[c] = g.u[k];
      vv[c] = g.v[k];
    }
#pragma unroll
    for (int c = 0; c < U; ++c) {
      const int k = k0 + c * kThreads;
      if (k < np) {
        g.un[k] = uu[c];
        g.vn[k] = vv[c];
        s += uu[c];
      }
    }
  }
  return s;
}

// u, v = 0 on the walls (:144-147).
__device__ __forceinline__ void wall_velocities(const ChannelArgs& g, int t) {
  const int nx = g.nx, ny = g.ny;
  for (int k = t; k < 2 * nx; k += kThreads) {
    const int a = k < nx ? k : (ny - 1) * nx + (k - nx);
    g.u[a] = 0.0;
    g.v[a] = 0.0;
  }
}

// Records the launch's outcome: steps run and whether the criterion ended it.
__device__ __forceinline__ void record(const ChannelArgs& g, int t, int steps, bool done) {
  if (t == 0) {
    g.stat[0] = static_cast<double>(steps);
    g.stat[1] = done ? 1.0 : 0.0;
  }
}

// Grids of more than kLdsPoints points: pressure in p and pn, b in memory.
__global__ __launch_bounds__(kThreads) void channel_flow_global(ChannelArgs g) {
  const int t = threadIdx.x, nx = g.nx, ny = g.ny, np = nx * ny;
  const Walk walk(t, nx, ny);
  const Divisors dv = prepare_divisors(g);
  double* cur = g.p;
  double* nxt = g.pn;
  int steps = 0;
  bool done = false;
  while (steps < g.max_steps && !done) {
    const double sum_un = block_sum(copy_velocities(g, t, np), g.red, t);
    walk.each<4>([&](const Pt(&q)[4], const bool(&ok)[4]) {
      double bv[4];
      chunk_eval(dv, bv, [&](int c, auto& div) { return source_at(g, dv, q[c], div); });
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (ok[c]) g.b[q[c].x()] = bv[c];
    });
    __syncthreads();
    for (int it = 0; it < g.nit; ++it) {
      walk.each<kChunkP>([&](const Pt(&q)[kChunkP], const bool(&ok)[kChunkP]) {
        double val[kChunkP];
        chunk_eval(dv, val, [&](int c, auto& div) {
          return pressure_at(g, dv, cur, q[c], g.b[q[c].x()], div);
        });
#pragma unroll
        for (int c = 0; c < kChunkP; ++c)
          if (ok[c]) store_pressure(nxt, q[c], nx, val[c]);
      });
      __syncthreads();
      double* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    double su = 0.0;
    walk.each<kChunkV>([&](const Pt(&q)[kChunkV], const bool(&ok)[kChunkV]) {
      double nu[kChunkV], nv[kChunkV];
      chunk_eval(dv, nu, [&](int c, auto& div) {
        double uu;
        velocity_at(g, dv, cur, q[c], uu, nv[c], div);
        return uu;
      });
#pragma unroll
      for (int c = 0; c < kChunkV; ++c) {
        if (ok[c]) {
          g.u[q[c].x()] = nu[c];
          g.v[q[c].x()] = nv[c];
          su += nu[c];
        }
      }
    });
    wall_velocities(g, t);
    const double sum_u = block_sum(su, g.red, t);  // the walls add zeros
    ++steps;
    done = !((sum_u - sum_un) / sum_u > .001);
  }
  if (cur != g.p) {
    for (int k = t; k < np; k += kThreads) g.p[k] = cur[k];
  }
  record(g, t, steps, done);
}

// Grids of up to kLdsPoints points: both pressure buffers in LDS, each
// thread's points and b values in registers.
__global__ __launch_bounds__(kThreads) void channel_flow_lds(ChannelArgs g) {
  __shared__ double lp[2 * kLdsPoints];
  const int t = threadIdx.x, nx = g.nx, ny = g.ny, np = nx * ny;
  const Walk walk(t, nx, ny);
  const Divisors dv = prepare_divisors(g);
  Pt pts[kMaxPts];
#pragma unroll
  for (int m = 0; m < kMaxPts; ++m) pts[m] = walk.point(0, 1, false);
  int cnt = 0;
  walk.each<kMaxPts>([&](const Pt(&q)[kMaxPts], const bool(&ok)[kMaxPts]) {
#pragma unroll
    for (int m = 0; m < kMaxPts; ++m) {
      pts[m] = q[m];
      cnt += ok[m] ? 1 : 0;
    }
  });
  for (int k = t; k < np; k += kThreads) lp[k] = g.p[k];
  double* cur = lp;
  double* nxt = lp + kLdsPoints;
  __syncthreads();
  int steps = 0;
  bool done = false;
  while (steps < g.max_steps && !done) {
    // opaque point data: the compiler recomputes addresses each step instead
    // of holding every one of them (and the lane masks) in registers
#pragma unroll
    for (int m = 0; m < kMaxPts; ++m) asm volatile("" : "+v"(pts[m].pk));
    asm volatile("" : "+v"(cnt));
    const double sum_un = block_sum(copy_velocities(g, t, np), g.red, t);
    double bq[kMaxPts];
#pragma unroll
    for (int m0 = 0; m0 < kMaxPts; m0 += kChunkV) {
      double bc[kChunkV] = {};
      if (m0 < cnt)
        chunk_eval(dv, bc, [&](int c, auto& div) { return source_at(g, dv, pts[m0 + c], div); });
#pragma unroll
      for (int c = 0; c < kChunkV; ++c) bq[m0 + c] = bc[c];
      asm volatile("" ::: "memory");
    }
    for (int it = 0; it < g.nit; ++it) {
#pragma unroll
      for (int m = 0; m < kMaxPts; ++m) asm volatile("" : "+v"(pts[m].pk));
      asm volatile("" : "+v"(cnt));
#pragma unroll
      for (int m0 = 0; m0 < kMaxPts; m0 += kMaxPts / 2) {
        if (m0 >= cnt) continue;
        double val[kMaxPts / 2];
        chunk_eval(dv, val, [&](int c, auto& div) {
          return pressure_at(g, dv, cur, pts[m0 + c], bq[m0 + c], div);
        });
#pragma unroll
        for (int c = 0; c < kMaxPts / 2; ++c)
          if (m0 + c < cnt) store_pressure(nxt, pts[m0 + c], nx, val[c]);
      }
      __syncthreads();
      double* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    double su = 0.0;
#pragma unroll
    for (int m0 = 0; m0 < kMaxPts; m0 += kChunkV) {
      if (m0 >= cnt) continue;
      double nu[kChunkV], nv[kChunkV];
      chunk_eval(dv, nu, [&](int c, auto& div) {
        double uu;
        velocity_at(g, dv, cur, pts[m0 + c], uu, nv[c], div);
        return uu;
      });
#pragma unroll
      for (int c = 0; c < kChunkV; ++c) {
        if (m0 + c < cnt) {
          g.u[pts[m0 + c].x()] = nu[c];
          g.v[pts[m0 + c].x()] = nv[c];
          su += nu[c];
        }
      }
    }
    wall_velocities(g, t);
    const double sum_u = block_sum(su, g.red, t);
    ++steps;
    done = !((sum_u - sum_un) / sum_u > .001);
  }
  for (int k = t; k < np; k += kThreads) g.p[k] = cur[k];
  record(g, t, steps, done);
}

}  // namespace
}  // namespace npbx

namespace npbx {
namespace {

// Checks the arguments and fills the launch arguments (scratch from the
// library workspace).
int channel_setup(int dtype, int64_t nit, void* u, void* v, double dt, double dx, double dy,
                  void* p, double rho, double nu, double F, int64_t nx, int64_t ny,
                  int64_t max_steps, ChannelArgs& g) {
  NPBX_REQUIRE(dtype == NPBX_F64, "channel_flow: dtype %s not supported (f64)",
               dtype_name(dtype));
  NPBX_REQUIRE(nx >= 3 && ny >= 3 && nx * ny <= (int64_t(1) << 26),
               "channel_flow: grid %lld x %lld out of range (3 x 3 .. 2^26 points)",
               static_cast<long long>(ny), static_cast<long long>(nx));
  NPBX_REQUIRE(nit >= 0 && nit <= (int64_t(1) << 30), "channel_flow: nit out of range");
  NPBX_REQUIRE(max_steps >= 1 && max_steps <= (int64_t(1) << 30),
               "channel_flow: max_steps out of range (1 .. 2^30)");
  NPBX_REQUIRE(u && v && p && u != v && v != p && u != p,
               "channel_flow: u, v, p must be distinct non-null arrays");
  const int64_t np = nx * ny;
  void* ws = nullptr;
  const int64_t stat_off = 4 * np, red_off = stat_off + 64;
  NPBX_TRY(workspace(static_cast<size_t>(red_off + kWaves) * sizeof(double), &ws));
  g.u = static_cast<double*>(u);
  g.v = static_cast<double*>(v);
  g.p = static_cast<double*>(p);
  g.un = static_cast<double*>(ws);
  g.vn = g.un + np;
  g.pn = g.vn + np;
  g.b = g.pn + np;
  g.stat = g.un + stat_off;
  g.red = g.un + red_off;
  g.nx = static_cast<int>(nx);
  g.ny = static_cast<int>(ny);
  g.nit = static_cast<int>(nit);
  g.rho = rho;
  g.nu = nu;
  g.F = F;
  g.dt = dt;
  g.dx = dx;
  g.dy = dy;
  g.inv_dt = 1.0 / dt;
  g.two_dx = 2.0 * dx;
  g.two_dy = 2.0 * dy;
  // CPython's x**2 calls the C library's pow; an opaque exponent keeps the
  // compiler from folding the call into x * x
  volatile double two = 2.0;
  g.dx2 = std::pow(dx, static_cast<double>(two));
  g.dy2 = std::pow(dy, static_cast<double>(two));
  g.den = 2.0 * (g.dx2 + g.dy2);
  g.cb = g.dx2 * g.dy2 / g.den;
  g.cpu = dt / (2.0 * rho * dx);
  g.cpv = dt / (2.0 * rho * dy);
  g.cdx2 = dt / g.dx2;
  g.cdy2 = dt / g.dy2;
  g.fdt = F * dt;
  g.max_steps = static_cast<int>(max_steps);
  return NPBX_OK;
}

// One launch of at most g.max_steps time steps, continuing from u, v, p; waits
// for it and returns the steps it ran and whether the criterion stopped it.
int channel_launch(const ChannelArgs& g, hipStream_t s, int64_t* steps, bool* done) {
  if (static_cast<int64_t>(g.nx) * g.ny <= kLdsPoints) {
    hipLaunchKernelGGL(channel_flow_lds, dim3(1), dim3(kThreads), 0, s, g);
    NPBX_LAUNCHED("channel_flow_lds");
  } else {
    hipLaunchKernelGGL(channel_flow_global, dim3(1), dim3(kThreads), 0, s, g);
    NPBX_LAUNCHED("channel_flow_global");
  }
  double stat[2];
  NPBX_HIP(hipMemcpyAsync(stat, g.stat, sizeof(stat), hipMemcpyDeviceToHost, s));
  NPBX_HIP(hipStreamSynchronize(s));
  *steps = static_cast<int64_t>(stat[0]);
  *done = stat[1] != 0.0;
  return NPBX_OK;
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_channel_flow(int dtype, int64_t nit, void* u, void* v, double dt, double dx,
                                 double dy, void* p, double rho, double nu, double F, int64_t nx,
                                 int64_t ny, int64_t max_steps, int64_t* stepcount,
                                 void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(stepcount != nullptr, "channel_flow: null stepcount");
  ChannelArgs g;
  NPBX_TRY(channel_setup(dtype, nit, u, v, dt, dx, dy, p, rho, nu, F, nx, ny, max_steps, g));
  // launches of at most max_steps steps until the criterion stops one
  int64_t total = 0;
  for (;;) {
    int64_t steps = 0;
    bool done = false;
    NPBX_TRY(channel_launch(g, as_stream(stream), &steps, &done));
    total += steps;
    if (done) break;
  }
  *stepcount = total;
  return NPBX_OK;
}

extern "C" int npbx_channel_flow_launch(int dtype, int64_t nit, void* u, void* v, double dt,
                                        double dx, double dy, void* p, double rho, double nu,
                                        double F, int64_t nx, int64_t ny, int64_t max_steps,
                                        int64_t* steps, int32_t* converged, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(steps != nullptr && converged != nullptr, "channel_flow_launch: null output");
  ChannelArgs g;
  NPBX_TRY(channel_setup(dtype, nit, u, v, dt, dx, dy, p, rho, nu, F, nx, ny, max_steps, g));
  bool done = false;
  NPBX_TRY(channel_launch(g, as_stream(stream), steps, &done));
  *converged = done ? 1 : 0;
  return NPBX_OK;
}
