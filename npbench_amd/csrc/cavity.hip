// Lid-driven cavity flow of npbench's cavity_flow
// (npbench/benchmarks/cavity_flow/cavity_flow_numpy.py:13-83), FP64, bitwise
// the NumPy reference.
//
// A time step is a chain of whole-grid phases -- copy u, v; build b; nit
// pressure-Poisson Jacobi iterations; update u, v -- on grids of ~10^4 points
// (the paper preset is 101 x 101 with 700 steps of 50 iterations: 37,000
// phases).  One launch per phase would be bound by launches, so one workgroup
// of kThreadsC threads runs every phase in a loop with a barrier between
// phases.  The pressure alternates between two buffers (the reference's
// pn = p.copy() each iteration).  Grids of up to kLdsPoints points (the paper
// and S presets) keep both buffers in LDS for the whole run and b in
// registers, so an iteration never leaves the CU (cavity_flow_lds); larger
// grids use p and a device scratch copy, which the caches hold
// (cavity_flow_global).  Divisions are by five uniform divisors whose
// reciprocals are prepared once (division.h).  b needs no zeroing: only its
// interior is read, and each step writes it first.
//
// The four boundary statements after each iteration (:30-33), applied in the
// reference's order, leave row 0, column 0 and column nx-1 equal to this
// iteration's interior next to them (the corners (0,0) and (0,nx-1) take
// (1,1) and (1,nx-2)) and the last row 0.  So the thread that computes an
// interior point next to one of those edges also writes the edge copies, and
// an iteration is a single phase.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.h"
#include "division.h"

namespace npbx {
namespace {

constexpr int kThreadsC = 1024;
constexpr int kLdsPoints = 10240;  // two FP64 pressure buffers: all 160 KiB of LDS

// The reference's scalar expressions, computed on the host as Python computes
// them (x**2 is the C library's pow, which CPython's float power calls).
struct CavityArgs {
  double* u;
  double* v;
  double* p;
  double* un;
  double* vn;
  double* pn;
  double* b;
  int nx, ny, nt, nit;
  double rho, nu;
  double inv_dt;    // 1 / dt
  double two_dx;    // 2 * dx
  double two_dy;    // 2 * dy
  double dx2, dy2;  // dx**2, dy**2
  double den;       // 2 * (dx**2 + dy**2)
  double cb;        // dx**2 * dy**2 / (2 * (dx**2 + dy**2))
  double dt, dx, dy;
  double cpu;       // dt / (2 * rho * dx)
  double cpv;       // dt / (2 * rho * dy)
  double cdx2;      // dt / dx**2
  double cdy2;      // dt / dy**2
};

// The kernel's divisors, prepared once per launch and held in scalar
// registers: every division of the time step is by one of them.  Only their
// reciprocals and one common no-rescale range of |m| stay live, so the
// kernel's uniform values fit the scalar register file.
struct CavityDivisors {
  Divisor den, two_dx, two_dy, dx, dy;
  double lo, hi;  // |m| in [lo, hi) for every division of a chunk: exact
};

// Walks this thread's interior points k = t, t + kThreadsC, ... in row-major
// order of the (ny-2) x (nx-2) interior without a division per point, C
// points at a time: f receives their flat grid indices a, edge flags and
// validity (past the end, a is a safe interior index to read, not to write).
// Handing a chunk to f at once lets it issue every load of the chunk before
// its first store, so the loads' latencies overlap.
enum : int { kRow1 = 1, kCol1 = 2, kColLast = 4 };  // interior row 0, column 0, column w-1

struct InteriorWalk {
  int i0, j0, si, sj, w, h, nx;
  __device__ InteriorWalk(int t, int nx_, int ny_) : nx(nx_) {
    w = nx_ - 2;
    h = ny_ - 2;
    i0 = t / w;
    j0 = t - i0 * w;
    si = kThreadsC / w;
    sj = kThreadsC - si * w;
  }
  template <int C, class F>
  __device__ __forceinline__ void each(F&& f) const {
    int i = i0, j = j0;
    while (i < h) {
      int a[C], fl[C];
      bool ok[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        ok[c] = i < h;
        a[c] = ok[c] ? (i + 1) * nx + j + 1 : nx + 1;
        fl[c] = (i == 0 ? kRow1 : 0) | (j == 0 ? kCol1 : 0) | (j == w - 1 ? kColLast : 0);
        i += si;
        j += sj;
        if (j >= w) {
          j -= w;
          ++i;
        }
      }
      f(a, fl, ok);
    }
  }
};

constexpr int kChunkP = 6;  // pressure points per chunk (5 loads each)
constexpr int kChunkB = 4;  // source points per chunk (8 loads each)
constexpr int kChunkV = 2;  // velocity points per chunk (14 loads each)
constexpr int kMaxPts = kLdsPoints / kThreadsC;  // interior points per thread, LDS grids

// The point computations take their division as a policy (ShortDiv or
// FullDiv, run by chunk_eval).

// build_up_b (:13-21) at interior point x.
template <class Div>
__device__ __forceinline__ double source_at(const CavityArgs& g, const CavityDivisors& dv, int x,
                                            Div& div) {
  const int nx = g.nx;
  const double* u = g.u;
  const double* v = g.v;
  const double dudx = div(u[x + 1] - u[x - 1], dv.two_dx);
  const double dvdy = div(v[x + nx] - v[x - nx], dv.two_dy);
  const double t1 = g.inv_dt * (dudx + dvdy);
  const double t2 = dudx * dudx;
  const double t3 = 2.0 * div(div(u[x + nx] - u[x - nx], dv.two_dy) * (v[x + 1] - v[x - 1]), dv.two_dx);
  const double t4 = dvdy * dvdy;
  return g.rho * (((t1 - t2) - t3) - t4);
}

// One pressure iterate (:30-33) at interior point x, with b there.
template <class P, class Div>
__device__ __forceinline__ double pressure_at(const CavityArgs& g, const CavityDivisors& dv,
                                              const P* src, int x, double bx, Div& div) {
  const int nx = g.nx;
  return div((src[x + 1] + src[x - 1]) * g.dy2 + (src[x + nx] + src[x - nx]) * g.dx2, dv.den) -
         g.cb * bx;
}

// Stores an iterate at interior point x and, for points next to the edges,
// the copies the boundary statements (:35-37) make of it.
template <class P>
__device__ __forceinline__ void store_pressure(P* dst, int x, int fl, int nx, double val) {
  dst[x] = val;
  if (!fl) return;                                // most points have no edge copy
  if (fl & kRow1) dst[x - nx] = val;              // row 0 = row 1
  if (fl & kCol1) {
    dst[x - 1] = val;                             // column 0 = column 1
    if (fl & kRow1) dst[x - nx - 1] = val;        // corner (0,0) = (1,1)
  }
  if (fl & kColLast) {
    dst[x + 1] = val;                             // column nx-1 = column nx-2
    if (fl & kRow1) dst[x - nx + 1] = val;        // corner (0,nx-1) = (1,nx-2)
  }
}

// u and v (:53-73) at interior point x from un, vn and the pressure p.
template <class P, class Div>
__device__ __forceinline__ void velocity_at(const CavityArgs& g, const CavityDivisors& dv,
                                            const P* p, int x, double& nu, double& nv, Div& div) {
  const int nx = g.nx;
  const double* un = g.un;
  const double* vn = g.vn;
  const double uc = un[x], vc = vn[x];
  const double udx = div(uc * g.dt, dv.dx);  // un * dt / dx
  const double vdy = div(vc * g.dt, dv.dy);  // vn * dt / dy
  const double x1 = udx * (uc - un[x - 1]);
  const double x2 = vdy * (uc - un[x - nx]);
  const double x3 = g.cpu * (p[x + 1] - p[x - 1]);
  const double x4 = g.nu * (g.cdx2 * (un[x + 1] - 2.0 * uc + un[x - 1]) +
                            g.cdy2 * (un[x + nx] - 2.0 * uc + un[x - nx]));
  nu = ((uc - x1 - x2) - x3) + x4;
  const double y1 = udx * (vc - vn[x - 1]);
  const double y2 = vdy * (vc - vn[x - nx]);
  const double y3 = g.cpv * (p[x + nx] - p[x - nx]);
  const double y4 = g.nu * (g.cdx2 * (vn[x + 1] - 2.0 * vc + vn[x - 1]) +
                            g.cdy2 * (vn[x + nx] - 2.0 * vc + vn[x - nx]));
  nv = ((vc - y1 - y2) - y3) + y4;
}

// The velocity boundary conditions (:75-82): the lid row is 1, corners
// included, every other edge point 0.
__device__ __forceinline__ void velocity_boundaries(const CavityArgs& g, int t) {
  const int nx = g.nx, ny = g.ny;
  for (int k = t; k < 2 * nx + 2 * ny; k += kThreadsC) {
    int a;
    if (k < nx) a = k;                                    // row 0
    else if (k < 2 * nx) a = (ny - 1) * nx + (k - nx);    // row ny-1
    else if (k < 2 * nx + ny) a = (k - 2 * nx) * nx;      // column 0
    else a = (k - 2 * nx - ny) * nx + nx - 1;             // column nx-1
    g.u[a] = a >= (ny - 1) * nx ? 1.0 : 0.0;
    g.v[a] = 0.0;
  }
}

// un = u.copy(); vn = v.copy() (:47-48), loads ahead of stores.
__device__ __forceinline__ void copy_velocities(const CavityArgs& g, int t, int np) {
  constexpr int U = 4;
  for (int k0 = t; k0 < np; k0 += U * kThreadsC) {
    double uu[U], vv[U];
#pragma unroll
    for (int c = 0; c < U; ++c) {
      const int k = min(k0 + c * kThreadsC, np - 1);
      uu[c] = g.u[k];
      vv[c] = g.v[k];
    }
#pragma unroll
    for (int c = 0; c < U; ++c) {
      const int k = k0 + c * kThreadsC;
      if (k < np) {
        g.un[k] = uu[c];
        g.vn[k] = vv[c];
      }
    }
  }
}

__device__ __forceinline__ CavityDivisors cavity_divisors(const CavityArgs& g) {
  double lo = 0.0, hi = INFINITY;
  CavityDivisors dv;
  dv.den = prepare_divisor(g.den, lo, hi);
  dv.two_dx = prepare_divisor(g.two_dx, lo, hi);
  dv.two_dy = prepare_divisor(g.two_dy, lo, hi);
  dv.dx = prepare_divisor(g.dx, lo, hi);
  dv.dy = prepare_divisor(g.dy, lo, hi);
  dv.lo = uniform_double(lo);
  dv.hi = uniform_double(hi);
  return dv;
}

// Grids of more than kLdsPoints points: the pressure alternates between p and
// the scratch copy pn, b lives in device memory, and the threads walk their
// points in chunks.
__global__ __launch_bounds__(kThreadsC) void cavity_flow_global(CavityArgs g) {
  const int t = threadIdx.x, nx = g.nx, ny = g.ny;
  const int np = nx * ny;
  const InteriorWalk walk(t, nx, ny);
  const CavityDivisors dv = cavity_divisors(g);
  double* cur = g.p;  // the current pressure
  double* nxt = g.pn;
  for (int step = 0; step < g.nt; ++step) {
    copy_velocities(g, t, np);
    // b's interior, from u and v, which the copy above only reads too
    walk.each<kChunkB>([&](const int(&a)[kChunkB], const int(&)[kChunkB],
                           const bool(&ok)[kChunkB]) {
      double bv[kChunkB];
      chunk_eval(dv, bv, [&](int c, auto& div) { return source_at(g, dv, a[c], div); });
#pragma unroll
      for (int c = 0; c < kChunkB; ++c)
        if (ok[c]) g.b[a[c]] = bv[c];
    });
    __syncthreads();
    for (int q = 0; q < g.nit; ++q) {
      walk.each<kChunkP>([&](const int(&a)[kChunkP], const int(&fl)[kChunkP],
                             const bool(&ok)[kChunkP]) {
        double val[kChunkP];
        chunk_eval(dv, val, [&](int c, auto& div) { return pressure_at(g, dv, cur, a[c], g.b[a[c]], div); });
#pragma unroll
        for (int c = 0; c < kChunkP; ++c)
          if (ok[c]) store_pressure(nxt, a[c], fl[c], nx, val[c]);
      });
      for (int k = t; k < nx; k += kThreadsC) nxt[(ny - 1) * nx + k] = 0.0;  // p[-1, :] = 0
      __syncthreads();
      double* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    walk.each<kChunkV>([&](const int(&a)[kChunkV], const int(&)[kChunkV],
                           const bool(&ok)[kChunkV]) {
      double nu[kChunkV], nv[kChunkV];
      chunk_eval(dv, nu, [&](int c, auto& div) {
        double uu;
        velocity_at(g, dv, cur, a[c], uu, nv[c], div);
        return uu;
      });
#pragma unroll
      for (int c = 0; c < kChunkV; ++c) {
        if (ok[c]) {
          g.u[a[c]] = nu[c];
          g.v[a[c]] = nv[c];
        }
      }
    });
    velocity_boundaries(g, t);
    __syncthreads();
  }
  if (cur != g.p) {  // an odd number of iterations in all: the result is in pn
    for (int k = t; k < np; k += kThreadsC) g.p[k] = cur[k];
  }
}

// Grids of up to kLdsPoints points: both pressure buffers live in LDS for the
// whole run, each thread keeps its (at most kMaxPts) interior points' indices,
// edge flags and b values in registers -- it computes b for the same points
// it then iterates -- so a pressure iteration touches LDS only.
__global__ __launch_bounds__(kThreadsC) void cavity_flow_lds(CavityArgs g) {
  __shared__ double lp[2 * kLdsPoints];
  const int t = threadIdx.x, nx = g.nx, ny = g.ny;
  const int np = nx * ny;
  const CavityDivisors dv = cavity_divisors(g);
  const InteriorWalk walk(t, nx, ny);
  int pa[kMaxPts];  // past cnt: the safe interior index nx + 1, read and never written
#pragma unroll
  for (int m = 0; m < kMaxPts; ++m) pa[m] = nx + 1;
  int pfl = 0;  // 3 edge-flag bits per point
  int cnt = 0;
  walk.each<kMaxPts>([&](const int(&a)[kMaxPts], const int(&fl)[kMaxPts],
                         const bool(&ok)[kMaxPts]) {
#pragma unroll
    for (int m = 0; m < kMaxPts; ++m) {
      pa[m] = a[m];
      pfl |= (ok[m] ? fl[m] : 0) << (3 * m);
      cnt += ok[m] ? 1 : 0;
    }
  });
  for (int k = t; k < np; k += kThreadsC) lp[k] = g.p[k];
  double* cur = lp;  // the current pressure
  double* nxt = lp + kLdsPoints;
  __syncthreads();
  for (int step = 0; step < g.nt; ++step) {
    // the indices are opaque to the compiler here, so it recomputes the
    // neighbours' addresses each step instead of holding every one of them
    // in registers across the loop
#pragma unroll
    for (int m = 0; m < kMaxPts; ++m) asm volatile("" : "+v"(pa[m]));
    asm volatile("" : "+v"(pfl), "+v"(cnt));
    copy_velocities(g, t, np);
    double bq[kMaxPts];
#pragma unroll
    for (int m0 = 0; m0 < kMaxPts; m0 += kChunkV) {
      double bc[kChunkV] = {};
      if (m0 < cnt)  // whole waves skip the chunks past their points
        chunk_eval(dv, bc, [&](int c, auto& div) { return source_at(g, dv, pa[m0 + c], div); });
#pragma unroll
      for (int c = 0; c < kChunkV; ++c) bq[m0 + c] = bc[c];
      asm volatile("" ::: "memory");  // keeps the next chunk's loads behind this one's
    }
    __syncthreads();  // un and vn complete
    for (int q = 0; q < g.nit; ++q) {
      // opaque here too, with the flags and the count: the lane masks the
      // compiler would otherwise hoist out of this loop overflow the scalar
      // registers
#pragma unroll
      for (int m = 0; m < kMaxPts; ++m) asm volatile("" : "+v"(pa[m]));
      asm volatile("" : "+v"(pfl), "+v"(cnt));
#pragma unroll
      for (int m0 = 0; m0 < kMaxPts; m0 += kMaxPts / 2) {
        if (m0 >= cnt) continue;
        double val[kMaxPts / 2];
        chunk_eval(dv, val, [&](int c, auto& div) {
          return pressure_at(g, dv, cur, pa[m0 + c], bq[m0 + c], div);
        });
#pragma unroll
        for (int c = 0; c < kMaxPts / 2; ++c)
          if (m0 + c < cnt) store_pressure(nxt, pa[m0 + c], (pfl >> (3 * (m0 + c))) & 7, nx, val[c]);
      }
      for (int k = t; k < nx; k += kThreadsC) nxt[(ny - 1) * nx + k] = 0.0;  // p[-1, :] = 0
      __syncthreads();
      double* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
#pragma unroll
    for (int m0 = 0; m0 < kMaxPts; m0 += kChunkV) {
      if (m0 >= cnt) continue;
      double nu[kChunkV], nv[kChunkV];
      chunk_eval(dv, nu, [&](int c, auto& div) {
        double uu;
        velocity_at(g, dv, cur, pa[m0 + c], uu, nv[c], div);
        return uu;
      });
#pragma unroll
      for (int c = 0; c < kChunkV; ++c) {
        if (m0 + c < cnt) {
          g.u[pa[m0 + c]] = nu[c];
          g.v[pa[m0 + c]] = nv[c];
        }
      }
    }
    velocity_boundaries(g, t);
    __syncthreads();
  }
  for (int k = t; k < np; k += kThreadsC) g.p[k] = cur[k];
}

}  // namespace
}  // namespace npbx

extern "C" int npbx_cavity_flow(int dtype, int64_t nx, int64_t ny, int64_t nt, int64_t nit,
                                void* u, void* v, double dt, double dx, double dy, void* p,
                                double rho, double nu, void* stream) {
  using namespace npbx;
  NPBX_REQUIRE(dtype == NPBX_F64, "cavity_flow: dtype %s not supported (f64)",
               dtype_name(dtype));
  NPBX_REQUIRE(nx >= 3 && ny >= 3 && nx * ny <= (int64_t(1) << 26),
               "cavity_flow: grid %lld x %lld out of range (3 x 3 .. 2^26 points)",
               static_cast<long long>(ny), static_cast<long long>(nx));
  NPBX_REQUIRE(nt >= 0 && nit >= 0 && nt <= (int64_t(1) << 30) && nit <= (int64_t(1) << 30),
               "cavity_flow: negative or huge step counts");
  if (nt == 0) return NPBX_OK;
  NPBX_REQUIRE(u && v && p && u != v && v != p && u != p, "cavity_flow: u, v, p must be distinct non-null arrays");
  const int64_t np = nx * ny;
  void* ws = nullptr;
  NPBX_TRY(workspace(static_cast<size_t>(4 * np) * sizeof(double), &ws));
  CavityArgs g;
  g.u = static_cast<double*>(u);
  g.v = static_cast<double*>(v);
  g.p = static_cast<double*>(p);
  g.un = static_cast<double*>(ws);
  g.vn = g.un + np;
  g.pn = g.vn + np;
  g.b = g.pn + np;
  g.nx = static_cast<int>(nx);
  g.ny = static_cast<int>(ny);
  g.nt = static_cast<int>(nt);
  g.nit = static_cast<int>(nit);
  g.rho = rho;
  g.nu = nu;
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
  // A time step depends on u, v and p only (b, un, vn are rebuilt each step),
  // so the steps split into launches of about 50 ms of work each: no launch
  // runs for long, whatever nt is.
  const int64_t per_launch = std::max<int64_t>(1, int64_t(150000000) / (np * (nit + 3)));
  for (int64_t done = 0; done < nt; done += per_launch) {
    g.nt = static_cast<int>(std::min<int64_t>(per_launch, nt - done));
    if (np <= kLdsPoints) {
      hipLaunchKernelGGL(cavity_flow_lds, dim3(1), dim3(kThreadsC), 0, as_stream(stream), g);
      NPBX_LAUNCHED("cavity_flow_lds");
    } else {
      hipLaunchKernelGGL(cavity_flow_global, dim3(1), dim3(kThreadsC), 0, as_stream(stream), g);
      NPBX_LAUNCHED("cavity_flow_global");
    }
  }
  return NPBX_OK;
}
