// HIP kernels for gfx950 (MI355X): the PIAFS hyperbolic/parabolic/source
// right-hand side and the explicit Runge-Kutta stage algebra, FP64 throughout.
//
// Arithmetic follows the reference's association order term by term (the
// file:line of the formula being restated is given at each device function);
// the compiler may contract a*b+c into FMA, which changes results only at the
// rounding level.
#include "kernels.hpp"

#include <cstdio>

namespace piafs {

namespace {

constexpr int TB = 256;  // threads per block for 1D-indexed kernels

inline unsigned nblk(long n, int tb = TB) { return (unsigned)((n + tb - 1) / tb); }

// ---------------------------------------------------------------------------
// WENO5 nonlinear weights, include/interpolation.h:255-490
// (JS :255-281, mapped :309-341, Z (Borges) :420-450, YC :478-508)
// ---------------------------------------------------------------------------
template <int KIND>
__device__ __forceinline__ void weno_w(double c1, double c2, double c3, double m3, double m2, double m1,
                                       double p1, double p2, double eps, double &w1, double &w2, double &w3) {
  const double t1312 = 13.0 / 12.0, q14 = 1.0 / 4.0;
  const double d1 = m3 - 2 * m2 + m1, e1 = m3 - 4 * m2 + 3 * m1;
  const double d2 = m2 - 2 * m1 + p1, e2 = m2 - p1;
  const double d3 = m1 - 2 * p1 + p2, e3 = 3 * m1 - 4 * p1 + p2;
  const double b1 = t1312 * d1 * d1 + q14 * e1 * e1;
  const double b2 = t1312 * d2 * d2 + q14 * e2 * e2;
  const double b3 = t1312 * d3 * d3 + q14 * e3 * e3;
  double a1, a2, a3;
  if (KIND == 2 || KIND == 3) {
    double tau;
    if (KIND == 2) {
      tau = fabs(b3 - b1);
    } else {
      const double t = m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2;
      tau = t * t;
    }
    const double r1 = tau / (b1 + eps), r2 = tau / (b2 + eps), r3 = tau / (b3 + eps);
    a1 = c1 * (1.0 + r1 * r1);
    a2 = c2 * (1.0 + r2 * r2);
    a3 = c3 * (1.0 + r3 * r3);
  } else {
    a1 = c1 / ((b1 + eps) * (b1 + eps));
    a2 = c2 / ((b2 + eps) * (b2 + eps));
    a3 = c3 / ((b3 + eps) * (b3 + eps));
  }
  double s = 1.0 / (a1 + a2 + a3);
  w1 = a1 * s;
  w2 = a2 * s;
  w3 = a3 * s;
  if (KIND == 1) {
    a1 = w1 * (c1 + c1 * c1 - 3 * c1 * w1 + w1 * w1) / (c1 * c1 + w1 * (1.0 - 2.0 * c1));
    a2 = w2 * (c2 + c2 * c2 - 3 * c2 * w2 + w2 * w2) / (c2 * c2 + w2 * (1.0 - 2.0 * c2));
    a3 = w3 * (c3 + c3 * c3 - 3 * c3 * w3 + w3 * w3) / (c3 * c3 + w3 * (1.0 - 2.0 * c3));
    s = 1.0 / (a1 + a2 + a3);
    w1 = a1 * s;
    w2 = a2 * s;
    w3 = a3 * s;
  }
}

__device__ __forceinline__ void weno_w_rt(int kind, double c1, double c2, double c3, double m3, double m2,
                                          double m1, double p1, double p2, double eps, double &w1, double &w2,
                                          double &w3) {
  switch (kind) {
    case 1: weno_w<1>(c1, c2, c3, m3, m2, m1, p1, p2, eps, w1, w2, w3); break;
    case 2: weno_w<2>(c1, c2, c3, m3, m2, m1, p1, p2, eps, w1, w2, w3); break;
    case 3: weno_w<3>(c1, c2, c3, m3, m2, m1, p1, p2, eps, w1, w2, w3); break;
    default: weno_w<0>(c1, c2, c3, m3, m2, m1, p1, p2, eps, w1, w2, w3); break;
  }
}

// Candidate stencils of WENO5 (Interp1PrimFifthOrderWENO.c:151-153) and of the
// compact scheme (Interp1PrimFifthOrderCRWENO.c:176-179).
__device__ __forceinline__ void weno5_cands(double m3, double m2, double m1, double p1, double p2, double &f1,
                                            double &f2, double &f3) {
  const double s6 = 1.0 / 6.0;
  f1 = (2 * s6) * m3 + (-7 * s6) * m2 + (11 * s6) * m1;
  f2 = (-s6) * m2 + (5 * s6) * m1 + (2 * s6) * p1;
  f3 = (2 * s6) * m1 + (5 * s6) * p1 + (-s6) * p2;
}
__device__ __forceinline__ void crweno5_cands(double m2, double m1, double p1, double &f1, double &f2,
                                              double &f3) {
  const double s6 = 1.0 / 6.0;
  f1 = (s6)*m2 + (5 * s6) * m1;
  f2 = (5 * s6) * m1 + (s6)*p1;
  f3 = (s6)*m1 + (5 * s6) * p1;
}

// ---------------------------------------------------------------------------
// Stage state + boundary conditions.
// TimeRK.c:46-57 (U = u + sum dt*A*Udot), then ApplyBoundaryConditions
// (BCPeriodic.c:22-62: ghost <- opposite interior; BCExtrapolate.c:16-51:
// mirror copy).  Each ghost point is evaluated at its boundary source point,
// which equals applying the BC to the stage state.  Points outside every
// boundary zone (the corners) are zero, as in the reference where they are
// never written.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bc_src(int i, int N, const FaceBC &lo, const FaceBC &hi, int other, bool &ok,
                                      int &other_shift) {
  other_shift = 0;
  if (i >= 0 && i < N) return i;
  const FaceBC &f = (i < 0) ? lo : hi;
  if (f.type < 0 || other < f.lo || other >= f.hi) {
    ok = false;
    return 0;
  }
  if (f.type == 0) {  // periodic: the reference indexes the source without the zone offset
    other_shift = -f.lo;
    return (i < 0) ? i + N : i - N;
  }
  return (i < 0) ? -i - 1 : 2 * N - i - 1;
}

template <int TILE>
__global__ void __launch_bounds__(TILE * 8) k_stage_bc2d(Grid G, BCMap bc, const double *__restrict__ u,
                                                         const double *__restrict__ d0, const double *__restrict__ d1,
                                                         const double *__restrict__ d2, double a0, double a1,
                                                         double a2, int nprev, double *__restrict__ U,
                                                         double *__restrict__ UT) {
  __shared__ double tile[TILE][TILE + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * TILE - G.g, j0 = blockIdx.y * TILE - G.gy;
  const int v = blockIdx.z;
  const long pl = (long)v * G.plane;
  for (int r = ty; r < TILE; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    double val = 0.0;
    if (i < G.nx + G.g && j < G.ny + G.gy) {
      bool ok = true;
      int sx, sy;
      int si = bc_src(i, G.nx, bc.f[0][0], bc.f[0][1], j, ok, sx);
      int sj = j;
      if (ok) {
        if (i >= 0 && i < G.nx) {
          sj = bc_src(j, G.ny, bc.f[1][0], bc.f[1][1], i, ok, sy);
          si = i + sy;
        } else {
          sj = j + sx;
        }
      }
      if (ok) {
        const long p = pl + (long)(sj + G.gy) * G.nxg + (si + G.g);
        val = u[p];
        if (nprev > 0) val += a0 * d0[p];
        if (nprev > 1) val += a1 * d1[p];
        if (nprev > 2) val += a2 * d2[p];
      }
      U[pl + (long)(j + G.gy) * G.nxg + (i + G.g)] = val;
    }
    tile[r][tx] = val;
  }
  if (!UT) return;
  __syncthreads();
  // transposed write: UT[v][i][j]
  for (int r = ty; r < TILE; r += 8) {
    const int i = i0 + r, j = j0 + tx;
    if (i < G.nx + G.g && j < G.ny + G.gy) UT[pl + (long)(i + G.g) * G.nyg + (j + G.gy)] = tile[tx][r];
  }
}

__global__ void k_stage_bc1d(Grid G, BCMap bc, const double *__restrict__ u, const double *__restrict__ d0,
                             const double *__restrict__ d1, const double *__restrict__ d2, double a0, double a1,
                             double a2, int nprev, double *__restrict__ U) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.nxg * G.nv) return;
  const int v = (int)(t / G.nxg), i = (int)(t % G.nxg) - G.g;
  bool ok = true;
  int sh;
  const int si = bc_src(i, G.nx, bc.f[0][0], bc.f[0][1], 0, ok, sh);
  double val = 0.0;
  if (ok) {
    const long p = (long)v * G.plane + si + G.g;
    val = u[p];
    if (nprev > 0) val += a0 * d0[p];
    if (nprev > 1) val += a1 * d1[p];
    if (nprev > 2) val += a2 * d2[p];
  }
  U[(long)v * G.plane + i + G.g] = val;
}

// ---------------------------------------------------------------------------
// Cell-centred flux at every position of every interior line.
// NavierStokes2DFlux.c:24-52 (navierstokes2d.h:70-110), Euler1DFlux.c:19
// (euler1d.h:61-85).
// ---------------------------------------------------------------------------
template <int MODEL, int DIR>
__global__ void __launch_bounds__(TB) k_flux(SweepGeom S, int nv, double gamma, const double *__restrict__ q,
                                             double *__restrict__ f) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long npos = S.N + 2 * S.g;
  if (t >= npos * S.nlines) return;
  const int l = (int)(t % S.nlines);
  const long k = t / S.nlines;
  const long b = k * S.pos_stride + S.line_off + l;
  if (MODEL == 1) {
    const double rho = q[b], v = q[S.plane + b] / rho, e = q[2 * S.plane + b];
    const double P = (e - 0.5 * rho * v * v) * (gamma - 1.0);
    f[b] = rho * v;
    f[S.plane + b] = rho * v * v + P;
    f[2 * S.plane + b] = (e + P) * v;
    for (int m = 3; m < nv; m++) f[m * S.plane + b] = v * q[m * S.plane + b];
  } else {
    const double rho = q[b];
    const double vx = (rho == 0) ? 0 : q[S.plane + b] / rho;
    const double vy = (rho == 0) ? 0 : q[2 * S.plane + b] / rho;
    const double e = q[3 * S.plane + b];
    const double vsq = (vx * vx) + (vy * vy);
    const double P = (e - 0.5 * rho * vsq) * (gamma - 1.0);
    const double vn = DIR == 0 ? vx : vy;
    f[b] = rho * vn;
    if (DIR == 0) {
      f[S.plane + b] = rho * vx * vx + P;
      f[2 * S.plane + b] = rho * vx * vy;
    } else {
      f[S.plane + b] = rho * vy * vx;
      f[2 * S.plane + b] = rho * vy * vy + P;
    }
    f[3 * S.plane + b] = (e + P) * vn;
    for (int m = 4; m < nv; m++) f[m * S.plane + b] = vn * q[m * S.plane + b];
  }
}

// ---------------------------------------------------------------------------
// Component-wise WENO5 / CRWENO5 reconstruction, one lane per
// (line, variable, reconstruction), sweeping the line with a six-value
// register window.  Reconstructions r: 0 uL, 1 uR (of u), 2 fL, 3 fR (of the
// flux); r even = left-biased.
//  * weights: WENOFifthOrderCalculateWeights{JS,M,Z,YC}
//    (WENOFifthOrderCalculateWeights.c:112-237 and the variants);
//    optimal weights :202-220 (WENO5 values at the two end interfaces of a
//    CRWENO line).
//  * WENO5: Interp1PrimFifthOrderWENO.c:124-166.
//  * CRWENO5: row assembly Interp1PrimFifthOrderCRWENO.c:132-201 (identity
//    rows with WENO5 right-hand sides at interfaces 0 and N), then the Thomas
//    algorithm of tridiagLU.c:133-150 (forward) and :243-255 (back
//    substitution).  The forward sweep keeps b' and c per row in scratch and
//    x' in the output array; the backward sweep overwrites it with x.
// ---------------------------------------------------------------------------
// Coefficients of row k of one line system: A, B, C and the right-hand side
// (for WENO5 the right-hand side is the interface value itself).
template <int CR, int KIND>
__device__ __forceinline__ void row_coeffs(int k, int N, bool left, const ReconParams &R, double w0, double w1,
                                           double w2, double w3, double w4, double w5, double &Ak, double &Bk,
                                           double &Ck, double &Rk) {
  double m3, m2, m1, p1, p2;
  if (left) { m3 = w0; m2 = w1; m1 = w2; p1 = w3; p2 = w4; }
  else { m3 = w5; m2 = w4; m1 = w3; p1 = w2; p2 = w1; }
  const bool bnd = (k == 0) || (k == N);
  double c1, c2, c3;
  if (CR && !bnd) { c1 = 0.2; c2 = 0.5; c3 = 0.3; }
  else { c1 = 0.1; c2 = 0.6; c3 = 0.3; }
  double a1, a2, a3;
  if (R.no_limiting) { a1 = c1; a2 = c2; a3 = c3; }
  else weno_w<KIND>(c1, c2, c3, m3, m2, m1, p1, p2, R.eps, a1, a2, a3);
  double f1, f2, f3;
  if (!CR || bnd) {
    weno5_cands(m3, m2, m1, p1, p2, f1, f2, f3);
    Ak = 0.0; Bk = 1.0; Ck = 0.0;
  } else {
    const double t3 = 1.0 / 3.0;
    crweno5_cands(m2, m1, p1, f1, f2, f3);
    const double lo = (2 * t3) * a1 + (t3)*a2;
    Bk = (t3)*a1 + (2 * t3) * a2 + (2 * t3) * a3;
    const double hi = (t3)*a3;
    if (left) { Ak = lo; Ck = hi; } else { Ak = hi; Ck = lo; }
  }
  Rk = a1 * f1 + a2 * f2 + a3 * f3;
}

// Rows are processed in blocks of BLK: the stencil values feeding the next
// block and, in the backward sweep, the next block's (b', c, x') are loaded
// one block ahead, so a memory latency is paid once per block and the
// block's independent weight evaluations overlap.
template <int CR, int KIND, int BLK>
__global__ void __launch_bounds__(TB) k_recon_comp(SweepGeom S, int nv, ReconParams R,
                                                   const double *__restrict__ q, const double *__restrict__ f,
                                                   double *__restrict__ uL, double *__restrict__ uR,
                                                   double *__restrict__ fL, double *__restrict__ fR,
                                                   double *__restrict__ scr_b, double *__restrict__ scr_c) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nl = S.nlines;
  if (t >= (long)nl * nv * 4) return;
  const int l = (int)(t % nl);
  const int vr = (int)(t / nl);
  const int v = vr % nv, r = vr / nv;
  const bool left = !(r & 1);
  const double *src = ((r < 2) ? q : f) + (long)v * S.plane + S.line_off + l;
  const long ps = S.pos_stride;
  const int N = S.N, g = S.g;
  const int last = N + g - 1;  // last position held by a line (ghosts included)
  double *out = ((r == 0) ? uL : (r == 1) ? uR : (r == 2) ? fL : fR) + (long)v * (N + 1) * nl + l;
  double *sb = nullptr, *sc = nullptr;
  if (CR) {
    const long so = (long)(r * nv + v) * (N + 1) * nl + l;
    sb = scr_b + so;
    sc = scr_c + so;
  }
  auto ld = [&](int pos) -> double { return (pos <= last) ? src[(long)(pos + g) * ps] : 0.0; };
  // window w0..w5 = positions k-3 .. k+2; cur[j] = position k0+3+j feeds row k0+j
  double w0 = ld(-3), w1 = ld(-2), w2 = ld(-1), w3 = ld(0), w4 = ld(1), w5 = ld(2);
  double cur[BLK];
#pragma unroll
  for (int j = 0; j < BLK; j++) cur[j] = ld(3 + j);
  double bprev = 1.0, cprev = 0.0, xprev = 0.0;
  for (int k0 = 0; k0 <= N; k0 += BLK) {
    double nxt[BLK];
#pragma unroll
    for (int j = 0; j < BLK; j++) nxt[j] = ld(k0 + BLK + 3 + j);
    double Ab[BLK], Bb[BLK], Cb[BLK], Rb[BLK];
#pragma unroll
    for (int j = 0; j < BLK; j++) {
      row_coeffs<CR, KIND>(k0 + j, N, left, R, w0, w1, w2, w3, w4, w5, Ab[j], Bb[j], Cb[j], Rb[j]);
      w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5; w5 = cur[j];
    }
#pragma unroll
    for (int j = 0; j < BLK; j++) {
      const int k = k0 + j;
      if (k > N) break;
      if (!CR) {
        out[(long)k * nl] = Rb[j];
      } else {
        double bk, xk;
        if (k == 0) { bk = Bb[j]; xk = Rb[j]; }
        else {
          const double fac = Ab[j] / bprev;
          bk = Bb[j] - fac * cprev;
          xk = Rb[j] - fac * xprev;
        }
        sb[(long)k * nl] = bk;
        sc[(long)k * nl] = Cb[j];
        out[(long)k * nl] = xk;
        bprev = bk; cprev = Cb[j]; xprev = xk;
      }
    }
#pragma unroll
    for (int j = 0; j < BLK; j++) cur[j] = nxt[j];
  }
  if (CR) {
    double xn = xprev / bprev;  // row N: c_N = 0
    out[(long)N * nl] = xn;
    double bb[BLK], cc[BLK], xx[BLK];
#pragma unroll
    for (int j = 0; j < BLK; j++) {
      const int k = N - 1 - j;
      bb[j] = k >= 0 ? sb[(long)k * nl] : 1.0;
      cc[j] = k >= 0 ? sc[(long)k * nl] : 0.0;
      xx[j] = k >= 0 ? out[(long)k * nl] : 0.0;
    }
    for (int k0 = N - 1; k0 >= 0; k0 -= BLK) {
      double nb[BLK], nc[BLK], nx[BLK];
#pragma unroll
      for (int j = 0; j < BLK; j++) {
        const int k = k0 - BLK - j;
        nb[j] = k >= 0 ? sb[(long)k * nl] : 1.0;
        nc[j] = k >= 0 ? sc[(long)k * nl] : 0.0;
        nx[j] = k >= 0 ? out[(long)k * nl] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < BLK; j++) {
        const int k = k0 - j;
        if (k < 0) break;
        const double xk = (xx[j] - cc[j] * xn) / bb[j];
        out[(long)k * nl] = xk;
        xn = xk;
      }
#pragma unroll
      for (int j = 0; j < BLK; j++) { bb[j] = nb[j]; cc[j] = nc[j]; xx[j] = nx[j]; }
    }
  }
}

// CRWENO5 reconstruction as a producer/consumer workgroup: 64 systems (64
// consecutive lines of one variable and reconstruction) per workgroup of four
// waves.  Waves 1-3 evaluate the row coefficients (weights, A, B, C, R) of a
// chunk of KR rows into one half of a double-buffered LDS ring while wave 0
// runs the Thomas forward recurrence (tridiagLU.c:133-150) over the previous
// chunk; wave 0 then back-substitutes (:243-255) with a one-block-ahead
// prefetch.  Per system the arithmetic and its order are those of
// k_recon_comp<1, KIND>.
template <int KIND, int KR>
__global__ void __launch_bounds__(256) k_crweno_pc(SweepGeom S, int nv, ReconParams R,
                                                   const double *__restrict__ q, const double *__restrict__ f,
                                                   double *__restrict__ uL, double *__restrict__ uR,
                                                   double *__restrict__ fL, double *__restrict__ fR,
                                                   double *__restrict__ scr_b, double *__restrict__ scr_c) {
  __shared__ double buf[2][KR][4][64];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int nl = S.nlines, N = S.N, g = S.g;
  const int l0 = blockIdx.x * 64;
  const int v = blockIdx.y % nv, r = blockIdx.y / nv;
  const bool left = !(r & 1);
  const double *src = ((r < 2) ? q : f) + (long)v * S.plane + S.line_off + l0;
  const long ps = S.pos_stride;
  const int nchunks = (N + 1 + KR - 1) / KR;
  const int my_l = l0 + lane;
  const bool lane_ok = my_l < nl;
  double *out = ((r == 0) ? uL : (r == 1) ? uR : (r == 2) ? fL : fR) + (long)v * (N + 1) * nl + my_l;
  const long so = (long)(r * nv + v) * (N + 1) * nl + my_l;
  double *sb = scr_b + so, *sc = scr_c + so;
  double bprev = 1.0, cprev = 0.0, xprev = 0.0;
  for (int c = 0; c <= nchunks; c++) {
    if (wave > 0 && c < nchunks) {
      const int half = c & 1;
      for (int it = tid - 64; it < KR * 64; it += 192) {
        const int kr = it >> 6, ln = it & 63;
        const int k = c * KR + kr;
        if (k <= N && l0 + ln < nl) {
          const double *sp = src + ln;
          const double w0 = sp[(long)(k - 3 + g) * ps], w1 = sp[(long)(k - 2 + g) * ps];
          const double w2 = sp[(long)(k - 1 + g) * ps], w3 = sp[(long)(k + g) * ps];
          const double w4 = sp[(long)(k + 1 + g) * ps], w5 = sp[(long)(k + 2 + g) * ps];
          double A, B, C, Rk;
          row_coeffs<1, KIND>(k, N, left, R, w0, w1, w2, w3, w4, w5, A, B, C, Rk);
          buf[half][kr][0][ln] = A;
          buf[half][kr][1][ln] = B;
          buf[half][kr][2][ln] = C;
          buf[half][kr][3][ln] = Rk;
        }
      }
    }
    if (wave == 0 && c > 0 && lane_ok) {
      const int cc = c - 1, half = cc & 1;
#pragma unroll
      for (int kr = 0; kr < KR; kr++) {
        const int k = cc * KR + kr;
        if (k > N) break;
        const double A = buf[half][kr][0][lane], B = buf[half][kr][1][lane];
        const double C = buf[half][kr][2][lane], Rk = buf[half][kr][3][lane];
        double bk, xk;
        if (k == 0) { bk = B; xk = Rk; }
        else {
          const double fac = A / bprev;
          bk = B - fac * cprev;
          xk = Rk - fac * xprev;
        }
        sb[(long)k * nl] = bk;
        sc[(long)k * nl] = C;
        out[(long)k * nl] = xk;
        bprev = bk; cprev = C; xprev = xk;
      }
    }
    __syncthreads();
  }
  if (wave != 0 || !lane_ok) return;
  constexpr int BLK = 8;
  double xn = xprev / bprev;  // row N: c_N = 0
  out[(long)N * nl] = xn;
  double bb[BLK], cc[BLK], xx[BLK];
#pragma unroll
  for (int j = 0; j < BLK; j++) {
    const int k = N - 1 - j;
    bb[j] = k >= 0 ? sb[(long)k * nl] : 1.0;
    cc[j] = k >= 0 ? sc[(long)k * nl] : 0.0;
    xx[j] = k >= 0 ? out[(long)k * nl] : 0.0;
  }
  for (int k0 = N - 1; k0 >= 0; k0 -= BLK) {
    double nb[BLK], nc[BLK], nx[BLK];
#pragma unroll
    for (int j = 0; j < BLK; j++) {
      const int k = k0 - BLK - j;
      nb[j] = k >= 0 ? sb[(long)k * nl] : 1.0;
      nc[j] = k >= 0 ? sc[(long)k * nl] : 0.0;
      nx[j] = k >= 0 ? out[(long)k * nl] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < BLK; j++) {
      const int k = k0 - j;
      if (k < 0) break;
      const double xk = (xx[j] - cc[j] * xn) / bb[j];
      out[(long)k * nl] = xk;
      xn = xk;
    }
#pragma unroll
    for (int j = 0; j < BLK; j++) { bb[j] = nb[j]; cc[j] = nc[j]; xx[j] = nx[j]; }
  }
}

// ---------------------------------------------------------------------------
// 1D Euler eigenstructure, euler1d.h:91-193
// ---------------------------------------------------------------------------
struct Eig3 {
  double L[3][3], R[3][3], D[3];
};

__device__ __forceinline__ void euler1d_roeavg(double g, double rhoL, double muL, double eL, double rhoR,
                                               double muR, double eR, double &r0, double &r1, double &r2) {
  const double vL = muL / rhoL, PL = (eL - 0.5 * rhoL * vL * vL) * (g - 1.0), cLsq = g * PL / rhoL;
  const double HL = 0.5 * vL * vL + cLsq / (g - 1.0);
  const double vR = muR / rhoR, PR = (eR - 0.5 * rhoR * vR * vR) * (g - 1.0), cRsq = g * PR / rhoR;
  const double HR = 0.5 * vR * vR + cRsq / (g - 1.0);
  const double tL = sqrt(rhoL), tR = sqrt(rhoR);
  const double rho = tL * tR, v = (tL * vL + tR * vR) / (tL + tR), H = (tL * HL + tR * HR) / (tL + tR);
  const double csq = (g - 1.0) * (H - 0.5 * v * v), P = csq * rho / g;
  r0 = rho;
  r1 = rho * v;
  r2 = P / (g - 1.0) + 0.5 * rho * v * v;
}

__device__ __forceinline__ void euler1d_eig(double g, double u0, double u1, double u2, Eig3 &E) {
  const double rho = u0, v = u1 / rho, e = u2;
  const double P = (e - 0.5 * rho * v * v) * (g - 1.0), c = sqrt(g * P / rho);
  E.D[0] = v; E.D[1] = v - c; E.D[2] = v + c;
  const double k = (g - 1) / (rho * c);
  E.L[1][0] = k * (-(v * v) / 2 - c * v / (g - 1));
  E.L[1][1] = k * (v + c / (g - 1));
  E.L[1][2] = k * (-1);
  E.L[0][0] = k * (rho * (-(v * v) / 2 + c * c / (g - 1)) / c);
  E.L[0][1] = k * (rho * v / c);
  E.L[0][2] = k * (-rho / c);
  E.L[2][0] = k * ((v * v) / 2 - c * v / (g - 1));
  E.L[2][1] = k * (-v + c / (g - 1));
  E.L[2][2] = k * (1);
  E.R[0][1] = -rho / (2 * c);
  E.R[1][1] = -rho * (v - c) / (2 * c);
  E.R[2][1] = -rho * ((v * v) / 2 + (c * c) / (g - 1) - c * v) / (2 * c);
  E.R[0][0] = 1; E.R[1][0] = v; E.R[2][0] = v * v / 2;
  E.R[0][2] = rho / (2 * c);
  E.R[1][2] = rho * (v + c) / (2 * c);
  E.R[2][2] = rho * ((v * v) / 2 + (c * c) / (g - 1) + c * v) / (2 * c);
}

// Characteristic WENO5 for euler1d, one thread per (interface, line).
// WENOFifthOrderCalculateWeightsChar* (WENOFifthOrderCalculateWeights.c:761-930)
// and Interp1PrimFifthOrderWENOChar.c:127-195: L and R from the Roe average of
// the two cells adjacent to the interface; species rows are the identity, so
// species reconstruct component-wise.
template <int KIND>
__global__ void __launch_bounds__(TB) k_recon_char_e1d(SweepGeom S, int nv, ReconParams R, double gam,
                                                       const double *__restrict__ q, const double *__restrict__ f,
                                                       double *__restrict__ uL, double *__restrict__ uR,
                                                       double *__restrict__ fL, double *__restrict__ fR) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nl = S.nlines, N = S.N;
  if (t >= (long)(N + 1) * nl) return;
  const int l = (int)(t % nl), k = (int)(t / nl);
  const long ps = S.pos_stride, pl = S.plane;
  const long base = S.line_off + l;
  auto at = [&](const double *a, int v, int pos) { return a[v * pl + (long)(pos + S.g) * ps + base]; };
  double a0, a1, a2;
  euler1d_roeavg(gam, at(q, 0, k - 1), at(q, 1, k - 1), at(q, 2, k - 1), at(q, 0, k), at(q, 1, k), at(q, 2, k),
                 a0, a1, a2);
  Eig3 E;
  euler1d_eig(gam, a0, a1, a2, E);
  const double s6 = 1.0 / 6.0;
  for (int r = 0; r < 4; r++) {
    const double *src = (r < 2) ? q : f;
    const bool left = !(r & 1);
    double *out = (r == 0) ? uL : (r == 1) ? uR : (r == 2) ? fL : fR;
    int pos[5];
    if (left) { pos[0] = k - 3; pos[1] = k - 2; pos[2] = k - 1; pos[3] = k; pos[4] = k + 1; }
    else { pos[0] = k + 2; pos[1] = k + 1; pos[2] = k; pos[3] = k - 1; pos[4] = k - 2; }
    double s[5][3];
    for (int p = 0; p < 5; p++) {
      const double x0 = at(src, 0, pos[p]), x1 = at(src, 1, pos[p]), x2 = at(src, 2, pos[p]);
      for (int vv = 0; vv < 3; vv++) s[p][vv] = E.L[vv][0] * x0 + E.L[vv][1] * x1 + E.L[vv][2] * x2;
    }
    double fch[3];
    for (int vv = 0; vv < 3; vv++) {
      double w1, w2, w3;
      if (R.no_limiting) { w1 = 0.1; w2 = 0.6; w3 = 0.3; }
      else weno_w<KIND>(0.1, 0.6, 0.3, s[0][vv], s[1][vv], s[2][vv], s[3][vv], s[4][vv], R.eps, w1, w2, w3);
      const double f1 = (2 * s6) * s[0][vv] - (7.0 * s6) * s[1][vv] + (11.0 * s6) * s[2][vv];
      const double f2 = (-s6) * s[1][vv] + (5.0 * s6) * s[2][vv] + (2 * s6) * s[3][vv];
      const double f3 = (2 * s6) * s[2][vv] + (5 * s6) * s[3][vv] - (s6)*s[4][vv];
      fch[vv] = w1 * f1 + w2 * f2 + w3 * f3;
    }
    for (int vv = 0; vv < 3; vv++)
      out[((long)vv * (N + 1) + k) * nl + l] = E.R[vv][0] * fch[0] + E.R[vv][1] * fch[1] + E.R[vv][2] * fch[2];
    for (int m = 3; m < nv; m++) {
      const double m3 = at(src, m, pos[0]), m2 = at(src, m, pos[1]), m1 = at(src, m, pos[2]);
      const double p1 = at(src, m, pos[3]), p2 = at(src, m, pos[4]);
      double w1, w2, w3, f1, f2, f3;
      if (R.no_limiting) { w1 = 0.1; w2 = 0.6; w3 = 0.3; }
      else weno_w<KIND>(0.1, 0.6, 0.3, m3, m2, m1, p1, p2, R.eps, w1, w2, w3);
      weno5_cands(m3, m2, m1, p1, p2, f1, f2, f3);
      out[((long)m * (N + 1) + k) * nl + l] = w1 * f1 + w2 * f2 + w3 * f3;
    }
  }
}

// ---------------------------------------------------------------------------
// Upwinding.
//  NS2D Roe: NavierStokes2DUpwind.c:26-92 with navierstokes2d.h:115-311 --
//   |A| = R |D| L built as the 4x4 flow block (the species block of the
//   reference's dense n x n product is diagonal: R = L = I, D = vn), Harten
//   entropy fix delta = 1e-6.
//  Euler1D Roe: Euler1DUpwind.c:27-93 (no entropy fix); rf-char :107-190.
// ---------------------------------------------------------------------------
template <int DIR>
__global__ void __launch_bounds__(TB) k_upwind_ns2d_roe(SweepGeom S, int nv, double g, const double *__restrict__ q,
                                                        const double *__restrict__ uL, const double *__restrict__ uR,
                                                        const double *__restrict__ fL, const double *__restrict__ fR,
                                                        double *__restrict__ fI) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nl = S.nlines, N = S.N;
  if (t >= (long)(N + 1) * nl) return;
  const int l = (int)(t % nl), k = (int)(t / nl);
  const long ps = S.pos_stride, pl = S.plane;
  const long cL = (long)(k - 1 + S.g) * ps + S.line_off + l, cR = cL + ps;
  // Roe average of the cell-centred neighbours (navierstokes2d.h:115-160)
  const double rhoL = q[cL], vxL = q[pl + cL] / rhoL, vyL = q[2 * pl + cL] / rhoL, eL = q[3 * pl + cL];
  const double vsqL = (vxL * vxL) + (vyL * vyL), PL = (eL - 0.5 * rhoL * vsqL) * (g - 1.0);
  const double cLsq = g * PL / rhoL, HL = 0.5 * (vxL * vxL + vyL * vyL) + cLsq / (g - 1.0);
  const double rhoR = q[cR], vxR = q[pl + cR] / rhoR, vyR = q[2 * pl + cR] / rhoR, eR = q[3 * pl + cR];
  const double vsqR = (vxR * vxR) + (vyR * vyR), PR = (eR - 0.5 * rhoR * vsqR) * (g - 1.0);
  const double cRsq = g * PR / rhoR, HR = 0.5 * (vxR * vxR + vyR * vyR) + cRsq / (g - 1.0);
  const double tL = sqrt(rhoL), tR = sqrt(rhoR);
  const double rA = tL * tR;
  const double vxA = (tL * vxL + tR * vxR) / (tL + tR), vyA = (tL * vyL + tR * vyR) / (tL + tR);
  const double HA = (tL * HL + tR * HR) / (tL + tR), vsqA = vxA * vxA + vyA * vyA;
  const double csqA = (g - 1.0) * (HA - 0.5 * vsqA), PA = csqA * rA / g;
  const double eA = PA / (g - 1.0) + 0.5 * rA * vsqA;
  const double ua0 = rA, ua1 = rA * vxA, ua2 = rA * vyA, ua3 = eA;
  // eigenstructure of the averaged state (navierstokes2d.h:166-311)
  const double gm1 = g - 1.0;
  const double rho = ua0, vx = (rho == 0) ? 0 : ua1 / rho, vy = (rho == 0) ? 0 : ua2 / rho, e = ua3;
  const double vsq = (vx * vx) + (vy * vy), P = (e - 0.5 * rho * vsq) * (g - 1.0);
  const double a = sqrt(g * P / rho), ek = 0.5 * (vx * vx + vy * vy);
  const double vn = DIR == 0 ? vx : vy;
  double D[4];
  D[0] = vn - a;
  if (DIR == 0) { D[1] = vn + a; D[2] = vn; }
  else { D[1] = vn; D[2] = vn + a; }
  D[3] = vn;
  const double delta = 0.000001, delta2 = delta * delta;
  double Dsp = (fabs(vn) < delta ? (vn * vn + delta2) / (2 * delta) : fabs(vn));
#pragma unroll
  for (int m = 0; m < 4; m++) D[m] = (fabs(D[m]) < delta ? (D[m] * D[m] + delta2) / (2 * delta) : fabs(D[m]));
  double L[4][4], R[4][4];
  const double aa2 = 2 * a * a, aa = a * a;
  if (DIR == 0) {
    const double un = vx, nx = 1.0, ny = 0.0;
    L[0][0] = (gm1 * ek + a * un) / aa2;
    L[0][1] = ((-gm1) * vx - a * nx) / aa2;
    L[0][2] = ((-gm1) * vy - a * ny) / aa2;
    L[0][3] = gm1 / aa2;
    L[3][0] = (aa - gm1 * ek) / aa;
    L[3][1] = (gm1 * vx) / aa;
    L[3][2] = (gm1 * vy) / aa;
    L[3][3] = (-gm1) / aa;
    L[1][0] = (gm1 * ek - a * un) / aa2;
    L[1][1] = ((-gm1) * vx + a * nx) / aa2;
    L[1][2] = ((-gm1) * vy + a * ny) / aa2;
    L[1][3] = gm1 / aa2;
    L[2][0] = (vy - un * ny) / nx;
    L[2][1] = ny;
    L[2][2] = (ny * ny - 1.0) / nx;
    L[2][3] = 0.0;
  } else {
    const double un = vy, nx = 0.0, ny = 1.0;
    L[0][0] = (gm1 * ek + a * un) / aa2;
    L[0][1] = ((1.0 - g) * vx - a * nx) / aa2;
    L[0][2] = ((1.0 - g) * vy - a * ny) / aa2;
    L[0][3] = gm1 / aa2;
    L[3][0] = (aa - gm1 * ek) / aa;
    L[3][1] = gm1 * vx / aa;
    L[3][2] = gm1 * vy / aa;
    L[3][3] = (1.0 - g) / aa;
    L[2][0] = (gm1 * ek - a * un) / aa2;
    L[2][1] = ((1.0 - g) * vx + a * nx) / aa2;
    L[2][2] = ((1.0 - g) * vy + a * ny) / aa2;
    L[2][3] = gm1 / aa2;
    L[1][0] = (un * nx - vx) / ny;
    L[1][1] = (1.0 - nx * nx) / ny;
    L[1][2] = -nx;
    L[1][3] = 0;
  }
  {
    const double h0 = a * a / gm1 + ek;
    const double nx = DIR == 0 ? 1.0 : 0.0, ny = DIR == 0 ? 0.0 : 1.0, un = vn;
    R[0][0] = 1.0; R[1][0] = vx - a * nx; R[2][0] = vy - a * ny; R[3][0] = h0 - a * un;
    R[0][3] = 1.0; R[1][3] = vx; R[2][3] = vy; R[3][3] = ek;
    const int cp = DIR == 0 ? 1 : 2, cz = DIR == 0 ? 2 : 1;
    R[0][cp] = 1.0; R[1][cp] = vx + a * nx; R[2][cp] = vy + a * ny; R[3][cp] = h0 + a * un;
    R[0][cz] = 0.0; R[1][cz] = ny; R[2][cz] = -nx; R[3][cz] = vx * ny - vy * nx;
  }
  // interface data, flow block
  const long ii = (long)k * nl + l, vstride = (long)(N + 1) * nl;
  double ud[4], fo[4];
#pragma unroll
  for (int m = 0; m < 4; m++) ud[m] = 0.5 * (uR[m * vstride + ii] - uL[m * vstride + ii]);
  // modA = R (D L), row-major sums in index order (matmult_native.h:21-34)
  double DL[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) DL[i][j] = D[i] * L[i][j];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double us = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double mA = R[i][0] * DL[0][j] + R[i][1] * DL[1][j] + R[i][2] * DL[2][j] + R[i][3] * DL[3][j];
      us += mA * ud[j];
    }
    fo[i] = 0.5 * (fL[i * vstride + ii] + fR[i * vstride + ii]) - us;
  }
#pragma unroll
  for (int m = 0; m < 4; m++) fI[m * vstride + ii] = fo[m];
  for (int m = 4; m < nv; m++) {
    const double udm = 0.5 * (uR[m * vstride + ii] - uL[m * vstride + ii]);
    fI[m * vstride + ii] = 0.5 * (fL[m * vstride + ii] + fR[m * vstride + ii]) - Dsp * udm;
  }
}

template <int RF>
__global__ void __launch_bounds__(TB) k_upwind_e1d(SweepGeom S, int nv, double g, const double *__restrict__ q,
                                                   const double *__restrict__ uL, const double *__restrict__ uR,
                                                   const double *__restrict__ fL, const double *__restrict__ fR,
                                                   double *__restrict__ fI) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nl = S.nlines, N = S.N;
  if (t >= (long)(N + 1) * nl) return;
  const int l = (int)(t % nl), k = (int)(t / nl);
  const long ps = S.pos_stride, pl = S.plane;
  const long cL = (long)(k - 1 + S.g) * ps + S.line_off + l, cR = cL + ps;
  double a0, a1, a2;
  euler1d_roeavg(g, q[cL], q[pl + cL], q[2 * pl + cL], q[cR], q[pl + cR], q[2 * pl + cR], a0, a1, a2);
  Eig3 E;
  euler1d_eig(g, a0, a1, a2, E);
  const long ii = (long)k * nl + l, vs = (long)(N + 1) * nl;
  if (!RF) {
    double ud[3], DL[3][3];
    for (int m = 0; m < 3; m++) ud[m] = 0.5 * (uR[m * vs + ii] - uL[m * vs + ii]);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) DL[i][j] = fabs(E.D[i]) * E.L[i][j];
    for (int i = 0; i < 3; i++) {
      double us = 0.0;
      for (int j = 0; j < 3; j++) {
        const double mA = E.R[i][0] * DL[0][j] + E.R[i][1] * DL[1][j] + E.R[i][2] * DL[2][j];
        us += mA * ud[j];
      }
      fI[i * vs + ii] = 0.5 * (fL[i * vs + ii] + fR[i * vs + ii]) - us;
    }
    const double Dsp = fabs(E.D[0]);
    for (int m = 3; m < nv; m++) {
      const double udm = 0.5 * (uR[m * vs + ii] - uL[m * vs + ii]);
      fI[m * vs + ii] = 0.5 * (fL[m * vs + ii] + fR[m * vs + ii]) - Dsp * udm;
    }
  } else {
    // cell eigenvalues (euler1d.h:131-147) of the two neighbours
    Eig3 EL, ER;
    euler1d_eig(g, q[cL], q[pl + cL], q[2 * pl + cL], EL);
    euler1d_eig(g, q[cR], q[pl + cR], q[2 * pl + cR], ER);
    double fc[3];
    for (int m = 0; m < 3; m++) {
      double ucL = 0, ucR = 0, fcL = 0, fcR = 0;
      for (int j = 0; j < 3; j++) {
        ucL += E.L[m][j] * uL[j * vs + ii];
        ucR += E.L[m][j] * uR[j * vs + ii];
        fcL += E.L[m][j] * fL[j * vs + ii];
        fcR += E.L[m][j] * fR[j * vs + ii];
      }
      const double eL = EL.D[m], eC = E.D[m], eR = ER.D[m];
      if (eL > 0 && eC > 0 && eR > 0) fc[m] = fcL;
      else if (eL < 0 && eC < 0 && eR < 0) fc[m] = fcR;
      else {
        const double x1 = fabs(eL), x2 = fabs(eC), x3 = fabs(eR);
        const double mx1 = x1 > x2 ? x1 : x2, mx2 = x2 > x3 ? x2 : x3;
        const double alpha = mx1 > mx2 ? mx1 : mx2;
        fc[m] = 0.5 * (fcL + fcR + alpha * (ucL - ucR));
      }
    }
    for (int i = 0; i < 3; i++) fI[i * vs + ii] = E.R[i][0] * fc[0] + E.R[i][1] * fc[1] + E.R[i][2] * fc[2];
    // species rows: identity eigenvectors, eigenvalue v
    const double eLs = EL.D[0], eCs = E.D[0], eRs = ER.D[0];
    for (int m = 3; m < nv; m++) {
      const double ucL = uL[m * vs + ii], ucR = uR[m * vs + ii], fcL = fL[m * vs + ii], fcR = fR[m * vs + ii];
      double r;
      if (eLs > 0 && eCs > 0 && eRs > 0) r = fcL;
      else if (eLs < 0 && eCs < 0 && eRs < 0) r = fcR;
      else {
        const double x1 = fabs(eLs), x2 = fabs(eCs), x3 = fabs(eRs);
        const double mx1 = x1 > x2 ? x1 : x2, mx2 = x2 > x3 ? x2 : x3;
        r = 0.5 * (fcL + fcR + (mx1 > mx2 ? mx1 : mx2) * (ucL - ucR));
      }
      fI[m * vs + ii] = r;
    }
  }
}

// hyp = dxinv * (fI[k+1] - fI[k]): HyperbolicFunction.c:97-112.  Output is the
// compact interior array h[(v*N + k)*nlines + l].
__global__ void __launch_bounds__(TB) k_divergence(SweepGeom S, int nv, const double *__restrict__ dxinv,
                                                   const double *__restrict__ fI, double *__restrict__ h) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nl = S.nlines, N = S.N;
  if (t >= (long)N * nl) return;
  const int l = (int)(t % nl), k = (int)(t / nl);
  const double di = dxinv[S.g + k];
  for (int v = 0; v < nv; v++) {
    const long a = ((long)v * (N + 1) + k) * nl + l;
    h[((long)v * N + k) * nl + l] = di * (fI[a + nl] - fI[a]);
  }
}

// [v][rows][cols] -> [v][cols][rows]
__global__ void __launch_bounds__(256) k_transpose(int rows, int cols, const double *__restrict__ src,
                                                   double *__restrict__ dst) {
  __shared__ double tile[32][33];
  const int v = blockIdx.z;
  const long off = (long)v * rows * cols;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int y = threadIdx.y; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + threadIdx.x;
    if (r < rows && c < cols) tile[y][threadIdx.x] = src[off + (long)r * cols + c];
  }
  __syncthreads();
  for (int y = threadIdx.y; y < 32; y += 8) {
    const int c = c0 + y, r = r0 + threadIdx.x;
    if (r < rows && c < cols) dst[off + (long)c * rows + r] = tile[threadIdx.x][y];
  }
}

// ---------------------------------------------------------------------------
// Parabolic term (navierstokes2d): NavierStokes2DParabolicFunction.c:41-204.
// Serial-run semantics: d/dx of Q is evaluated on interior rows only and d/dy
// on interior columns only (FirstDerivativeFourthOrder.c:66-71 loops over the
// interior lines of the other dimension); elsewhere the derivative is zero.
// ---------------------------------------------------------------------------
struct Prim {
  double vx, vy, T;
};
__device__ __forceinline__ Prim prim_at(const Grid &G, double gam, const double *U, long p) {
  const double rho = U[p];
  const double vx = (rho == 0) ? 0 : U[G.plane + p] / rho, vy = (rho == 0) ? 0 : U[2 * G.plane + p] / rho;
  const double e = U[3 * G.plane + p];
  const double vsq = (vx * vx) + (vy * vy);
  const double P = (e - 0.5 * rho * vsq) * (gam - 1.0);
  Prim r;
  r.vx = vx;
  r.vy = vy;
  r.T = gam * P / rho;
  return r;
}

// FirstDerivativeFourthOrder.c:39-132 (PAR4) / FirstDerivativeFirstOrder.c (bias +1)
// of the three primitives at position i of a line of N interior points.
template <int PAR4>
__device__ __forceinline__ void deriv_prim(const Grid &G, double gam, const double *U, long p0, long st, int i,
                                           int N, double out[3]) {
  const int g = G.g;
  auto P = [&](int k) { return prim_at(G, gam, U, p0 + (long)k * st); };
  if (PAR4) {
    const double t12 = 1.0 / 12.0;
    if (i == -g) {
      Prim a = P(i), b = P(i + 1), c = P(i + 2), d = P(i + 3), e = P(i + 4);
      out[0] = (-25 * a.vx + 48 * b.vx - 36 * c.vx + 16 * d.vx - 3 * e.vx) * t12;
      out[1] = (-25 * a.vy + 48 * b.vy - 36 * c.vy + 16 * d.vy - 3 * e.vy) * t12;
      out[2] = (-25 * a.T + 48 * b.T - 36 * c.T + 16 * d.T - 3 * e.T) * t12;
    } else if (i == -g + 1) {
      Prim a = P(i - 1), b = P(i), c = P(i + 1), d = P(i + 2), e = P(i + 3);
      out[0] = (-3 * a.vx - 10 * b.vx + 18 * c.vx - 6 * d.vx + e.vx) * t12;
      out[1] = (-3 * a.vy - 10 * b.vy + 18 * c.vy - 6 * d.vy + e.vy) * t12;
      out[2] = (-3 * a.T - 10 * b.T + 18 * c.T - 6 * d.T + e.T) * t12;
    } else if (i == N + g - 2) {
      Prim a = P(i - 3), b = P(i - 2), c = P(i - 1), d = P(i), e = P(i + 1);
      out[0] = (-a.vx + 6 * b.vx - 18 * c.vx + 10 * d.vx + 3 * e.vx) * t12;
      out[1] = (-a.vy + 6 * b.vy - 18 * c.vy + 10 * d.vy + 3 * e.vy) * t12;
      out[2] = (-a.T + 6 * b.T - 18 * c.T + 10 * d.T + 3 * e.T) * t12;
    } else if (i == N + g - 1) {
      Prim a = P(i - 4), b = P(i - 3), c = P(i - 2), d = P(i - 1), e = P(i);
      out[0] = (3 * a.vx - 16 * b.vx + 36 * c.vx - 48 * d.vx + 25 * e.vx) * t12;
      out[1] = (3 * a.vy - 16 * b.vy + 36 * c.vy - 48 * d.vy + 25 * e.vy) * t12;
      out[2] = (3 * a.T - 16 * b.T + 36 * c.T - 48 * d.T + 25 * e.T) * t12;
    } else {
      Prim a = P(i - 2), b = P(i - 1), d = P(i + 1), e = P(i + 2);
      out[0] = (a.vx - 8 * b.vx + 8 * d.vx - e.vx) * t12;
      out[1] = (a.vy - 8 * b.vy + 8 * d.vy - e.vy) * t12;
      out[2] = (a.T - 8 * b.T + 8 * d.T - e.T) * t12;
    }
  } else {
    if (i == -g) {
      Prim a = P(i), b = P(i + 1);
      out[0] = b.vx - a.vx; out[1] = b.vy - a.vy; out[2] = b.T - a.T;
    } else if (i == N + g - 1) {
      Prim a = P(i - 1), b = P(i);
      out[0] = b.vx - a.vx; out[1] = b.vy - a.vy; out[2] = b.T - a.T;
    } else {
      Prim a = P(i - 1), b = P(i), c = P(i + 1);
      out[0] = 1.0 * c.vx - 1.0 * b.vx + 0.0 * a.vx;
      out[1] = 1.0 * c.vy - 1.0 * b.vy + 0.0 * a.vy;
      out[2] = 1.0 * c.T - 1.0 * b.T + 0.0 * a.T;
    }
  }
}

// Qd planes: 0 ux, 1 vx, 2 Tx (d/dx), 3 uy, 4 vy, 5 Ty (d/dy), each scaled by
// dxinv (NavierStokes2DParabolicFunction.c:96-106).
template <int PAR4>
__global__ void __launch_bounds__(TB) k_qderiv(Grid G, double gam, const double *__restrict__ U,
                                               const double *__restrict__ dxinv, double *__restrict__ Qd) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.nxg * G.nyg) return;
  const int i = (int)(t % G.nxg) - G.g, j = (int)(t / G.nxg) - G.g;
  const long p = t;
  const long pl = (long)G.nxg * G.nyg;
  double dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0};
  if (j >= 0 && j < G.ny) {
    const long p0 = (long)(j + G.g) * G.nxg + G.g;  // (0, j)
    deriv_prim<PAR4>(G, gam, U, p0, 1, i, G.nx, dx);
    const double s = dxinv[G.g + i];
    dx[0] *= s; dx[1] *= s; dx[2] *= s;
  }
  if (i >= 0 && i < G.nx) {
    const long p0 = (long)G.g * G.nxg + (i + G.g);  // (i, 0)
    deriv_prim<PAR4>(G, gam, U, p0, G.nxg, j, G.ny, dy);
    const double s = dxinv[(G.nx + 2 * G.g) + G.g + j];
    dy[0] *= s; dy[1] *= s; dy[2] *= s;
  }
  Qd[p] = dx[0]; Qd[pl + p] = dx[1]; Qd[2 * pl + p] = dx[2];
  Qd[3 * pl + p] = dy[0]; Qd[4 * pl + p] = dy[1]; Qd[5 * pl + p] = dy[2];
}

// Viscous fluxes (NavierStokes2DParabolicFunction.c:112-143 and :155-186),
// Sutherland viscosity and the conductivity law of navierstokes2d.h:315-331
// (raiseto(x,1.5) = exp(1.5 log x), include/math_ops.h:40).
// FV planes: 0..2 = x-flux components 1..3, 3..5 = y-flux components 1..3.
__global__ void __launch_bounds__(TB) k_viscflux(Grid G, PhysParams P, const double *__restrict__ U,
                                                 const double *__restrict__ Qd, double *__restrict__ FV) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.nxg * G.nyg) return;
  const int i = (int)(t % G.nxg) - G.g, j = (int)(t / G.nxg) - G.g;
  const bool ix = (i >= 0 && i < G.nx), jy = (j >= 0 && j < G.ny);
  if (!ix && !jy) return;  // corners are never read
  const long pl = (long)G.nxg * G.nyg;
  const Prim q = prim_at(G, P.gamma, U, t);
  const double ux = Qd[t], vx = Qd[pl + t], Tx = Qd[2 * pl + t];
  const double uy = Qd[3 * pl + t], vy = Qd[4 * pl + t], Ty = Qd[5 * pl + t];
  const double Td = q.T * P.Tref;
  const double r15 = exp(1.5 * log(Td / P.T0));
  const double mu = r15 * (P.T0 + P.TS) / (Td + P.TS);
  const double kap = r15 * P.kappa_num / (Td + P.TA * exp(-P.TB / Td));
  const double muRe = mu * P.inv_Re;
  const double kc = kap * P.inv_Re * P.inv_gm1 * P.inv_Pr;
  const double txx = P.two_third * muRe * (2 * ux - vy);
  const double txy = muRe * (uy + vx);
  FV[t] = txx;
  FV[pl + t] = txy;
  FV[2 * pl + t] = q.vx * txx + q.vy * txy + kc * Tx;
  const double tyx = muRe * (uy + vx);
  const double tyy = P.two_third * muRe * (-ux + 2 * vy);
  FV[3 * pl + t] = tyx;
  FV[4 * pl + t] = tyy;
  FV[5 * pl + t] = q.vx * tyx + q.vy * tyy + kc * Ty;
}

// ---------------------------------------------------------------------------
// RHS assembly at interior points: rhs = -hyp + par + source
// (TimeRHSFunctionExplicit.c:55-74).  hyp: x part from the x-sweep (compact
// interior, natural order) plus the y divergence; par: second derivative of
// the viscous fluxes (central stencil at interior points); source:
// ChemistrySource (ChemistrySolve.c:18-115, chemistry.h:30-127).
// ---------------------------------------------------------------------------
template <int ND, int VISC, int PAR4, int CHEM>
__global__ void __launch_bounds__(TB) k_assemble(Grid G, PhysParams P, const double *__restrict__ U,
                                                 const double *__restrict__ hx, const double *__restrict__ fIy,
                                                 const double *__restrict__ FV, const double *__restrict__ dxinv,
                                                 const double *__restrict__ imap, double hnu_scale,
                                                 double *__restrict__ udot) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nint = (long)G.nx * G.ny;
  if (t >= nint) return;
  const int i = (int)(t % G.nx), j = (int)(t / G.nx);
  const long p = (long)(j + G.gy) * G.nxg + (i + G.g);
  const int nv = G.nv;
  double dyi = 0.0;
  if (ND == 2) dyi = dxinv[(G.nx + 2 * G.g) + G.g + j];
  // parabolic part for the momentum and energy components
  double par[4] = {0, 0, 0, 0};
  if (VISC) {
    const long pl = (long)G.nxg * G.nyg;
    const double dxi = dxinv[G.g + i];
    for (int c = 0; c < 3; c++) {
      const double *Fx = FV + c * pl, *Fy = FV + (3 + c) * pl;
      double dfx, dfy;
      if (PAR4) {
        const double t12 = 1.0 / 12.0;
        dfx = (Fx[p - 2] - 8 * Fx[p - 1] + 8 * Fx[p + 1] - Fx[p + 2]) * t12;
        dfy = (Fy[p - 2 * G.nxg] - 8 * Fy[p - G.nxg] + 8 * Fy[p + G.nxg] - Fy[p + 2 * G.nxg]) * t12;
      } else {
        dfx = 0.0 * Fx[p + 1] - (-1.0) * Fx[p] + (-1.0) * Fx[p - 1];
        dfy = 0.0 * Fy[p + G.nxg] - (-1.0) * Fy[p] + (-1.0) * Fy[p - G.nxg];
      }
      par[c + 1] = dxi * dfx;
      par[c + 1] = par[c + 1] + dyi * dfy;
    }
  }
  // chemistry source: photon density per z-layer, species rates, heating
  double src_e = 0.0;
  const int ie = (ND == 2) ? 3 : 2;
  for (int v = 0; v < nv; v++) {
    double hyp = hx[(long)v * nint + t];
    if (ND == 2) {
      const long a = ((long)v * (G.ny + 1) + j) * G.nx + i;
      hyp = hyp + dyi * (fIy[a + G.nx] - fIy[a]);
    }
    double r = -hyp;
    if (VISC && v >= 1 && v <= 3) r = r + par[v];
    else if (ND == 2) r = r + 0.0;
    if (!CHEM || v < P.nflow) {
      udot[(long)v * G.plane + p] = r;
    } else {
      udot[(long)v * G.plane + p] = r;  // species source added below
    }
  }
  if (CHEM) {
    const long pl = G.plane;
    double nh = (hnu_scale * imap[p]) / P.hnu_den;
    for (int iz = 0; iz < P.nlayers; iz++) {
      const int nfv = P.nflow + P.z_stride * iz;
      if (iz > 0) nh = nh * (1.0 - P.damp_coef * U[(long)(P.nflow + P.z_stride * (iz - 1) + 1) * pl + p]);
      const double nO2 = U[(long)(nfv + 0) * pl + p], nO3 = U[(long)(nfv + 1) * pl + p];
      const double n1D = U[(long)(nfv + 2) * pl + p], n1Dg = U[(long)(nfv + 3) * pl + p];
      const double n3Su = U[(long)(nfv + 4) * pl + p], n1Sg = U[(long)(nfv + 5) * pl + p];
      const double nCO2 = U[(long)(nfv + 6) * pl + p];
      double s[7];
      s[0] = 0.0;
      s[1] = -(P.k0a + P.k0b) * nh * nO3 - (P.k2a + P.k2b) * nO3 * n1D - (P.k3a + P.k3b) * nO3 * n1Sg -
             P.k5 * nO3 * n3Su - P.k6 * n1Dg * nO3;
      s[2] = P.k0a * nh * nO3 - (P.k1a + P.k1b) * n1D * nO2 - (P.k2a + P.k2b) * n1D * nO3 - P.k4 * n1D * nCO2;
      s[3] = P.k0a * nh * nO3 + P.k5 * nO3 * n3Su - P.k6 * n1Dg * nO3;
      s[4] = P.k2a * nO3 * n1D - P.k5 * nO3 * n3Su;
      s[5] = P.k1a * n1D * nO2 - (P.k3a + P.k3b) * nO3 * n1Sg;
      s[6] = 0.0;
      for (int m = 0; m < 7; m++) {
        double *o = udot + (long)(nfv + m) * pl + p;
        *o = *o + s[m];
      }
      src_e = ((P.q0a * P.k0a + P.q0b * P.k0b) * nh * nO3 + (P.q1a * P.k1a + P.q1b * P.k1b) * nO2 * n1D +
               (P.q2a * P.k2a + P.q2b * P.k2b) * nO3 * n1D + (P.q3a * P.k3a + P.q3b * P.k3b) * nO3 * n1Sg +
               P.q4 * P.k4 * n1D * nCO2 + P.q5 * P.k5 * nO3 * n3Su + P.q6 * P.k6 * n1Dg * nO3) /
              P.ch_gm1;
    }
    double *o = udot + (long)ie * pl + p;
    *o = *o + src_e;
  }
}

// u += (dt*b_s) * Udot_s, stage by stage (TimeRK.c:86-99), interior points.
__global__ void __launch_bounds__(TB) k_update(Grid G, double *__restrict__ u, const double *__restrict__ d0,
                                               const double *__restrict__ d1, const double *__restrict__ d2,
                                               const double *__restrict__ d3, double b0, double b1, double b2,
                                               double b3, int nst) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nint = (long)G.nx * G.ny;
  if (t >= nint * G.nv) return;
  const int v = (int)(t / nint);
  const long q = t % nint;
  const int i = (int)(q % G.nx), j = (int)(q / G.nx);
  const long p = (long)v * G.plane + (long)(j + G.gy) * G.nxg + (i + G.g);
  double x = u[p];
  x += b0 * d0[p];
  if (nst > 1) x += b1 * d1[p];
  if (nst > 2) x += b2 * d2[p];
  if (nst > 3) x += b3 * d3[p];
  u[p] = x;
}

// Per-block partial max CFL (Euler1DComputeCFL.c / NavierStokes2DComputeCFL.c)
// and sum of squared step differences (TimePostStep.c:41-53).
__global__ void __launch_bounds__(TB) k_cfl_norm(Grid G, int model, double gam, double dt,
                                                 const double *__restrict__ u, const double *__restrict__ u_prev,
                                                 const double *__restrict__ dxinv, double *__restrict__ partials) {
  __shared__ double smax[TB], ssum[TB];
  const long nint = (long)G.nx * G.ny;
  double mx = 0.0, sm = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nint; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % G.nx), j = (int)(t / G.nx);
    const long p = (long)(j + G.gy) * G.nxg + (i + G.g);
    if (model == 1) {
      const double rho = u_prev[p], v = u_prev[G.plane + p] / rho;
      const double P = (u_prev[2 * G.plane + p] - 0.5 * rho * v * v) * (gam - 1.0);
      const double c = sqrt(gam * P / rho);
      const double l = (fabs(v) + c) * dt * dxinv[G.g + i];
      mx = l > mx ? l : mx;
    } else {
      const double rho = u_prev[p];
      const double vx = (rho == 0) ? 0 : u_prev[G.plane + p] / rho, vy = (rho == 0) ? 0 : u_prev[2 * G.plane + p] / rho;
      const double vsq = (vx * vx) + (vy * vy);
      const double P = (u_prev[3 * G.plane + p] - 0.5 * rho * vsq) * (gam - 1.0);
      const double c = sqrt(gam * P / rho);
      const double lx = (fabs(vx) + c) * dt * dxinv[G.g + i];
      const double ly = (fabs(vy) + c) * dt * dxinv[(G.nx + 2 * G.g) + G.g + j];
      mx = lx > mx ? lx : mx;
      mx = ly > mx ? ly : mx;
    }
    for (int v = 0; v < G.nv; v++) {
      const double d = u_prev[(long)v * G.plane + p] - u[(long)v * G.plane + p];
      sm += d * d;
    }
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int s = TB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smax[threadIdx.x] = smax[threadIdx.x] > smax[threadIdx.x + s] ? smax[threadIdx.x] : smax[threadIdx.x + s];
      ssum[threadIdx.x] += ssum[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = smax[0];
    partials[2 * blockIdx.x + 1] = ssum[0];
  }
}

// Batched Thomas solve, one lane per system: tridiagLU.c:133-150, :243-255.
__global__ void __launch_bounds__(TB) k_tridiag(const double *__restrict__ a, const double *__restrict__ b,
                                                const double *__restrict__ c, double *__restrict__ x, int n, int ns,
                                                int *__restrict__ flag, double *__restrict__ sb) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ns) return;
  double bprev = b[d], xprev = x[d], cprev = c[d];
  sb[d] = bprev;
  if (bprev == 0) { atomicOr(flag, 1); return; }
  for (int i = 1; i < n; i++) {
    const long q = (long)i * ns + d;
    if (bprev == 0) { atomicOr(flag, 1); return; }
    const double fac = a[q] / bprev;
    const double bk = b[q] - fac * cprev, xk = x[q] - fac * xprev;
    sb[q] = bk;
    x[q] = xk;
    bprev = bk; xprev = xk; cprev = c[q];
  }
  {
    const long q = (long)(n - 1) * ns + d;
    if (sb[q] == 0) { atomicOr(flag, 1); return; }
    x[q] = x[q] / sb[q];
  }
  for (int i = n - 2; i >= 0; i--) {
    const long q = (long)i * ns + d;
    if (sb[q] == 0) { atomicOr(flag, 1); return; }
    x[q] = (x[q] - c[q] * x[q + ns]) / sb[q];
  }
}

__global__ void __launch_bounds__(TB) k_weno_weights(int kind, double c1, double c2, double c3,
                                                     const double *__restrict__ st, int count, double eps,
                                                     double *__restrict__ w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const double *s = st + 5 * (long)t;
  double w1, w2, w3;
  weno_w_rt(kind, c1, c2, c3, s[0], s[1], s[2], s[3], s[4], eps, w1, w2, w3);
  w[3 * (long)t] = w1;
  w[3 * (long)t + 1] = w2;
  w[3 * (long)t + 2] = w3;
}

}  // namespace

// ===========================================================================
// launchers
// ===========================================================================

void launch_stage_bc(const Grid &G, const BCMap &bc, const double *u, const double *const *udot, const double *a,
                     int nprev, double *U, double *UT, hipStream_t s) {
  const double *d0 = nprev > 0 ? udot[0] : nullptr, *d1 = nprev > 1 ? udot[1] : nullptr,
               *d2 = nprev > 2 ? udot[2] : nullptr;
  const double a0 = nprev > 0 ? a[0] : 0, a1 = nprev > 1 ? a[1] : 0, a2 = nprev > 2 ? a[2] : 0;
  if (G.ndims == 1) {
    hipLaunchKernelGGL(k_stage_bc1d, dim3(nblk((long)G.nxg * G.nv)), dim3(TB), 0, s, G, bc, u, d0, d1, d2, a0, a1,
                       a2, nprev, U);
  } else {
    constexpr int T = 32;
    dim3 grid((G.nxg + T - 1) / T, (G.nyg + T - 1) / T, G.nv);
    hipLaunchKernelGGL(k_stage_bc2d<T>, grid, dim3(T, 8), 0, s, G, bc, u, d0, d1, d2, a0, a1, a2, nprev, U, UT);
  }
}

void launch_flux(const Grid &G, const SweepGeom &S, int model, int dir, double gamma, const double *q, double *f,
                 hipStream_t s) {
  const long n = (long)(S.N + 2 * S.g) * S.nlines;
  if (model == 1) hipLaunchKernelGGL((k_flux<1, 0>), dim3(nblk(n)), dim3(TB), 0, s, S, G.nv, gamma, q, f);
  else if (dir == 0) hipLaunchKernelGGL((k_flux<2, 0>), dim3(nblk(n)), dim3(TB), 0, s, S, G.nv, gamma, q, f);
  else hipLaunchKernelGGL((k_flux<2, 1>), dim3(nblk(n)), dim3(TB), 0, s, S, G.nv, gamma, q, f);
}

void launch_recon_components(const Grid &G, const SweepGeom &S, const ReconParams &R, const double *q,
                             const double *f, double *uL, double *uR, double *fL, double *fR, double *scr_b,
                             double *scr_c, hipStream_t s) {
  if (R.crweno && R.kind < 0) {  // producer/consumer variant: kept for measurement, not dispatched
    const dim3 gr((S.nlines + 63) / 64, G.nv * 4), bl(256);
#define PC(K) hipLaunchKernelGGL((k_crweno_pc<K, 8>), gr, bl, 0, s, S, G.nv, R, q, f, uL, uR, fL, fR, scr_b, scr_c)
    switch (R.kind) { case 1: PC(1); break; case 2: PC(2); break; case 3: PC(3); break; default: PC(0); }
#undef PC
    return;
  }
  const long n = (long)S.nlines * G.nv * 4;
  const dim3 gr(nblk(n)), bl(TB);
#define RC(CR, K) hipLaunchKernelGGL((k_recon_comp<CR, K, 8>), gr, bl, 0, s, S, G.nv, R, q, f, uL, uR, fL, fR, scr_b, scr_c)
  if (R.crweno) {
    switch (R.kind) { case 1: RC(1, 1); break; case 2: RC(1, 2); break; case 3: RC(1, 3); break; default: RC(1, 0); }
  } else {
    switch (R.kind) { case 1: RC(0, 1); break; case 2: RC(0, 2); break; case 3: RC(0, 3); break; default: RC(0, 0); }
  }
#undef RC
}

void launch_recon_char_euler1d(const Grid &G, const SweepGeom &S, const ReconParams &R, double gamma,
                               const double *q, const double *f, double *uL, double *uR, double *fL, double *fR,
                               hipStream_t s) {
  const long n = (long)(S.N + 1) * S.nlines;
  const dim3 gr(nblk(n)), bl(TB);
#define RCH(K) hipLaunchKernelGGL((k_recon_char_e1d<K>), gr, bl, 0, s, S, G.nv, R, gamma, q, f, uL, uR, fL, fR)
  switch (R.kind) { case 1: RCH(1); break; case 2: RCH(2); break; case 3: RCH(3); break; default: RCH(0); }
#undef RCH
}

void launch_upwind(const Grid &G, const SweepGeom &S, int model, int upw, int dir, double gamma, const double *q,
                   const double *uL, const double *uR, const double *fL, const double *fR, double *fI,
                   hipStream_t s) {
  const long n = (long)(S.N + 1) * S.nlines;
  const dim3 gr(nblk(n)), bl(TB);
  if (model == 1) {
    if (upw == 1) hipLaunchKernelGGL((k_upwind_e1d<1>), gr, bl, 0, s, S, G.nv, gamma, q, uL, uR, fL, fR, fI);
    else hipLaunchKernelGGL((k_upwind_e1d<0>), gr, bl, 0, s, S, G.nv, gamma, q, uL, uR, fL, fR, fI);
  } else if (dir == 0) {
    hipLaunchKernelGGL((k_upwind_ns2d_roe<0>), gr, bl, 0, s, S, G.nv, gamma, q, uL, uR, fL, fR, fI);
  } else {
    hipLaunchKernelGGL((k_upwind_ns2d_roe<1>), gr, bl, 0, s, S, G.nv, gamma, q, uL, uR, fL, fR, fI);
  }
}

void launch_divergence(const Grid &G, const SweepGeom &S, const double *dxinv_dir, const double *fI, double *h,
                       hipStream_t s) {
  const long n = (long)S.N * S.nlines;
  hipLaunchKernelGGL(k_divergence, dim3(nblk(n)), dim3(TB), 0, s, S, G.nv, dxinv_dir, fI, h);
}

void launch_transpose(int nv, int rows, int cols, const double *src, double *dst, hipStream_t s) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32, nv);
  hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, s, rows, cols, src, dst);
}

void launch_qderiv(const Grid &G, const PhysParams &P, const double *U, const double *dxinv, double *Qd,
                   hipStream_t s) {
  const long n = (long)G.nxg * G.nyg;
  if (P.par4) hipLaunchKernelGGL((k_qderiv<1>), dim3(nblk(n)), dim3(TB), 0, s, G, P.gamma, U, dxinv, Qd);
  else hipLaunchKernelGGL((k_qderiv<0>), dim3(nblk(n)), dim3(TB), 0, s, G, P.gamma, U, dxinv, Qd);
}

void launch_viscflux(const Grid &G, const PhysParams &P, const double *U, const double *Qd, double *FV,
                     hipStream_t s) {
  const long n = (long)G.nxg * G.nyg;
  hipLaunchKernelGGL(k_viscflux, dim3(nblk(n)), dim3(TB), 0, s, G, P, U, Qd, FV);
}

void launch_assemble(const Grid &G, const PhysParams &P, const double *U, const double *hx, const double *fIy,
                     const double *FV, const double *dxinv, const double *imap, double hnu_scale, double *udot,
                     hipStream_t s) {
  const long n = (long)G.nx * G.ny;
  const dim3 gr(nblk(n)), bl(TB);
#define AS(ND, VI, P4, CH) \
  hipLaunchKernelGGL((k_assemble<ND, VI, P4, CH>), gr, bl, 0, s, G, P, U, hx, fIy, FV, dxinv, imap, hnu_scale, udot)
  if (G.ndims == 1) {
    if (P.chem) AS(1, 0, 0, 1); else AS(1, 0, 0, 0);
  } else if (P.viscous) {
    if (P.par4) { if (P.chem) AS(2, 1, 1, 1); else AS(2, 1, 1, 0); }
    else { if (P.chem) AS(2, 1, 0, 1); else AS(2, 1, 0, 0); }
  } else {
    if (P.chem) AS(2, 0, 0, 1); else AS(2, 0, 0, 0);
  }
#undef AS
}

void launch_update(const Grid &G, double *u, const double *const *udot, const double *bcoef, int nst, hipStream_t s) {
  const long n = (long)G.nx * G.ny * G.nv;
  hipLaunchKernelGGL(k_update, dim3(nblk(n)), dim3(TB), 0, s, G, u, udot[0], nst > 1 ? udot[1] : udot[0],
                     nst > 2 ? udot[2] : udot[0], nst > 3 ? udot[3] : udot[0], bcoef[0], nst > 1 ? bcoef[1] : 0.0,
                     nst > 2 ? bcoef[2] : 0.0, nst > 3 ? bcoef[3] : 0.0, nst);
}

int cfl_norm_blocks(const Grid &G) {
  const long n = (long)G.nx * G.ny;
  long b = (n + TB - 1) / TB;
  return (int)(b < 1024 ? b : 1024);
}

void launch_cfl_norm(const Grid &G, int model, double gamma, double dt, const double *u, const double *u_prev,
                     const double *dxinv, double *partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_cfl_norm, dim3(nblocks), dim3(TB), 0, s, G, model, gamma, dt, u, u_prev, dxinv, partials);
}

void launch_tridiag(const double *a, const double *b, const double *c, double *x, int n, int ns, int *flag,
                    double *scr_b, hipStream_t s) {
  hipLaunchKernelGGL(k_tridiag, dim3(nblk(ns)), dim3(TB), 0, s, a, b, c, x, n, ns, flag, scr_b);
}

void launch_weno_weights(int kind, const double *copt, const double *st, int count, double eps, double *w,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_weno_weights, dim3(nblk(count)), dim3(TB), 0, s, kind, copt[0], copt[1], copt[2], st, count,
                     eps, w);
}

}  // namespace piafs
