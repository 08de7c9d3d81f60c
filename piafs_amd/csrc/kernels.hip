// HIP kernels for gfx950 (MI355X): the PIAFS hyperbolic/parabolic/source
// right-hand side and the explicit Runge-Kutta stage algebra, FP64 throughout.
//
// Arithmetic follows the reference's association order term by term (the
// file:line of the formula being restated is given at each device function),
// and the library is compiled without floating-point contraction so that the
// rounding of every +, -, *, / and sqrt matches the reference build.
#include "kernels.hpp"

#include <cstdio>
#include <type_traits>

namespace piafs {

namespace {

constexpr int TB = 256;  // threads per block

inline unsigned nblk(long n, int tb = TB) { return (unsigned)((n + tb - 1) / tb); }

// ---------------------------------------------------------------------------
// Quotient n / d for the WENO weights and the CRWENO5 multipliers: the
// target's IEEE double division (a reciprocal estimate, two Newton-Raphson
// refinements and one residual correction, Markstein's method) without its
// range scaling and special-case fix-up steps, in three fewer instructions
// and with no condition register between them.  Those steps are identities,
// and the result is the IEEE quotient's bits, while n is +0 or of magnitude at
// least 2^-969, d is normal and positive, and the quotient lies within
// 2^-969 .. 2^1022: for the weights, whenever the smoothness indicators stay
// within about 2^400 of epsilon (data of ordinary magnitude with the usual
// epsilons).  Far outside (epsilon 1e-100 with data of 1e70), a weight below
// about 1e-290 can differ in its last bits, and where a Z or YC ratio
// overflows the row's weights are all NaN rather than partly NaN: the
// interpolated value is still NaN as the reference's is, and a tiny weight's
// term lies far below the last bit of a sum led by the large weights.
// ---------------------------------------------------------------------------
#ifndef PIAFS_FAST_WDIV
#define PIAFS_FAST_WDIV 1
#endif
__device__ __forceinline__ double wdiv(double n, double d) {
#if PIAFS_FAST_WDIV
  const double r0 = __builtin_amdgcn_rcp(d);
  const double e0 = __builtin_fma(-d, r0, 1.0);
  const double r1 = __builtin_fma(r0, e0, r0);
  const double e1 = __builtin_fma(-d, r1, 1.0);
  const double r2 = __builtin_fma(r1, e1, r1);
  const double q = n * r2;
  const double t = __builtin_fma(-d, q, n);
  return __builtin_fma(t, r2, q);
#else
  return n / d;
#endif
}

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
    const double r1 = wdiv(tau, b1 + eps), r2 = wdiv(tau, b2 + eps), r3 = wdiv(tau, b3 + eps);
    a1 = c1 * (1.0 + r1 * r1);
    a2 = c2 * (1.0 + r2 * r2);
    a3 = c3 * (1.0 + r3 * r3);
  } else {
    a1 = wdiv(c1, (b1 + eps) * (b1 + eps));
    a2 = wdiv(c2, (b2 + eps) * (b2 + eps));
    a3 = wdiv(c3, (b3 + eps) * (b3 + eps));
  }
  double s = wdiv(1.0, a1 + a2 + a3);
  w1 = a1 * s;
  w2 = a2 * s;
  w3 = a3 * s;
  if (KIND == 1) {
    a1 = wdiv(w1 * (c1 + c1 * c1 - 3 * c1 * w1 + w1 * w1), c1 * c1 + w1 * (1.0 - 2.0 * c1));
    a2 = wdiv(w2 * (c2 + c2 * c2 - 3 * c2 * w2 + w2 * w2), c2 * c2 + w2 * (1.0 - 2.0 * c2));
    a3 = wdiv(w3 * (c3 + c3 * c3 - 3 * c3 * w3 + w3 * w3), c3 * c3 + w3 * (1.0 - 2.0 * c3));
    s = wdiv(1.0, a1 + a2 + a3);
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

// Linear fifth-order upwind interpolation (upw5),
// Interp1PrimFifthOrderUpwind.c:92-96 and :138-142, in the reference's
// operation order (sums left to right).
__device__ __forceinline__ double upw5_value(double m3, double m2, double m1, double p1, double p2) {
  const double c30 = 1.0 / 30.0, c13 = 13.0 / 60.0, c47 = 47.0 / 60.0, c27 = 27.0 / 60.0, c20 = 1.0 / 20.0;
  return c30 * m3 - c13 * m2 + c47 * m1 + c27 * p1 - c20 * p2;
}

// Coefficients of row k of one line system: A, B, C and the right-hand side
// (for WENO5 and upw5, KIND 4, the right-hand side is the interface value
// itself).
template <int CR, int KIND>
__device__ __forceinline__ void row_coeffs(int k, int N, bool left, const ReconParams &R, double w0, double w1,
                                           double w2, double w3, double w4, double w5, double &Ak, double &Bk,
                                           double &Ck, double &Rk) {
  double m3, m2, m1, p1, p2;
  if (left) { m3 = w0; m2 = w1; m1 = w2; p1 = w3; p2 = w4; }
  else { m3 = w5; m2 = w4; m1 = w3; p1 = w2; p2 = w1; }
  if (KIND == 4) {
    Ak = 0.0; Bk = 1.0; Ck = 0.0;
    Rk = upw5_value(m3, m2, m1, p1, p2);
    return;
  }
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

// ---------------------------------------------------------------------------
// 2D Navier-Stokes flow eigenstructure along direction DIR (0: x, 1: y).
// Only the 4x4 flow block is formed: the species block of the reference's
// n x n matrices is the identity (L, R) with eigenvalue vn.
// ---------------------------------------------------------------------------
// Roe average of two conserved flow states (navierstokes2d.h:115-160)
__device__ __forceinline__ void ns2d_roeavg(double g, const double uL[4], const double uR[4], double ua[4]) {
  const double rhoL = uL[0], vxL = uL[1] / rhoL, vyL = uL[2] / rhoL, eL = uL[3];
  const double vsqL = (vxL * vxL) + (vyL * vyL), PL = (eL - 0.5 * rhoL * vsqL) * (g - 1.0);
  const double cLsq = g * PL / rhoL, HL = 0.5 * (vxL * vxL + vyL * vyL) + cLsq / (g - 1.0);
  const double rhoR = uR[0], vxR = uR[1] / rhoR, vyR = uR[2] / rhoR, eR = uR[3];
  const double vsqR = (vxR * vxR) + (vyR * vyR), PR = (eR - 0.5 * rhoR * vsqR) * (g - 1.0);
  const double cRsq = g * PR / rhoR, HR = 0.5 * (vxR * vxR + vyR * vyR) + cRsq / (g - 1.0);
  const double tL = sqrt(rhoL), tR = sqrt(rhoR);
  const double rA = tL * tR;
  const double vxA = (tL * vxL + tR * vxR) / (tL + tR), vyA = (tL * vyL + tR * vyR) / (tL + tR);
  const double HA = (tL * HL + tR * HR) / (tL + tR), vsqA = vxA * vxA + vyA * vyA;
  const double csqA = (g - 1.0) * (HA - 0.5 * vsqA), PA = csqA * rA / g;
  ua[0] = rA;
  ua[1] = rA * vxA;
  ua[2] = rA * vyA;
  ua[3] = PA / (g - 1.0) + 0.5 * rA * vsqA;
}

// Flow eigenvalues, in the reference's order (navierstokes2d.h:161-184)
template <int DIR>
__device__ __forceinline__ void ns2d_eigvals(double g, const double u[4], double D[4]) {
  const double rho = u[0], vx = (rho == 0) ? 0 : u[1] / rho, vy = (rho == 0) ? 0 : u[2] / rho, e = u[3];
  const double vsq = (vx * vx) + (vy * vy), P = (e - 0.5 * rho * vsq) * (g - 1.0);
  const double a = sqrt(g * P / rho), vn = DIR == 0 ? vx : vy;
  D[0] = vn - a;
  if (DIR == 0) { D[1] = vn + a; D[2] = vn; }
  else { D[1] = vn; D[2] = vn + a; }
  D[3] = vn;
}

// |v_n| of the Roe average of two cells with the Harten entropy fix
// (delta = 1e-6, NavierStokes2DUpwind.c:75-78): the species' dissipation
// coefficient of the Roe flux, as upwind_ns2d_roe computes it
template <int DIR>
__device__ __forceinline__ double ns2d_roe_dsp(double g, const double uLc[4], const double uRc[4]) {
  double ua[4], D[4];
  ns2d_roeavg(g, uLc, uRc, ua);
  ns2d_eigvals<DIR>(g, ua, D);
  const double vn = D[3];
  const double delta = 0.000001, delta2 = delta * delta;
  return fabs(vn) < delta ? (vn * vn + delta2) / (2 * delta) : fabs(vn);
}

// Left (rows) and right (columns) flow eigenvectors (navierstokes2d.h:186-311)
template <int DIR>
__device__ __forceinline__ void ns2d_eigvecs(double g, const double u[4], double L[4][4], double R[4][4]) {
  const double gm1 = g - 1.0;
  const double rho = u[0], vx = (rho == 0) ? 0 : u[1] / rho, vy = (rho == 0) ? 0 : u[2] / rho, e = u[3];
  const double vsq = (vx * vx) + (vy * vy), P = (e - 0.5 * rho * vsq) * (g - 1.0);
  const double a = sqrt(g * P / rho), ek = 0.5 * (vx * vx + vy * vy);
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
  const double h0 = a * a / gm1 + ek;
  const double nx = DIR == 0 ? 1.0 : 0.0, ny = DIR == 0 ? 0.0 : 1.0, un = DIR == 0 ? vx : vy;
  R[0][0] = 1.0; R[1][0] = vx - a * nx; R[2][0] = vy - a * ny; R[3][0] = h0 - a * un;
  R[0][3] = 1.0; R[1][3] = vx; R[2][3] = vy; R[3][3] = ek;
  const int cp = DIR == 0 ? 1 : 2, cz = DIR == 0 ? 2 : 1;
  R[0][cp] = 1.0; R[1][cp] = vx + a * nx; R[2][cp] = vy + a * ny; R[3][cp] = h0 + a * un;
  R[0][cz] = 0.0; R[1][cz] = ny; R[2][cz] = -nx; R[3][cz] = vx * ny - vy * nx;
}

// rf-char selection for one characteristic field (Euler1DUpwindRF,
// Euler1DUpwind.c:160-173; NavierStokes2DUpwindRF, NavierStokes2DUpwind.c:152-166):
// upwind when the field's eigenvalue has one sign at the left, averaged and
// right states, else a local Lax-Friedrichs flux with the largest |eigenvalue|.
// With llf set, the local Lax-Friedrichs flux is taken for every field
// (Euler1DUpwindLLF, Euler1DUpwind.c:259-261; NavierStokes2DUpwindLLF,
// NavierStokes2DUpwind.c:248-250).
__device__ __forceinline__ double rf_pick(double eL, double eC, double eR, double ucL, double ucR, double fcL,
                                          double fcR, bool llf) {
  if (!llf && eL > 0 && eC > 0 && eR > 0) return fcL;
  if (!llf && eL < 0 && eC < 0 && eR < 0) return fcR;
  const double x1 = fabs(eL), x2 = fabs(eC), x3 = fabs(eR);
  const double mx1 = x1 > x2 ? x1 : x2, mx2 = x2 > x3 ? x2 : x3;
  const double alpha = mx1 > mx2 ? mx1 : mx2;
  return 0.5 * (fcL + fcR + alpha * (ucL - ucR));
}

// ---------------------------------------------------------------------------
// Ensemble indexing: block -> case, local work item
// ---------------------------------------------------------------------------
#define CASE_ARGS const CaseDev *__restrict__ cases, const int *__restrict__ blockcase, \
                  const int *__restrict__ firstblock, int stage
// Workgroups go to the eight XCDs in turn, each XCD with its own L2.  The tile
// kernels take their tiles in this order: each XCD gets one contiguous band of
// the case's nb tiles, so that the cache lines two neighbouring tiles share are
// fetched into one L2.  It is a bijection of 0..nb-1; a case whose blocks do not
// start on a multiple of eight keeps workgroup order.
__device__ __forceinline__ int xcd_order(int id, int nb, int first) {
  if (!PIAFS_XCD_ORDER || (first & 7)) return id;
  const int q = nb >> 3, r = nb & 7, x = id & 7, k = id >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

#define CASE_PREAMBLE                                                      \
  const int cidx = blockcase[blockIdx.x];                                  \
  const CaseDev &C = cases[cidx];                                          \
  const long t = (long)(blockIdx.x - firstblock[cidx]) * blockDim.x + threadIdx.x; \
  (void)stage;

// ---------------------------------------------------------------------------
// Stage state + boundary conditions.
// TimeRK.c:46-57 (U = u + sum_j dt*A[s][j]*Udot_j, zero coefficients skipped:
// adding 0*Udot leaves U unchanged), then ApplyBoundaryConditions
// (BCPeriodic.c:22-62: ghost <- opposite interior; BCExtrapolate.c:16-51:
// mirror copy).  Each ghost point is evaluated at its boundary source point,
// which equals applying the BC to the stage state.  Points outside every
// boundary zone (the corners) are zero, as in the reference where they are
// never written.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double stage_value(const CaseDev &C, int s, long p) {
  double val = C.u[p];
  const int ns = C.nstages;
  for (int j = 0; j < s; j++) {
    const double a = C.dt * C.A[s * ns + j];
    if (a != 0.0) val += a * C.Udot[j][p];
  }
  return val;
}

// ---------------------------------------------------------------------------
// Boundary conditions as a gather: each ghost point finds the last zone (in
// list order) that sets its variable and evaluates that zone's rule on the
// stage values of the interior points it reads.  The zones read interior
// points only, so list order matters only where zones overlap.
//  periodic BCPeriodic.c:22-62, extrapolate BCExtrapolate.c:16-51,
//  dirichlet BCDirichlet.c:16-45, reflect BCReflect.c:17-50, walls
//  BCSlipWall.c:24-135 / BCNoslipWall.c:22-80, BCSubsonicInflow.c:22-90,
//  BCSubsonicOutflow.c:21-78, BCSubsonicAmbivalent.c:30-125,
//  BCSupersonicInflow.c:26-80, BCSupersonicOutflow.c:25-80.
// A ghost point no zone sets keeps the value the reference never changes
// from its zero initialisation.
// ---------------------------------------------------------------------------
// _NavierStokes2DGetFlowVar_, navierstokes2d.h:70-79
__device__ __forceinline__ void ns2d_flowvar(const double u[4], double gam, double &rho, double &vx, double &vy,
                                             double &e, double &P) {
  rho = u[0];
  vx = (rho == 0) ? 0 : u[1] / rho;
  vy = (rho == 0) ? 0 : u[2] / rho;
  e = u[3];
  const double vsq = (vx * vx) + (vy * vy);
  P = (e - 0.5 * rho * vsq) * (gam - 1.0);
}

// does zone z set variable v (the flow BCs exist for ndims 2 only, the slip
// wall also for ndims 1; walls and outflows leave passive scalars alone)
__device__ __forceinline__ bool zone_sets(const Zone &z, int v, int ndims) {
  switch (z.kind) {
    case Z_PERIODIC: case Z_EXTRAPOLATE: case Z_DIRICHLET: case Z_REFLECT: return true;
    case Z_SPONGE: return false;
    case Z_SLIP_WALL: return ndims == 1 ? v < 3 : v < 4;
    case Z_SUBSONIC_INFLOW: case Z_SUPERSONIC_INFLOW: return ndims == 2;
    default: return ndims == 2 && v < 4;
  }
}


// stage values of the four flow variables at point p
__device__ __forceinline__ void stage_flow(const CaseDev &C, int s, long p, double u[4]) {
#pragma unroll
  for (int m = 0; m < 4; m++) u[m] = stage_value(C, s, (long)m * C.G.plane + p);
}

// value of variable v at ghost point (i, j) for stage s
__device__ double ghost_value(const CaseDev &C, int s, int v, int i, int j) {
  const Grid &G = C.G;
  const int nd = G.ndims;
  int d, face, gd, o;
  if (i < 0 || i >= G.nx) {
    d = 0; face = (i < 0) ? 1 : -1; gd = i; o = j;
    if (nd == 2 && (j < 0 || j >= G.ny)) return 0.0;  // corner: no zone reaches it
  } else {
    d = 1; face = (j < 0) ? 1 : -1; gd = j; o = i;
  }
  const int N = d ? G.ny : G.nx, od = 1 - d;
  int z = C.nzones - 1;
  for (; z >= 0; z--) {
    const Zone &Z = C.zones[z];
    if (Z.kind == Z_SPONGE || Z.dim != d || Z.face != face || !zone_sets(Z, v, nd)) continue;
    if (nd == 2 && (o < Z.is[od] || o >= Z.ie[od])) continue;
    break;
  }
  if (z < 0) return 0.0;
  const Zone &Z = C.zones[z];
  auto at = [&](int a, int b) -> long {  // ghosted offset of (index along d, index across d)
    const int ii = d ? b : a, jj = d ? a : b;
    return (long)(jj + G.gy) * G.nxg + (ii + G.g);
  };
  if (Z.kind == Z_DIRICHLET) return C.bcvals[Z.vals + v];
  if (Z.kind == Z_PERIODIC) {
    // the source index across d is taken without the zone offset (BCPeriodic.c:44-52)
    const int src = (face == 1) ? gd + N : gd - N;
    return stage_value(C, s, (long)v * G.plane + at(src, nd == 2 ? o - Z.is[od] : 0));
  }
  const int mir = (face == 1) ? -gd - 1 : 2 * N - gd - 1;
  const long pm = at(mir, nd == 2 ? o : 0);
  if (Z.kind == Z_EXTRAPOLATE) return stage_value(C, s, (long)v * G.plane + pm);
  if (Z.kind == Z_REFLECT) return (-1.0) * stage_value(C, s, (long)v * G.plane + pm);
  const double gam = Z.gamma, igm1 = 1.0 / (gam - 1.0);
  if (nd == 1) {  // slip wall, euler1d
    const double rho = stage_value(C, s, pm), uvel = stage_value(C, s, G.plane + pm) / rho;
    const double energy = stage_value(C, s, 2 * G.plane + pm);
    const double pressure = (energy - 0.5 * rho * uvel * uvel) * (gam - 1.0);
    const double rg = rho, pg = pressure, ug = 2.0 * Z.vel[0] - uvel;
    if (v == 0) return rg;
    if (v == 1) return rg * ug;
    return igm1 * pg + 0.5 * rg * ug * ug;
  }
  if (v >= 4) return C.bcvals[Z.vals + v - 4];  // inflow passive scalars
  double um[4], rho, uvel, vvel, energy, pressure, rg, ug, vg, pg;
  if (Z.kind == Z_SUBSONIC_AMBIVALENT) {
    const double nx = (d == 0 ? 1.0 : 0.0) * (double)face, ny = (d == 1 ? 1.0 : 0.0) * (double)face;
    const int i1 = (face == 1) ? 0 : N - 1, i2 = (face == 1) ? 1 : N - 2;
    double u1, v1, u2, v2;
    stage_flow(C, s, at(i1, o), um);
    ns2d_flowvar(um, gam, rho, u1, v1, energy, pressure);
    stage_flow(C, s, at(i2, o), um);
    ns2d_flowvar(um, gam, rho, u2, v2, energy, pressure);
    const double vn = (1.5 * u1 - 0.5 * u2) * nx + (1.5 * v1 - 0.5 * v2) * ny;
    stage_flow(C, s, pm, um);
    ns2d_flowvar(um, gam, rho, uvel, vvel, energy, pressure);
    if (vn > 0) { rg = Z.rho; pg = pressure; ug = Z.vel[0]; vg = Z.vel[1]; }
    else { rg = rho; pg = Z.p; ug = uvel; vg = vvel; }
  } else {
    stage_flow(C, s, pm, um);
    ns2d_flowvar(um, gam, rho, uvel, vvel, energy, pressure);
    rg = rho; pg = pressure; ug = uvel; vg = vvel;  // supersonic outflow
    switch (Z.kind) {
      case Z_SLIP_WALL:
        if (d == 0) ug = 2.0 * Z.vel[0] - uvel;
        else vg = 2.0 * Z.vel[1] - vvel;
        break;
      case Z_NOSLIP_WALL:
        ug = 2.0 * Z.vel[0] - uvel;
        vg = 2.0 * Z.vel[1] - vvel;
        break;
      case Z_SUBSONIC_INFLOW: rg = Z.rho; ug = Z.vel[0]; vg = Z.vel[1]; break;
      case Z_SUBSONIC_OUTFLOW: pg = Z.p; break;
      case Z_SUPERSONIC_INFLOW: rg = Z.rho; pg = Z.p; ug = Z.vel[0]; vg = Z.vel[1]; break;
      default: break;
    }
  }
  if (v == 0) return rg;
  if (v == 1) return rg * ug;
  if (v == 2) return rg * vg;
  return igm1 * pg + 0.5 * rg * (ug * ug + vg * vg);
}

template <int ST>
__device__ __forceinline__ void stage_body(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const Grid &G = C.G;
  const int s = stage;
  if (G.ndims == 1) {
    if (t >= (long)G.nx * G.nv) return;
    const int v = (int)(t / G.nx), i = (int)(t % G.nx);
    const long p = (long)v * G.plane + i + G.g;
    C.U[p] = stage_value(C, s, p);
    return;
  }
  // 2D interior: ST x ST tiles, 256 threads; natural write, then transposed
  // write via LDS.  The ghost layer is k_ghost's.
  constexpr int SR = 256 / ST;
  __shared__ double tile[ST][ST + 1];
  const int ntx = (G.nx + ST - 1) / ST, nty = (G.ny + ST - 1) / ST;
  // (64-point tiles: workgroup order, which timed faster for them at 4096^2);
  // the case's block count is padded to a multiple of eight
  const int nb = G.nv * ntx * nty;
  const int lid = (int)(blockIdx.x - firstblock[cidx]);
  const int id = ST == 32 ? xcd_order(lid, (nb + 7) & ~7, firstblock[cidx]) : lid;
  if (id >= nb) return;
  const int v = id / (ntx * nty), rem = id % (ntx * nty);
  const int i0 = (rem % ntx) * ST, j0 = (rem / ntx) * ST;
  const int tx = threadIdx.x & (ST - 1), ty = threadIdx.x / ST;
  const long pl = (long)v * G.plane;
  for (int r = ty; r < ST; r += SR) {
    const int i = i0 + tx, j = j0 + r;
    double val = 0.0;
    if (i < G.nx && j < G.ny) {
      const long p = pl + (long)(j + G.gy) * G.nxg + (i + G.g);
      val = stage_value(C, s, p);
      C.U[p] = val;
    }
    tile[r][tx] = val;
  }
  __syncthreads();
  for (int r = ty; r < ST; r += SR) {
    const int i = i0 + r, j = j0 + tx;
    if (i < G.nx && j < G.ny) C.UT[pl + (long)(i + G.g) * G.nyg + (j + G.gy)] = tile[tx][r];
  }
}

// 32 x 32 tiles, or 64 x 64 when every case of the plan has at least
// 2048 x 2048 points (rows of 512 B in both layouts; 16 % faster there, 12 %
// slower at 1024 x 1024, where the smaller tiles give twice the blocks)
__global__ void __launch_bounds__(256) k_stage(CASE_ARGS) { stage_body<32>(cases, blockcase, firstblock, stage); }
__global__ void __launch_bounds__(256) k_stage64(CASE_ARGS) { stage_body<64>(cases, blockcase, firstblock, stage); }

// Ghost points of the stage state, one thread per (variable, ghost point):
// the g rows below and above the interior (full ghosted width), then the g
// columns left and right of it (interior height).  Written to both layouts.
__host__ __device__ inline long ghost_points(const Grid &G) {
  return G.ndims == 1 ? 2L * G.g : 2L * G.gy * G.nxg + 2L * G.g * G.ny;
}

__global__ void __launch_bounds__(256) k_ghost(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const Grid &G = C.G;
  const long ng = ghost_points(G);
  if (t >= ng * G.nv) return;
  const int v = (int)(t / ng);
  long q = t % ng;
  int i, j = 0;
  if (G.ndims == 1) {
    i = (q < G.g) ? (int)q - G.g : G.nx + (int)(q - G.g);
  } else if (q < 2L * G.gy * G.nxg) {
    const int row = (int)(q / G.nxg);
    i = (int)(q % G.nxg) - G.g;
    j = row < G.gy ? row - G.gy : G.ny + (row - G.gy);
  } else {
    q -= 2L * G.gy * G.nxg;
    const int col = (int)(q / G.ny);
    j = (int)(q % G.ny);
    i = col < G.g ? col - G.g : G.nx + (col - G.g);
  }
  const double val = ghost_value(C, stage, v, i, j);
  const long pl = (long)v * G.plane;
  C.U[pl + (long)(j + G.gy) * G.nxg + (i + G.g)] = val;
  if (G.ndims == 2) C.UT[pl + (long)(i + G.g) * G.nyg + (j + G.gy)] = val;
}


// ---------------------------------------------------------------------------
// Cell-centred flux at every position of every interior line, both sweeps:
// the x-sweep from the natural state, the y-sweep from the transposed copy.
// NavierStokes2DFlux.c:24-52 (navierstokes2d.h:70-110), Euler1DFlux.c:19
// (euler1d.h:61-85).
// ---------------------------------------------------------------------------
// The flux along dir at one point (NavierStokes2DFlux.c:24-52 with
// navierstokes2d.h:70-110; Euler1DFlux.c:19 with euler1d.h:61-85): q(m) reads
// variable m of the point, put(m, value) writes component m of the flux.
template <class Q, class Put>
__device__ __forceinline__ void point_flux(int model, int dir, double gamma, int nv, Q q, Put put) {
  if (model == 1) {
    const double rho = q(0), v = q(1) / rho, e = q(2);
    const double P_ = (e - 0.5 * rho * v * v) * (gamma - 1.0);
    put(0, rho * v);
    put(1, rho * v * v + P_);
    put(2, (e + P_) * v);
    for (int m = 3; m < nv; m++) put(m, v * q(m));
  } else {
    const double rho = q(0);
    const double vx = (rho == 0) ? 0 : q(1) / rho;
    const double vy = (rho == 0) ? 0 : q(2) / rho;
    const double e = q(3);
    const double vsq = (vx * vx) + (vy * vy);
    const double P_ = (e - 0.5 * rho * vsq) * (gamma - 1.0);
    const double vn = dir == 0 ? vx : vy;
    put(0, rho * vn);
    if (dir == 0) {
      put(1, rho * vx * vx + P_);
      put(2, rho * vx * vy);
    } else {
      put(1, rho * vy * vx);
      put(2, rho * vy * vy + P_);
    }
    put(3, (e + P_) * vn);
    for (int m = 4; m < nv; m++) put(m, vn * q(m));
  }
}

__global__ void __launch_bounds__(256) k_flux(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const long nX = (long)(C.Sx.N + 2 * C.Sx.g) * C.Sx.nlines;
  const long nY = (C.G.ndims == 2) ? (long)(C.Sy.N + 2 * C.Sy.g) * C.Sy.nlines : 0;
  if (t >= nX + nY) return;
  const int dir = t < nX ? 0 : 1;
  const SweepGeom &S = dir == 0 ? C.Sx : C.Sy;
  const long tt = dir == 0 ? t : t - nX;
  const double *q = dir == 0 ? C.U : C.UT;
  double *f = dir == 0 ? C.flux_x : C.flux_y;
  const int P = S.N + 2 * S.g;
  const int pos = (int)(tt % P) - S.g, l = (int)(tt / P);
  const long b = S.base + (long)l * S.line_stride + pos;
  const long pl = S.plane;
  point_flux(C.model, dir, C.P.gamma, C.G.nv, [&](int m) { return q[m * pl + b]; },
             [&](int m, double x) { f[m * pl + b] = x; });
  if (C.fuse_sp && pos >= 0 && pos <= S.N) {  // |v_n| at interface pos, between cells pos-1 and pos
    const double uLc[4] = {q[b - 1], q[pl + b - 1], q[2 * pl + b - 1], q[3 * pl + b - 1]};
    const double uRc[4] = {q[b], q[pl + b], q[2 * pl + b], q[3 * pl + b]};
    C.dsp[dir][(long)l * (S.N + 1) + pos] =
        dir == 0 ? ns2d_roe_dsp<0>(C.P.gamma, uLc, uRc) : ns2d_roe_dsp<1>(C.P.gamma, uLc, uRc);
  }
}

// ---------------------------------------------------------------------------
// Tridiagonal line systems solved by teams of lanes, with the reference's
// arithmetic.
//
// A team of T lanes (a power of two; teams of more than 64 lanes span whole
// waves of one block) owns one system, and each lane keeps a segment of at
// most kMaxRows consecutive rows in registers.  The result is the one of the
// sequential Thomas sweep of tridiagLU.c:133-150 (fac = a_k/b'_{k-1},
// b'_k = b_k - fac c_{k-1}, x'_k = x_k - fac x'_{k-1}) and :243-255
// (x_k = (x'_k - c_k x_{k+1}) / b'_k), bit for bit:
//  1. each segment gets an estimate of the state entering it -- the pivot
//     b' from a scan of the segments' Moebius maps b'_{k-1} -> b'_k (2x2
//     matrices), x' from a scan of their affine maps;
//  2. every lane runs the sweep over its rows from its entry state, and each
//     lane's entry becomes the state its predecessor ended with; this repeats
//     until no entry changes.  Lane 0 starts from the exact state, so the
//     fixed point is reached after at most T passes and is exactly the
//     sequential sweep; with the estimates (and the sweep's contraction) the
//     ends are usually exact after the first pass, so two passes are typical;
//  3. the back substitution is treated the same way from the top.
// The systems are never written to memory; only the solution is.
// ---------------------------------------------------------------------------
struct Mob {
  double a, b, c, d;  // [[a, b], [c, d]]
};
struct Aff {
  double a, b;  // x -> a*x + b
};

// Rescale by a power of two (exact) so that products of many segment maps
// neither overflow nor underflow; a Moebius map is unchanged by scaling.
__device__ __forceinline__ Mob mob_norm(Mob r) {
  const double mx = fmax(fmax(fabs(r.a), fabs(r.b)), fmax(fabs(r.c), fabs(r.d)));
  if (mx > 0.0 && mx < INFINITY) {
    const int e = ilogb(mx);
    r.a = ldexp(r.a, -e);
    r.b = ldexp(r.b, -e);
    r.c = ldexp(r.c, -e);
    r.d = ldexp(r.d, -e);
  }
  return r;
}
__device__ __forceinline__ Mob mob_mul(const Mob &x, const Mob &y) {
  Mob r;
  r.a = x.a * y.a + x.b * y.c;
  r.b = x.a * y.b + x.b * y.d;
  r.c = x.c * y.a + x.d * y.c;
  r.d = x.c * y.b + x.d * y.d;
  return mob_norm(r);
}
// outer o inner
__device__ __forceinline__ Aff aff_cmp(const Aff &o, const Aff &i) { return Aff{o.a * i.a, o.a * i.b + o.b}; }

// Lane moves of a double by DPP (a VALU operand modifier, no LDS round trip):
// lanes without a source (or outside ROWMASK) keep their own value.  CTRL:
// row_shr:d 0x110+d and row_shl:d 0x100+d move within rows of 16 lanes,
// wave_shr:1 0x138 and wave_shl:1 0x130 across the wave, row_bcast:15 0x142
// and row_bcast:31 0x143 broadcast a row's last lane to the following rows.
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int l0 = (int)(unsigned)b, h0 = (int)(unsigned)(b >> 32);
  const int lo = __builtin_amdgcn_update_dpp(l0, l0, CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(h0, h0, CTRL, ROWMASK, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ Mob dpp_mob(const Mob &x) {
  return Mob{dpp_move<CTRL, ROWMASK>(x.a), dpp_move<CTRL, ROWMASK>(x.b), dpp_move<CTRL, ROWMASK>(x.c),
             dpp_move<CTRL, ROWMASK>(x.d)};
}
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ Aff dpp_aff(const Aff &x) {
  return Aff{dpp_move<CTRL, ROWMASK>(x.a), dpp_move<CTRL, ROWMASK>(x.b)};
}
// value of lane i-1 (UP) or i+1 of the wave
template <bool UP>
__device__ __forceinline__ double lane_step(double v) {
  return UP ? dpp_move<0x138>(v) : dpp_move<0x130>(v);
}

// Exclusive scans over the lanes of a team.  `lane` is the lane within the
// team; `red` is LDS for one Mob per wave of the block.
__device__ Mob team_scan_mob(Mob x, int lane, int T, double *red) {
  const int W = T < 64 ? T : 64;
  const int wl = lane & (W - 1);
#if PIAFS_SCAN_DPP
  // within rows of 16 lanes (a team of fewer lanes lies inside one row), then
  // across the rows of the wave from each row's last lane
  const int rl = wl & 15;
  if (W > 1) { const Mob o = dpp_mob<0x111>(x); if (rl >= 1) x = mob_mul(x, o); }
  if (W > 2) { const Mob o = dpp_mob<0x112>(x); if (rl >= 2) x = mob_mul(x, o); }
  if (W > 4) { const Mob o = dpp_mob<0x114>(x); if (rl >= 4) x = mob_mul(x, o); }
  if (W > 8) { const Mob o = dpp_mob<0x118>(x); if (rl >= 8) x = mob_mul(x, o); }
  if (W > 16) { const Mob o = dpp_mob<0x142, 0xA>(x); if (wl & 16) x = mob_mul(x, o); }
  if (W > 32) { const Mob o = dpp_mob<0x143, 0xC>(x); if (wl & 32) x = mob_mul(x, o); }
#else
  for (int d = 1; d < W; d <<= 1) {
    Mob o;
    o.a = __shfl_up(x.a, d, W);
    o.b = __shfl_up(x.b, d, W);
    o.c = __shfl_up(x.c, d, W);
    o.d = __shfl_up(x.d, d, W);
    if (wl >= d) x = mob_mul(x, o);
  }
#endif
  Mob pre = {1.0, 0.0, 0.0, 1.0};
  if (T > 64) {
    const int w = lane >> 6, wave = threadIdx.x >> 6;
    if (wl == 63) {
      double *s = red + 4 * wave;
      s[0] = x.a; s[1] = x.b; s[2] = x.c; s[3] = x.d;
    }
    __syncthreads();
    for (int k = 0; k < w; k++) {
      const double *s = red + 4 * (wave - w + k);
      pre = mob_mul(Mob{s[0], s[1], s[2], s[3]}, pre);
    }
    __syncthreads();
  }
  Mob e;
#if PIAFS_SCAN_DPP
  e = Mob{lane_step<true>(x.a), lane_step<true>(x.b), lane_step<true>(x.c), lane_step<true>(x.d)};
#else
  e.a = __shfl_up(x.a, 1, W);
  e.b = __shfl_up(x.b, 1, W);
  e.c = __shfl_up(x.c, 1, W);
  e.d = __shfl_up(x.d, 1, W);
#endif
  if (wl == 0) e = pre;
  else if (T > 64) e = mob_mul(e, pre);
  return e;
}

__device__ Aff team_scan_aff_fwd(Aff x, int lane, int T, double *red) {
  const int W = T < 64 ? T : 64;
  const int wl = lane & (W - 1);
#if PIAFS_SCAN_DPP
  const int rl = wl & 15;
  if (W > 1) { const Aff o = dpp_aff<0x111>(x); if (rl >= 1) x = aff_cmp(x, o); }
  if (W > 2) { const Aff o = dpp_aff<0x112>(x); if (rl >= 2) x = aff_cmp(x, o); }
  if (W > 4) { const Aff o = dpp_aff<0x114>(x); if (rl >= 4) x = aff_cmp(x, o); }
  if (W > 8) { const Aff o = dpp_aff<0x118>(x); if (rl >= 8) x = aff_cmp(x, o); }
  if (W > 16) { const Aff o = dpp_aff<0x142, 0xA>(x); if (wl & 16) x = aff_cmp(x, o); }
  if (W > 32) { const Aff o = dpp_aff<0x143, 0xC>(x); if (wl & 32) x = aff_cmp(x, o); }
#else
  for (int d = 1; d < W; d <<= 1) {
    const Aff o = {__shfl_up(x.a, d, W), __shfl_up(x.b, d, W)};
    if (wl >= d) x = aff_cmp(x, o);
  }
#endif
  Aff pre = {1.0, 0.0};
  if (T > 64) {
    const int w = lane >> 6, wave = threadIdx.x >> 6;
    if (wl == 63) {
      red[2 * wave] = x.a;
      red[2 * wave + 1] = x.b;
    }
    __syncthreads();
    for (int k = 0; k < w; k++) pre = aff_cmp(Aff{red[2 * (wave - w + k)], red[2 * (wave - w + k) + 1]}, pre);
    __syncthreads();
  }
#if PIAFS_SCAN_DPP
  Aff e = {lane_step<true>(x.a), lane_step<true>(x.b)};
#else
  Aff e = {__shfl_up(x.a, 1, W), __shfl_up(x.b, 1, W)};
#endif
  if (wl == 0) e = pre;
  else if (T > 64) e = aff_cmp(e, pre);
  return e;
}

__device__ Aff team_scan_aff_bwd(Aff x, int lane, int T, double *red) {
  const int W = T < 64 ? T : 64;
  const int wl = lane & (W - 1);
#if PIAFS_SCAN_DPP
  // within rows of 16 lanes, then from the first lane of the following rows
  // (no DPP broadcast runs backwards: two LDS permutes)
  const int rl = wl & 15, rw = W < 16 ? W : 16;
  if (W > 1) { const Aff o = dpp_aff<0x101>(x); if (rl + 1 < rw) x = aff_cmp(x, o); }
  if (W > 2) { const Aff o = dpp_aff<0x102>(x); if (rl + 2 < rw) x = aff_cmp(x, o); }
  if (W > 4) { const Aff o = dpp_aff<0x104>(x); if (rl + 4 < rw) x = aff_cmp(x, o); }
  if (W > 8) { const Aff o = dpp_aff<0x108>(x); if (rl + 8 < rw) x = aff_cmp(x, o); }
  if (W > 16) {
    const int src = ((int)(threadIdx.x & 63) | 15) + 1;  // first lane of the next row
    const Aff o = {__shfl(x.a, src & 63, 64), __shfl(x.b, src & 63, 64)};
    if (!(wl & 16)) x = aff_cmp(x, o);
  }
  if (W > 32) {
    const Aff o = {__shfl(x.a, 32, 64), __shfl(x.b, 32, 64)};
    if (!(wl & 32)) x = aff_cmp(x, o);
  }
#else
  for (int d = 1; d < W; d <<= 1) {
    const Aff o = {__shfl_down(x.a, d, W), __shfl_down(x.b, d, W)};
    if (wl + d < W) x = aff_cmp(x, o);
  }
#endif
  Aff suf = {1.0, 0.0};
  if (T > 64) {
    const int w = lane >> 6, nw = T >> 6, wave = threadIdx.x >> 6;
    if (wl == 0) {
      red[2 * wave] = x.a;
      red[2 * wave + 1] = x.b;
    }
    __syncthreads();
    for (int k = nw - 1; k > w; k--) suf = aff_cmp(Aff{red[2 * (wave - w + k)], red[2 * (wave - w + k) + 1]}, suf);
    __syncthreads();
  }
#if PIAFS_SCAN_DPP
  Aff e = {lane_step<false>(x.a), lane_step<false>(x.b)};
#else
  Aff e = {__shfl_down(x.a, 1, W), __shfl_down(x.b, 1, W)};
#endif
  if (wl == W - 1) e = suf;
  else if (T > 64) e = aff_cmp(e, suf);
  return e;
}

// (a, b) of the previous lane of the team (dir = +1) or of the next one
// (dir = -1); the first (last) lane of the team keeps its own values.
__device__ __forceinline__ void team_shift2(double &a, double &b, int lane, int T, int dir, double *red) {
  const int W = T < 64 ? T : 64;
  const int wl = lane & (W - 1);
  double pa, pb;
#if PIAFS_SCAN_DPP
  // across team boundaries too: a team's first (last) lane receives another
  // team's value, which it does not use
  if (dir > 0) {
    pa = lane_step<true>(a);
    pb = lane_step<true>(b);
  } else {
    pa = lane_step<false>(a);
    pb = lane_step<false>(b);
  }
#else
  if (dir > 0) {
    pa = __shfl_up(a, 1, W);
    pb = __shfl_up(b, 1, W);
  } else {
    pa = __shfl_down(a, 1, W);
    pb = __shfl_down(b, 1, W);
  }
#endif
  if (T > 64) {
    const int wave = threadIdx.x >> 6;
    const bool edge_out = dir > 0 ? (wl == 63) : (wl == 0);
    if (edge_out) {
      red[2 * wave] = a;
      red[2 * wave + 1] = b;
    }
    __syncthreads();
    const bool edge_in = dir > 0 ? (wl == 0 && lane != 0) : (wl == 63 && lane != T - 1);
    if (edge_in) {
      pa = red[2 * (wave - dir)];
      pb = red[2 * (wave - dir) + 1];
    }
    __syncthreads();
  }
  a = pa;
  b = pb;
}

// true on every lane of the wave (T <= 64) or block (T > 64) when `p` holds
// on any of them
__device__ __forceinline__ bool team_any(bool p, int T) {
  if (T > 64) return __syncthreads_or(p) != 0;
  return __any(p) != 0;
}

// v_rcp_f64 refined by two Newton steps: a reciprocal good to the last bits,
// used only for the estimates that start the exact passes
__device__ __forceinline__ double rcp_est(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(r, fma(-b, r, 1.0), r);
  r = fma(r, fma(-b, r, 1.0), r);
  return r;
}

// Forward elimination of team_solve: on exit B holds b' and X holds x' of
// every row, Fk the multipliers.  The row before the team's first row has
// pivot bfirst, super-diagonal cfirst and x' xfirst (an identity row: 1, 0,
// its value).  zp is set when a pivot is zero.  QD: the system is a CRWENO5
// coefficient system (a_k >= +0, true pivots positive and of order one), whose
// multipliers a_k / b'_{k-1} divide with wdiv; the relay ends on the true
// pivots, where wdiv gives the IEEE quotient's bits within its envelope (the
// weights in a_k not tiny).
template <int MR, bool FULL = false, bool QD = false>
__device__ __forceinline__ void team_forward(double (&A)[MR], double (&B)[MR], double (&Cc)[MR], double (&X)[MR],
                                             double (&Fk)[MR], int n, int lane, int T, double bfirst, double cfirst,
                                             double xfirst, double *red, bool plain_start, bool &zp) {
  const bool has = FULL || n > 0;
  double cin = 0.0, unused = 0.0;
#pragma unroll
  for (int i = 0; i < MR; i++)
    if (FULL ? i == MR - 1 : i == n - 1) cin = Cc[i];
  team_shift2(cin, unused, lane, T, +1, red);
  if (lane == 0) cin = cfirst;
  // pivot entering the segment: scan of M_k = [[b_k, -a_k c_{k-1}], [1, 0]] on (b'_{k-1}, 1)
  Mob P = {1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int i = 0; i < MR; i++) {
    if (FULL || i < n) {
      const double ac = A[i] * (i == 0 ? cin : Cc[i > 0 ? i - 1 : 0]);
      const Mob R = {B[i] * P.a - ac * P.c, B[i] * P.b - ac * P.d, P.a, P.b};
      P = R;
    }
  }
  P = mob_norm(P);
  const Mob E = team_scan_mob(P, lane, T, red);
  double bin = lane == 0 ? bfirst : (plain_start ? 1.0 : (E.a + E.b) / (E.c + E.d));
  // multipliers along the estimated pivots; affine maps of x'
  double bend = bin;
  Aff F = {lane == 0 ? 0.0 : 1.0, lane == 0 ? xfirst : 0.0};
  {
    double c = cin;
#pragma unroll
    for (int i = 0; i < MR; i++) {
      if (FULL || i < n) {
        const double fac = QD ? wdiv(A[i], bend) : A[i] / bend;
        bend = B[i] - fac * c;
        c = Cc[i];
        Fk[i] = fac;
        F = Aff{-(fac * F.a), X[i] - fac * F.b};
      } else {
        Fk[i] = 0.0;
      }
    }
  }
  const Aff Ef = team_scan_aff_fwd(F, lane, T, red);
  double xin = lane == 0 ? xfirst : (plain_start ? 0.0 : Ef.a * xfirst + Ef.b);
  // relay of segment ends until every entry equals its predecessor's end
  for (int it = 0;; it++) {
    double xend = xin;
#pragma unroll
    for (int i = 0; i < MR; i++)
      if (FULL || i < n) xend = X[i] - Fk[i] * xend;
    double nb = bend;
    team_shift2(nb, xend, lane, T, +1, red);
    const bool bchg = lane != 0 && !(nb == bin);
    const bool changed = has && lane != 0 && (bchg || !(xend == xin));
    if (lane != 0) {
      bin = nb;
      xin = xend;
    }
    if (!team_any(changed, T) || it > T) break;
    if (bchg) {  // pivots of this segment derived again from the new entry, until they repeat
      const double bend_old = bend;
      double c = cin, bb = bin;
      bool live = true;
#pragma unroll
      for (int i = 0; i < MR; i++) {
        if ((FULL || i < n) && live) {
          const double fac = QD ? wdiv(A[i], bb) : A[i] / bb;
          if (__double_as_longlong(fac) == __double_as_longlong(Fk[i])) {
            live = false;  // same multiplier, so the same pivot here and in every later row
          } else {
            bb = B[i] - fac * c;
            c = Cc[i];
            Fk[i] = fac;
          }
        }
      }
      bend = live ? bb : bend_old;
    }
  }
  // b' and x' of every row, in the reference's arithmetic
  zp = zp || ((bin == 0.0) && has);
  {
    double b = bin, x = xin, c = cin;
#pragma unroll
    for (int i = 0; i < MR; i++) {
      if (FULL || i < n) {
        b = B[i] - Fk[i] * c;
        x = X[i] - Fk[i] * x;
        c = Cc[i];
        zp = zp || (b == 0.0);
        B[i] = b;
        X[i] = x;
      }
    }
  }
}

// Back substitution of team_solve: B holds b', X holds x' of every row; on
// exit X holds the solution.  The row after the team's last row (held by the
// lane with is_last) has solution xlast.  W is scratch.
template <int MR, bool FULL = false>
__device__ __forceinline__ void team_backward(double (&B)[MR], double (&Cc)[MR], double (&X)[MR], double (&W)[MR],
                                              int n, int lane, int T, bool is_last, double xlast, double *red,
                                              bool plain_start) {
  const bool has = FULL || n > 0;
  // estimate of x entering from above, then the relay
  Aff Gm = {is_last ? 0.0 : 1.0, is_last ? xlast : 0.0};
#pragma unroll
  for (int i = MR - 1; i >= 0; i--) {
    if (FULL || i < n) {
      const double rb = rcp_est(B[i]);
      Gm = Aff{-(Cc[i] * rb) * Gm.a, (X[i] - Cc[i] * Gm.b) * rb};
    }
  }
  const Aff Eb = team_scan_aff_bwd(Gm, lane, T, red);
  double yin = is_last ? xlast : (plain_start ? 0.0 : Eb.a * xlast + Eb.b);
  for (int it = 0;; it++) {
    double xv = yin, dummy = 0.0;
    bool live = true;
#pragma unroll
    for (int i = MR - 1; i >= 0; i--) {
      if ((FULL || i < n) && live) {
        const double nx = (X[i] - Cc[i] * xv) / B[i];
        // from the second pass on: a row that repeats its value leaves every row above unchanged
        if (it > 0 && __double_as_longlong(nx) == __double_as_longlong(W[i])) live = false;
        W[i] = nx;
        xv = nx;
      }
    }
    if (FULL || n > 0) xv = W[0];
    team_shift2(xv, dummy, lane, T, -1, red);
    const bool changed = has && !is_last && !(xv == yin);
    if (!is_last) yin = xv;
    if (!team_any(changed, T) || it > T) break;
  }
#pragma unroll
  for (int i = 0; i < MR; i++)
    if (FULL || i < n) X[i] = W[i];
}

// Solve one team's system.  On entry the n rows of this lane hold
// (A, B, Cc, X) = (sub-, main, super-diagonal, right-hand side); on exit X
// holds the solution and B the pivots.  The row before the team's first row
// is an identity row carrying xfirst; the row after the team's last row (held
// by the lane with is_last) is one carrying xlast.  Every lane of every team
// of the block must call this (it may synchronise the block).
// The multipliers of the
// pivot chain (fac = a_k / b'_{k-1}, one division each) are kept per row, so
// the x' chain of a pass costs no division: every pass takes the segments'
// entry states, runs the x' chain along the kept multipliers, and relays the
// segment ends; only when an entry pivot changes are the multipliers derived
// again.  Once no entry changes, one pass without divisions writes b' and x'
// (fac * c and fac * x' exactly as the reference computes them).  The back
// substitution iterates the same way from the top.  If `plain_start` is set
// the scanned estimates are replaced by the trivial ones (b' = 1, x' = 0), so
// that the relay runs through many passes (testing).  FULL: every lane holds
// MR rows (n is not read), rows past the system's end being identity rows.
template <int MR, bool FULL = false, bool QD = false>
__device__ __forceinline__ void team_solve(double (&A)[MR], double (&B)[MR], double (&Cc)[MR], double (&X)[MR],
                                                int n, int lane, int T, bool is_last, double xfirst, double xlast,
                                                double *red, int *status, bool plain_start) {
  double Fk[MR];
  bool zp = false;
  team_forward<MR, FULL, QD>(A, B, Cc, X, Fk, n, lane, T, 1.0, 0.0, xfirst, red, plain_start, zp);
  team_backward<MR, FULL>(B, Cc, X, Fk, n, lane, T, is_last, xlast, red, plain_start);
  if (zp) atomicOr(status, 1);
}

// ---------------------------------------------------------------------------
// Component-wise WENO5 / CRWENO5 reconstruction.  One system per (line,
// variable, reconstruction r): r = 0 uL, 1 uR (of the state), 2 fL, 3 fR (of
// the flux); r even = left-biased.  A team of lanes covers the interfaces of
// one system: lane j computes rows 1 + j*m .. of the interior rows 1..N-1,
// lane 0 also interface 0 and the lane holding row N-1 also interface N.
//  * weights: WENOFifthOrderCalculateWeights{JS,M,Z,YC}
//    (WENOFifthOrderCalculateWeights.c:112-237 and the variants);
//    optimal weights :202-220 (WENO5 values at the two end interfaces of a
//    CRWENO line).
//  * WENO5: Interp1PrimFifthOrderWENO.c:124-166.
//  * CRWENO5: row assembly Interp1PrimFifthOrderCRWENO.c:132-201 (identity
//    rows with WENO5 right-hand sides at interfaces 0 and N), then the line
//    solve (team_solve above).
// ---------------------------------------------------------------------------
// Systems of a sweep: split mode, one per (line, var, r); paired mode, one
// team per (line, var, pair) solving the left- and right-biased systems of
// the pair in turn (pair 0: uL, uR; pair 1: fL, fR) and writing the two
// combinations the Roe upwinding uses, 0.5*(uR - uL) and 0.5*(fL + fR)
// (NavierStokes2DUpwind.c:58-60, Euler1DUpwind.c:45-52).
__host__ __device__ inline long recon_systems(const SweepGeom &S, int nv, int paired) {
  return S.longl ? 0 : (long)S.nlines * nv * (paired ? 2 : 4);
}

template <int CR, int KIND>
__device__ __forceinline__ void recon_team(const CaseDev &C, int dir, long sys, bool active, int lane, double *red,
                                           double *stash) {
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const int nv = C.G.nv, N = S.N, nl = S.nlines, m = S.rows, g = S.g;
  const bool paired = C.paired != 0;
  const long sy = active ? sys : 0;
  const int per = paired ? 2 : 4;
  const int q = (int)(sy % per), v = (int)((sy / per) % nv), l = (int)((sy / per) / nv);
  const int r0 = paired ? 2 * q : q;
  const double *src = ((r0 < 2) ? (dir ? C.UT : C.U) : (dir ? C.flux_y : C.flux_x)) + (long)v * S.plane + S.base +
                      (long)l * S.line_stride;
  double *const *io = dir ? C.iy : C.ix;
  const long line = ((long)v * nl + l) * (N + 1);
  const int lo = 1 + lane * m;
  int n = N - lo;
  n = n < 0 ? 0 : (n > m ? m : n);
  if (!active) n = 0;
  const int lastlane = (N - 2) / m;
  const int hi_pos = N + g - 1;
  const double *ends = C.ends + 4 * ((dir ? recon_systems(C.Sx, nv, 1) : 0) + sy);
  double X[kMaxRows];
  // solve the system of reconstruction r into X
  auto solve = [&](int r, int side) {
    const bool left = !(r & 1);
    // the end interfaces 0 and N (identity rows carrying WENO5 values, from k_recon_ends)
    double x0 = 0.0, xN = 0.0;
    if (CR && active && lane == 0) x0 = paired ? ends[2 * side] : io[r][line];
    if (CR && active && lane == lastlane) xN = paired ? ends[2 * side + 1] : io[r][line + N];
    // stencil window: positions lo-3 .. lo+n+1
    double win[kMaxRows + 5];
#pragma unroll
    for (int j = 0; j < kMaxRows + 5; j++) {
      const int pos = lo - 3 + j;
      win[j] = (n > 0 && j < n + 5 && pos <= hi_pos) ? src[pos] : 0.0;
    }
    double A[kMaxRows], B[kMaxRows], Cc[kMaxRows];
#pragma unroll
    for (int i = 0; i < kMaxRows; i++) {
      if (i < n) {
        row_coeffs<CR, KIND>(lo + i, N, left, C.R, win[i], win[i + 1], win[i + 2], win[i + 3], win[i + 4],
                             win[i + 5], A[i], B[i], Cc[i], X[i]);
      } else {
        A[i] = 0.0; B[i] = 1.0; Cc[i] = 0.0; X[i] = 0.0;
      }
    }
    if (CR)
      team_solve<kMaxRows, false, PIAFS_FAST_WDIV>(A, B, Cc, X, n, lane, S.team, lane == lastlane, x0, xN, red, C.status,
                           C.R.plain_start != 0);
  };
  solve(r0, 0);
  if (paired) {
#pragma unroll
    for (int i = 0; i < kMaxRows; i++) stash[i * 256 + threadIdx.x] = X[i];  // this lane's own rows
    solve(r0 + 1, 1);
  }
  if (!paired) {
#pragma unroll
    for (int i = 0; i < kMaxRows; i++)
      if (i < n) io[r0][line + lo + i] = X[i];
  } else {
    double *out = io[q] + line;  // io[0]: 0.5*(uR - uL), io[1]: 0.5*(fL + fR)
#pragma unroll
    for (int i = 0; i < kMaxRows; i++) {
      if (i < n) {
        const double xl = stash[i * 256 + threadIdx.x];
        out[lo + i] = q ? 0.5 * (xl + X[i]) : 0.5 * (X[i] - xl);
      }
    }
  }
}

// Paired mode with the team's line staged in LDS.  The two systems of a pair
// read the same line (the state for pair 0, the flux for pair 1), so the team
// loads it once, with consecutive lanes reading consecutive positions, into a
// region of its own; each lane then reads its stencil window from there for
// both systems.  The combinations are written back through the same region so
// that the global stores are consecutive too.  The bias of each system is a
// compile-time constant.  Arithmetic and results are those of recon_team.
// Region layout: line position p (-2 .. T*m+2) at lds_slot(p + 2), one unused
// slot after every 16 so that lanes reading 16 apart hit different banks.
__device__ __forceinline__ int lds_slot(int i) { return i + (i >> 4); }
__host__ __device__ inline int line_region(int T, int m) {
  const int n = T * m + 5;
  return n + (n >> 4) + 1;
}
// LDS doubles of the staged lines of one 256-thread block: the largest
// (256 / T) * line_region(T, m) over teams T = 1 .. 256 and m <= 16 (T = 1),
// plus the overrun of the last lane's window when its segment is short
constexpr int kLineLds = 256 * (16 + 5 + 2) + 32;

// Block barrier that orders LDS accesses only: global loads and stores issued
// before it stay in flight (a __syncthreads also waits for those).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// all lanes of the team have finished their LDS accesses before any lane goes on
__device__ __forceinline__ void team_lds_sync(int T) {
  if (T > 64) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
}

template <int CR, int KIND, bool FULL>
__device__ __forceinline__ void recon_pair_lds(const CaseDev &C, int dir, long sys, bool active, int lane,
                                               double *red, double *stash, double *lines) {
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const int nv = C.G.nv, N = S.N, nl = S.nlines, m = S.rows, g = S.g, T = S.team;
  const long sy = active ? sys : 0;
  const int q = (int)(sy % 2), v = (int)((sy / 2) % nv), l = (int)((sy / 2) / nv);
  const double *src = (q == 0 ? (dir ? C.UT : C.U) : (dir ? C.flux_y : C.flux_x)) + (long)v * S.plane + S.base +
                      (long)l * S.line_stride;
  const long line = ((long)v * nl + l) * (N + 1);
  const int lo = 1 + lane * m;
  int n = N - lo;
  n = n < 0 ? 0 : (n > m ? m : n);
  if (!active) n = 0;
  const int lastlane = (N - 2) / m;
  const int hi_pos = N + g - 1;
  const double *ends = C.ends + 4 * ((dir ? recon_systems(C.Sx, nv, 1) : 0) + sy);
  double *L = lines + (threadIdx.x / T) * line_region(T, m);
  // species pairs whose two teams are different waves (T = 64) meet through
  // an arrival counter per pair, cleared before any wave can reach it
  __shared__ unsigned pair_arrived[2];
  if (C.fuse_sp && T == 64) {
    if (threadIdx.x < 2) pair_arrived[threadIdx.x] = 0u;
    __syncthreads();
  }
  // the line, positions -2 .. T*m+2 (zero past its last ghost)
  const int tot = T * m + 5;
  for (int b = 0; b < tot; b += T) {
    const int idx = b + lane;
    if (idx < tot) {
      const int pos = idx - 2;
      L[lds_slot(idx)] = (active && pos <= hi_pos) ? src[pos] : 0.0;
    }
  }
  team_lds_sync(T);
  double X[kMaxRows];
  // FULL (every lane holds kMaxRows rows): the rows past row N-1 are identity
  // rows, the first of them -- interface N -- carrying xN, so that no row
  // needs a bound check; the identity row at interface N lies in the team
  // unless T*kMaxRows == N-1, when it follows the team as in the general case
  const int nlane = (N - 1) / kMaxRows;  // the lane holding row N in FULL mode
  auto solve = [&](auto left_c, int side) {
    constexpr bool LEFT = decltype(left_c)::value;
    const int xn_lane = FULL ? (nlane < T ? nlane : T - 1) : lastlane;
    // rows assembled: all kMaxRows in FULL mode (m == kMaxRows there; the
    // row bound stays a run-time value, which keeps the row assembly in
    // order and the register count at two waves per SIMD)
    const int nrow = FULL ? m : n;
    double x0 = 0.0, xN = 0.0;
    if (CR && active && lane == 0) x0 = ends[2 * side];
    if (CR && active && lane == xn_lane) xN = ends[2 * side + 1];
    // stencil window: positions lo-3 .. lo+n+1
    double win[kMaxRows + 5];
#pragma unroll
    for (int j = 0; j < kMaxRows + 5; j++) win[j] = (j < nrow + 5) ? L[lds_slot(lo - 1 + j)] : 0.0;
    double A[kMaxRows], B[kMaxRows], Cc[kMaxRows];
#pragma unroll
    for (int i = 0; i < kMaxRows; i++) {
      if (i < nrow) {
        row_coeffs<CR, KIND>(lo + i, N, LEFT, C.R, win[i], win[i + 1], win[i + 2], win[i + 3], win[i + 4],
                             win[i + 5], A[i], B[i], Cc[i], X[i]);
      } else {
        A[i] = 0.0; B[i] = 1.0; Cc[i] = 0.0; X[i] = 0.0;
      }
    }
    if (FULL && lo + kMaxRows > N) {
#pragma unroll
      for (int i = 0; i < kMaxRows; i++) {
        if (lo + i >= N) {
          A[i] = 0.0; B[i] = 1.0; Cc[i] = 0.0;
          X[i] = (lo + i == N) ? xN : 0.0;
        }
      }
    }
    if (CR) {
      if (FULL)
        team_solve<kMaxRows, true, PIAFS_FAST_WDIV>(A, B, Cc, X, kMaxRows, lane, T, lane == T - 1, x0, nlane < T ? 0.0 : xN, red,
                                   C.status, C.R.plain_start != 0);
      else
        team_solve<kMaxRows, false, PIAFS_FAST_WDIV>(A, B, Cc, X, n, lane, T, lane == lastlane, x0, xN, red, C.status,
                             C.R.plain_start != 0);
    }
  };
  solve(std::integral_constant<bool, true>(), 0);
#pragma unroll
  for (int i = 0; i < kMaxRows; i++) stash[i * 256 + threadIdx.x] = X[i];  // this lane's own rows
  solve(std::integral_constant<bool, false>(), 1);
  // a species of a case whose reconstruction writes the species fluxes: the
  // |v_n| of this lane's interfaces, loaded now so that the combination step
  // below hides their latency (the line's N + 1 <= T kMaxRows + 2 interfaces
  // take at most kMaxRows + 2 per lane)
  constexpr int KO = kMaxRows + 2;
  const bool fsp = C.fuse_sp && v >= 4;
  double dvv[KO];
#pragma unroll
  for (int it = 0; it < KO; it++) dvv[it] = 0.0;
  if (fsp && (T == 64 || q == 1)) {  // a team of at most 32 lanes writes when it is the flux team
    const double *dv = C.dsp[dir] + (long)l * (N + 1);
#pragma unroll
    for (int it = 0; it < KO; it++) {
      const int k = it * T + lane;
      if (active && k <= N) dvv[it] = dv[k];
    }
  }
  team_lds_sync(T);  // every window of the team is read
#pragma unroll
  for (int i = 0; i < kMaxRows; i++) {
    if (i < n) {
      const double xl = stash[i * 256 + threadIdx.x];
      L[lds_slot(lo + i)] = q ? 0.5 * (xl + X[i]) : 0.5 * (X[i] - xl);
    }
  }
  team_lds_sync(T);
  if (fsp) {
    // A species: its Roe flux f_I = 0.5 (fL + fR) - |v_n| 0.5 (uR - uL), the
    // upwinder's arithmetic, is written here from the combinations of the
    // pair's two teams (u and f sources of one line and species), which sit
    // side by side in the block.  A team of at most 32 lanes shares its wave
    // with its partner, so the wave's own LDS sync has made both visible and
    // the flux team writes.  Teams of 64 lanes are two waves: each publishes
    // its combination with a release increment of the pair's counter, and the
    // one that arrives second (acquire) writes, so neither waits.  The two end
    // interfaces take the combinations k_recon_ends wrote.
    bool writer = q == 1;
    if (T == 64) {
      unsigned prev = 0u;
      if (lane == 0)
        prev = __hip_atomic_fetch_add(&pair_arrived[(threadIdx.x >> 6) >> 1], 1u, __ATOMIC_ACQ_REL,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
      writer = __shfl(prev, 0, 64) == 1u;
    }
    if (writer && active) {
      const double *Lp = lines + ((threadIdx.x / T) ^ 1) * line_region(T, m);
      const double *Lf = q ? L : Lp, *Lu = q ? Lp : L;
      double *const *io = dir ? C.iy : C.ix;
      double *fo = (dir ? C.fIy : C.fIx) + line;
#pragma unroll
      for (int it = 0; it < KO; it++) {
        const int k = it * T + lane;
        if (k <= N) {
          const bool end = k == 0 || k == N;
          const double fs = end ? io[1][line + k] : Lf[lds_slot(k)];
          const double ud = end ? io[0][line + k] : Lu[lds_slot(k)];
          fo[k] = fs - dvv[it] * ud;
        }
      }
    }
    return;
  }
  double *out = (dir ? C.iy : C.ix)[q] + line;  // [0]: 0.5*(uR - uL), [1]: 0.5*(fL + fR)
  for (int b = 1; b < N; b += T) {
    const int k = b + lane;
    if (active && k < N) out[k] = L[lds_slot(k)];
  }
}

// The two end interfaces (0 and N) of every line system, one thread each:
// WENO5 values on identity rows (Interp1PrimFifthOrderCRWENO.c:132-201 with
// the WENO5 weights at the ends; plain WENO5 for the non-compact scheme).  In
// paired mode a thread does both systems of its pair: the raw values go to
// the case's `ends` table for the line solves, the combination to the output.
template <int CR, int KIND>
__device__ __forceinline__ double end_value(const CaseDev &C, const SweepGeom &S, const double *src, int r, int k) {
  double a_, b_, c_, x_;
  row_coeffs<CR, KIND>(k, S.N, !(r & 1), C.R, src[k - 3], src[k - 2], src[k - 1], src[k], src[k + 1], src[k + 2],
                       a_, b_, c_, x_);
  return x_;
}

template <int CR, int KIND>
__device__ __forceinline__ void recon_end(const CaseDev &C, int dir, long sys, int e) {
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const int nv = C.G.nv, N = S.N, nl = S.nlines;
  const bool paired = C.paired != 0;
  const int per = paired ? 2 : 4;
  const int q = (int)(sys % per), v = (int)((sys / per) % nv), l = (int)((sys / per) / nv);
  const int r0 = paired ? 2 * q : q;
  const double *src = ((r0 < 2) ? (dir ? C.UT : C.U) : (dir ? C.flux_y : C.flux_x)) + (long)v * S.plane + S.base +
                      (long)l * S.line_stride;
  const int k = e ? N : 0;
  double *const *io = dir ? C.iy : C.ix;
  const long at = ((long)v * nl + l) * (N + 1) + k;
  if (!paired) {
    io[r0][at] = end_value<CR, KIND>(C, S, src, r0, k);
    return;
  }
  const double xl = end_value<CR, KIND>(C, S, src, r0, k), xr = end_value<CR, KIND>(C, S, src, r0 + 1, k);
  double *ends = C.ends + 4 * ((dir ? recon_systems(C.Sx, nv, 1) : 0) + sys);
  ends[e] = xl;
  ends[2 + e] = xr;
  io[q][at] = q ? 0.5 * (xl + xr) : 0.5 * (xr - xl);
}

// Characteristic WENO5 for euler1d, one thread per (interface, line).
// WENOFifthOrderCalculateWeightsChar* (WENOFifthOrderCalculateWeights.c:761-930)
// and Interp1PrimFifthOrderWENOChar.c:127-195: L and R from the Roe average of
// the two cells adjacent to the interface; species rows are the identity, so
// species reconstruct component-wise.
template <int KIND>
__device__ __forceinline__ void recon_char_e1d(const SweepGeom &S, int nv, const ReconParams &R, double gam, int l,
                                               int k, const double *__restrict__ q, const double *__restrict__ f,
                                               double *const *io) {
  const int nl = S.nlines, N = S.N;
  const long pl = S.plane;
  const long base = S.base + (long)l * S.line_stride;
  auto at = [&](const double *a, int v, int pos) { return a[v * pl + base + pos]; };
  double a0, a1, a2;
  euler1d_roeavg(gam, at(q, 0, k - 1), at(q, 1, k - 1), at(q, 2, k - 1), at(q, 0, k), at(q, 1, k), at(q, 2, k),
                 a0, a1, a2);
  Eig3 E;
  euler1d_eig(gam, a0, a1, a2, E);
  const double s6 = 1.0 / 6.0;
  for (int r = 0; r < 4; r++) {
    const double *src = (r < 2) ? q : f;
    const bool left = !(r & 1);
    double *out = io[r];
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
      if (KIND == 4) {
        fch[vv] = upw5_value(s[0][vv], s[1][vv], s[2][vv], s[3][vv], s[4][vv]);
        continue;
      }
      double w1, w2, w3;
      if (R.no_limiting) { w1 = 0.1; w2 = 0.6; w3 = 0.3; }
      else weno_w<KIND>(0.1, 0.6, 0.3, s[0][vv], s[1][vv], s[2][vv], s[3][vv], s[4][vv], R.eps, w1, w2, w3);
      const double f1 = (2 * s6) * s[0][vv] - (7.0 * s6) * s[1][vv] + (11.0 * s6) * s[2][vv];
      const double f2 = (-s6) * s[1][vv] + (5.0 * s6) * s[2][vv] + (2 * s6) * s[3][vv];
      const double f3 = (2 * s6) * s[2][vv] + (5 * s6) * s[3][vv] - (s6)*s[4][vv];
      fch[vv] = w1 * f1 + w2 * f2 + w3 * f3;
    }
    for (int vv = 0; vv < 3; vv++)
      out[((long)vv * nl + l) * (N + 1) + k] = E.R[vv][0] * fch[0] + E.R[vv][1] * fch[1] + E.R[vv][2] * fch[2];
    for (int m = 3; m < nv; m++) {
      const double m3 = at(src, m, pos[0]), m2 = at(src, m, pos[1]), m1 = at(src, m, pos[2]);
      const double p1 = at(src, m, pos[3]), p2 = at(src, m, pos[4]);
      if (KIND == 4) {
        out[((long)m * nl + l) * (N + 1) + k] = upw5_value(m3, m2, m1, p1, p2);
        continue;
      }
      double w1, w2, w3, f1, f2, f3;
      if (R.no_limiting) { w1 = 0.1; w2 = 0.6; w3 = 0.3; }
      else weno_w<KIND>(0.1, 0.6, 0.3, m3, m2, m1, p1, p2, R.eps, w1, w2, w3);
      weno5_cands(m3, m2, m1, p1, p2, f1, f2, f3);
      out[((long)m * nl + l) * (N + 1) + k] = w1 * f1 + w2 * f2 + w3 * f3;
    }
  }
}

// Characteristic WENO5 for navierstokes2d along DIR, one thread per
// (interface, line): Interp1PrimFifthOrderWENOChar.c:89-201 with
// WENOFifthOrderCalculateWeightsChar* (WENOFifthOrderCalculateWeights.c:761-930).
// L and R come from the Roe average of the two cells adjacent to the
// interface (navierstokes2d.h:115-311).  The flow values of the six stencil
// positions k-3 .. k+2 are projected once and serve both biases; the species
// rows of L and R are the identity, so species reconstruct component-wise.
// Products are summed in index order, as MatVecMult does (the reference's
// zero entries add exact zeros).
template <int DIR, int KIND>
__device__ __forceinline__ void recon_char_ns2d(const SweepGeom &S, int nv, const ReconParams &P, double gam, int l,
                                                int k, const double *__restrict__ q, const double *__restrict__ f,
                                                double *const *io) {
  const int nl = S.nlines, N = S.N;
  const long pl = S.plane;
  const double *ql = q + S.base + (long)l * S.line_stride, *fl = f + S.base + (long)l * S.line_stride;
  const double uLc[4] = {ql[k - 1], ql[pl + k - 1], ql[2 * pl + k - 1], ql[3 * pl + k - 1]};
  const double uRc[4] = {ql[k], ql[pl + k], ql[2 * pl + k], ql[3 * pl + k]};
  double ua[4], L[4][4], R[4][4];
  ns2d_roeavg(gam, uLc, uRc, ua);
  ns2d_eigvecs<DIR>(gam, ua, L, R);
  const long vs = (long)nl * (N + 1), ii = (long)l * (N + 1) + k;
  const double s6 = 1.0 / 6.0;
  auto weno = [&](double m3, double m2, double m1, double p1, double p2) {
    if (KIND == 4) return upw5_value(m3, m2, m1, p1, p2);
    double w1, w2, w3;
    if (P.no_limiting) { w1 = 0.1; w2 = 0.6; w3 = 0.3; }
    else weno_w<KIND>(0.1, 0.6, 0.3, m3, m2, m1, p1, p2, P.eps, w1, w2, w3);
    const double f1 = (2 * s6) * m3 - (7.0 * s6) * m2 + (11.0 * s6) * m1;
    const double f2 = (-s6) * m2 + (5.0 * s6) * m1 + (2 * s6) * p1;
    const double f3 = (2 * s6) * m1 + (5 * s6) * p1 - (s6)*p2;
    return w1 * f1 + w2 * f2 + w3 * f3;
  };
#pragma unroll
  for (int src = 0; src < 2; src++) {  // 0: the solution (uL, uR), 1: the flux (fL, fR)
    const double *a = src ? fl : ql;
    double c[6][4];  // characteristic values at k-3 .. k+2
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const int pos = k - 3 + p;
      const double x0 = a[pos], x1 = a[pl + pos], x2 = a[2 * pl + pos], x3 = a[3 * pl + pos];
#pragma unroll
      for (int v = 0; v < 4; v++) c[p][v] = L[v][0] * x0 + L[v][1] * x1 + L[v][2] * x2 + L[v][3] * x3;
    }
#pragma unroll
    for (int side = 0; side < 2; side++) {  // 0: left-biased (k-3 .. k+1), 1: right-biased (k+2 .. k-2)
      double fch[4];
#pragma unroll
      for (int v = 0; v < 4; v++)
        fch[v] = side ? weno(c[5][v], c[4][v], c[3][v], c[2][v], c[1][v])
                      : weno(c[0][v], c[1][v], c[2][v], c[3][v], c[4][v]);
      double *out = io[2 * src + side] + ii;
#pragma unroll
      for (int i = 0; i < 4; i++) out[i * vs] = R[i][0] * fch[0] + R[i][1] * fch[1] + R[i][2] * fch[2] + R[i][3] * fch[3];
    }
    for (int m = 4; m < nv; m++) {
      const double *am = a + m * pl + k;
      const double x[6] = {am[-3], am[-2], am[-1], am[0], am[1], am[2]};
      io[2 * src][m * vs + ii] = weno(x[0], x[1], x[2], x[3], x[4]);
      io[2 * src + 1][m * vs + ii] = weno(x[5], x[4], x[3], x[2], x[1]);
    }
  }
}

// Characteristic reconstruction: one thread per (interface, line) of each
// sweep, x sweep first (from the natural layout), then y (from the
// transposed copy).  Kept apart from the line-solve kernel so that neither's
// register allocation limits the other's occupancy.
__global__ void __launch_bounds__(256) k_recon_char(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const long nX = (long)(C.Sx.N + 1) * C.Sx.nlines;
  const long nY = (C.G.ndims == 2) ? (long)(C.Sy.N + 1) * C.Sy.nlines : 0;
  if (t >= nX + nY) return;
  const int dir = t < nX ? 0 : 1;
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const long u = dir ? t - nX : t;
  const int k = (int)(u % (S.N + 1)), l = (int)(u / (S.N + 1));
  const int nv = C.G.nv;
  const double gam = C.P.gamma;
  if (C.model == 1) {
    switch (C.R.kind) {
      case 1: recon_char_e1d<1>(S, nv, C.R, gam, l, k, C.U, C.flux_x, C.ix); break;
      case 2: recon_char_e1d<2>(S, nv, C.R, gam, l, k, C.U, C.flux_x, C.ix); break;
      case 3: recon_char_e1d<3>(S, nv, C.R, gam, l, k, C.U, C.flux_x, C.ix); break;
      case 4: recon_char_e1d<4>(S, nv, C.R, gam, l, k, C.U, C.flux_x, C.ix); break;
      default: recon_char_e1d<0>(S, nv, C.R, gam, l, k, C.U, C.flux_x, C.ix); break;
    }
    return;
  }
  const double *q = dir ? C.UT : C.U, *f = dir ? C.flux_y : C.flux_x;
  double *const *io = dir ? C.iy : C.ix;
#define RCHAR(KIND)                                                                  \
  (dir ? recon_char_ns2d<1, KIND>(S, nv, C.R, gam, l, k, q, f, io)                  \
       : recon_char_ns2d<0, KIND>(S, nv, C.R, gam, l, k, q, f, io))
  switch (C.R.kind) {
    case 1: RCHAR(1); break;
    case 2: RCHAR(2); break;
    case 3: RCHAR(3); break;
    case 4: RCHAR(4); break;
    default: RCHAR(0); break;
  }
#undef RCHAR
}

// ---------------------------------------------------------------------------
// Characteristic CRWENO5: Interp1PrimFifthOrderCRWENOChar.c:96-280 with the
// reference's block solver (blocktridiagLU.c:159-180 and :279-301 on one
// process; _MatrixInvert_ etc., matops.h:16-130).  Each interface row ties
// the characteristic fields through L of the Roe average at that interface,
// so the unknowns are the physical interface values and every row is a block
// row.  L's species block is the identity, and the reference's dense n x n
// arithmetic then reduces exactly (up to the sign of zeros) to the flow block
// plus one scalar chain per species; both are computed here in the
// reference's operation order.
// ---------------------------------------------------------------------------

// block rows of every system at interface k of line l, one thread each, the
// line index fastest: a wave reads the same position of consecutive lines,
// contiguous in the other layout (the x sweep reads the transposed copy, the
// y sweep the natural one), and computes the flux there as k_flux does
template <int MODEL, int DIR, int KIND>
__device__ __forceinline__ void cwc_rows(const CaseDev &C, const SweepGeom &S, int l, int k, double *rows) {
  constexpr int NF = MODEL == 1 ? 3 : 4;
  const int nv = C.G.nv, nl = S.nlines, N = S.N;
  const double gam = C.P.gamma;
  // state at position pos of line l
  const double *q;
  long o0, ostr, pl;
  if (MODEL == 1) {
    q = C.U; o0 = S.base + (long)l * S.line_stride; ostr = 1; pl = S.plane;
  } else {
    const SweepGeom &O = DIR ? C.Sx : C.Sy;
    q = DIR ? C.U : C.UT; o0 = O.base + l; ostr = O.line_stride; pl = O.plane;
  }
  auto at = [&](int v, int pos) { return q[(long)v * pl + o0 + (long)pos * ostr]; };
  double L[NF][NF];
  if (MODEL == 1) {
    double a0, a1, a2;
    euler1d_roeavg(gam, at(0, k - 1), at(1, k - 1), at(2, k - 1), at(0, k), at(1, k), at(2, k), a0, a1, a2);
    Eig3 E;
    euler1d_eig(gam, a0, a1, a2, E);
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
      for (int j = 0; j < NF; j++) L[i][j] = E.L[i][j];
  } else {
    // the reference averages (k-1, k) for the left bias and (k, k-1) for the
    // right one; the Roe average is symmetric to the bit
    const double uLc[4] = {at(0, k - 1), at(1, k - 1), at(2, k - 1), at(3, k - 1)};
    const double uRc[4] = {at(0, k), at(1, k), at(2, k), at(3, k)};
    double ua[4], L4[4][4], R4[4][4];
    ns2d_roeavg(gam, uLc, uRc, ua);
    ns2d_eigvecs<DIR>(gam, ua, L4, R4);
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
      for (int j = 0; j < NF; j++) L[i][j] = L4[i][j];
  }
  const long vs = (long)nl * (N + 1), ii = (long)k * nl + l;
#pragma unroll
  for (int i = 0; i < NF; i++)
#pragma unroll
    for (int j = 0; j < NF; j++) rows[(long)(i * NF + j) * vs + ii] = L[i][j];
  const bool bnd = (k == 0 || k == N);
  // optimal weights: WENO5 at the physical ends, CRWENO5 inside
  // (WENOFifthOrderCalculateWeights.c:846-862)
  const double c1 = bnd ? 0.1 : 0.2, c2 = bnd ? 0.6 : 0.5, c3 = 0.3;
  const double s6 = 1.0 / 6.0, t3 = 1.0 / 3.0;
  const ReconParams &P = C.R;
  auto row = [&](int sys, int v, double m3, double m2, double m1, double p1, double p2) {
    double w1, w2, w3, f1, f2, f3, ca, cb, cc;
    if (P.no_limiting) { w1 = c1; w2 = c2; w3 = c3; }
    else weno_w<KIND>(c1, c2, c3, m3, m2, m1, p1, p2, P.eps, w1, w2, w3);
    if (bnd) {  // WENO5 row, B = L
      f1 = (2 * s6) * m3 - (7.0 * s6) * m2 + (11.0 * s6) * m1;
      f2 = (-s6) * m2 + (5.0 * s6) * m1 + (2 * s6) * p1;
      f3 = (2 * s6) * m1 + (5 * s6) * p1 - (s6)*p2;
      ca = 0.0; cb = 1.0; cc = 0.0;
    } else {
      f1 = (s6) * (m2 + 5 * m1);
      f2 = (s6) * (5 * m1 + p1);
      f3 = (s6) * (m1 + 5 * p1);
      const double wa = ((2 * t3) * w1 + (t3)*w2), wb = ((t3)*w1 + (2 * t3) * (w2 + w3)), wc = ((t3)*w3);
      const bool left = !(sys & 1);  // the right bias swaps the sub- and super-diagonal coefficients
      ca = left ? wa : wc;
      cb = wb;
      cc = left ? wc : wa;
    }
    const long b = (long)(NF * NF + (sys * nv + v) * 4) * vs + ii;
    rows[b] = ca;
    rows[b + vs] = cb;
    rows[b + 2 * vs] = cc;
    rows[b + 3 * vs] = w1 * f1 + w2 * f2 + w3 * f3;
  };
  // flow state, flux and characteristic values at k-3 .. k+2
  double cu[6][NF], cf[6][NF], vn[6];
#pragma unroll
  for (int p = 0; p < 6; p++) {
    const int pos = k - 3 + p;
    double x[NF], fx[NF];
#pragma unroll
    for (int j = 0; j < NF; j++) x[j] = at(j, pos);
    if (MODEL == 1) {
      const double rho = x[0], v = x[1] / rho, e = x[2], P_ = (e - 0.5 * rho * v * v) * (gam - 1.0);
      fx[0] = rho * v;
      fx[1] = rho * v * v + P_;
      fx[2] = (e + P_) * v;
      vn[p] = v;
    } else {
      const double rho = x[0], vx = (rho == 0) ? 0 : x[1] / rho, vy = (rho == 0) ? 0 : x[2] / rho, e = x[3];
      const double vsq = (vx * vx) + (vy * vy), P_ = (e - 0.5 * rho * vsq) * (gam - 1.0);
      const double v = DIR == 0 ? vx : vy;
      fx[0] = rho * v;
      if (DIR == 0) { fx[1] = rho * vx * vx + P_; fx[2] = rho * vx * vy; }
      else { fx[1] = rho * vy * vx; fx[2] = rho * vy * vy + P_; }
      fx[NF - 1] = (e + P_) * v;
      vn[p] = v;
    }
#pragma unroll
    for (int v = 0; v < NF; v++) {
      double au = L[v][0] * x[0], af = L[v][0] * fx[0];
#pragma unroll
      for (int j = 1; j < NF; j++) {
        au = au + L[v][j] * x[j];
        af = af + L[v][j] * fx[j];
      }
      cu[p][v] = au;
      cf[p][v] = af;
    }
  }
#pragma unroll
  for (int v = 0; v < NF; v++) {
    row(0, v, cu[0][v], cu[1][v], cu[2][v], cu[3][v], cu[4][v]);
    row(1, v, cu[5][v], cu[4][v], cu[3][v], cu[2][v], cu[1][v]);
    row(2, v, cf[0][v], cf[1][v], cf[2][v], cf[3][v], cf[4][v]);
    row(3, v, cf[5][v], cf[4][v], cf[3][v], cf[2][v], cf[1][v]);
  }
  for (int m = NF; m < nv; m++) {  // passive scalars: identity rows, flux vn u_m
    double x[6], fx[6];
#pragma unroll
    for (int p = 0; p < 6; p++) {
      x[p] = at(m, k - 3 + p);
      fx[p] = vn[p] * x[p];
    }
    row(0, m, x[0], x[1], x[2], x[3], x[4]);
    row(1, m, x[5], x[4], x[3], x[2], x[1]);
    row(2, m, fx[0], fx[1], fx[2], fx[3], fx[4]);
    row(3, m, fx[5], fx[4], fx[3], fx[2], fx[1]);
  }
}

__global__ void __launch_bounds__(256) k_cw_rows(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const long nX = (long)(C.Sx.N + 1) * C.Sx.nlines;
  const long nY = (C.G.ndims == 2) ? (long)(C.Sy.N + 1) * C.Sy.nlines : 0;
  if (t >= nX + nY) return;
  const int dir = t < nX ? 0 : 1;
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const long u = dir ? t - nX : t;
  const int l = (int)(u % S.nlines), k = (int)(u / S.nlines);
  double *rows = C.cwr[dir];
#define CWROWS(KIND)                                                \
  (C.model == 1 ? cwc_rows<1, 0, KIND>(C, S, l, k, rows)           \
                : (dir ? cwc_rows<2, 1, KIND>(C, S, l, k, rows)    \
                       : cwc_rows<2, 0, KIND>(C, S, l, k, rows)))
  switch (C.R.kind) {
    case 1: CWROWS(1); break;
    case 2: CWROWS(2); break;
    case 3: CWROWS(3); break;
    default: CWROWS(0); break;
  }
#undef CWROWS
}

// A species chain (chain >= 1): the scalar chain of the same arithmetic as the
// flow block's, for species NF + chain - 1.
// Segments of a block chain: rows 1..N of the forward elimination split into
// `segs` runs of cw_seg_rows rows, each run owned by a lane (species) or a
// quad of lanes (flow) of the same wave.  A run starts from an estimate of
// the state entering it (its first row taken as if it were the system's
// first), runs the reference's row steps, and hands its exit state on; a run
// whose entry differs from its predecessor's exit runs again from that exit
// until a row reproduces, bit for bit, what the previous pass stored there
// (the chain depends on the past only through that row's stored values, so
// every later row would repeat too).  Run 0 starts from row 0 exactly, so
// the fixed point is the sequential chain's, bit for bit; the recurrences
// contract, so a run of a few dozen rows forgets its estimate and one
// correcting pass usually suffices.  The back substitution is relayed the
// same way from the top.  segs = 1 is the sequential chain.  After relay pass
// p, runs 0 .. p+1 hold the sequential chain (run p+1 has started from run
// p's exact exit), so segs - 1 passes always suffice and the loops' bound of
// segs + 1 passes is never what ends them.
#ifndef PIAFS_CW_RMIN
#define PIAFS_CW_RMIN 128
#endif
#ifndef PIAFS_CW_LANES
#define PIAFS_CW_LANES 65536
#endif
__host__ __device__ inline int cw_segs(const SweepGeom &S, int nv, int nf) {
  const long lanes = 4L * S.nlines * (4 + nv - nf);  // lanes of the sweep's chains unsegmented
  int s = 1;
  while (s < 16 && 2 * s * PIAFS_CW_RMIN <= S.N + 1 && 2 * s * lanes <= PIAFS_CW_LANES) s <<= 1;
  return s;
}
__host__ __device__ inline int cw_seg_rows(int N, int segs) { return (N + segs - 1) / segs; }

__device__ __forceinline__ void cwc_species(const CaseDev &C, const SweepGeom &S, int dir, int l, int sys,
                                           int chain, int NF, int seg, int segs) {
  const int nv = C.G.nv, nl = S.nlines, N = S.N;
  const long vs = (long)nl * (N + 1);
  const long bs = 4L * nl * (N + 1), bb = (long)sys * (N + 1) * nl + l;
  const double *rows = C.cwr[dir];
  double *back = C.cwb[dir];
  double *out = (dir ? C.iy : C.ix)[sys];
  auto rf = [&](int f, int k) { return rows[(long)f * vs + (long)k * nl + l]; };
  auto bk = [&](int f, int k) -> double & { return back[(long)f * bs + bb + (long)k * nl]; };
  auto coef = [&](int v, int which, int k) { return rf(NF * NF + (sys * nv + v) * 4 + which, k); };
  const int m = NF + chain - 1, f0 = 2 * NF * NF + NF + 3 * (chain - 1);
  const int R = cw_seg_rows(N, segs);
  const int k0 = 1 + seg * R, k1 = min(N + 1, k0 + R);  // forward rows of this run
  const bool has = k0 < k1;
  const int lane = threadIdx.x & 63;
  // forward: rows k0..k1-1, row k storing (1/b'_{k-1}, c_{k-1}, x'_{k-1}) at k-1
  double bp, xp, cprev;
  {
    const int ks = seg == 0 ? 0 : (has ? k0 - 1 : 0);
    bp = coef(m, 1, ks);
    xp = coef(m, 3, ks);
    cprev = coef(m, 2, ks);
  }
  // one pass from the entry (bp, xp); early: stop where a row repeats its stored values
  auto forward = [&](bool early) {
    double a = has ? coef(m, 0, k0) : 0.0, b = has ? coef(m, 1, k0) : 0.0, c = has ? coef(m, 2, k0) : 0.0,
           F = has ? coef(m, 3, k0) : 0.0;
    // what the previous pass stored at k-1, loaded one row ahead like the coefficients
    double srb = early && has ? bk(f0, k0 - 1) : 0.0, sxp = early && has ? bk(f0 + 2, k0 - 1) : 0.0;
    for (int k = k0; k < k1; k++) {
      const double na = k + 1 < k1 ? coef(m, 0, k + 1) : 0.0, nb = k + 1 < k1 ? coef(m, 1, k + 1) : 0.0,
                   nc = k + 1 < k1 ? coef(m, 2, k + 1) : 0.0, nF = k + 1 < k1 ? coef(m, 3, k + 1) : 0.0;
      const double crb = srb, cxp = sxp;
      if (early && k + 1 < k1) {
        srb = bk(f0, k);
        sxp = bk(f0 + 2, k);
      }
      const double rb = 1.0 / bp, fac = a * rb;
      if (early && __double_as_longlong(rb) == __double_as_longlong(crb) &&
          __double_as_longlong(xp) == __double_as_longlong(cxp))
        return false;  // this row and every later one as stored
      bk(f0, k - 1) = rb;
      bk(f0 + 1, k - 1) = cprev;
      bk(f0 + 2, k - 1) = xp;
      bp = b - fac * cprev;
      xp = F - fac * xp;
      cprev = c;
      a = na; b = nb; c = nc; F = nF;
    }
    return true;
  };
  forward(false);
  double ebp = bp, exp_ = xp;  // exit state
  double nbp = bp, nxp = xp;   // the entry this run last started from (set below)
  {
    // the entry of the first pass, for the comparison with the predecessor's exit
    const int ks = seg == 0 ? 0 : (has ? k0 - 1 : 0);
    nbp = coef(m, 1, ks);
    nxp = coef(m, 3, ks);
    if (seg == 0) {  // run 0 starts at row 0: b'_0 = b_0, x'_0 = F_0
      nbp = coef(m, 1, 0);
      nxp = coef(m, 3, 0);
    }
  }
  for (int pass = 0;; pass++) {
    const double pbp = __shfl(ebp, lane - 1), pxp = __shfl(exp_, lane - 1);
    const bool changed = seg > 0 && (__double_as_longlong(pbp) != __double_as_longlong(nbp) ||
                                     __double_as_longlong(pxp) != __double_as_longlong(nxp));
    if (!__any(changed) || pass > segs) break;
    if (changed) {
      nbp = pbp;
      nxp = pxp;
      if (!has) {
        ebp = pbp;
        exp_ = pxp;
      } else {
        bp = pbp;
        xp = pxp;
        cprev = coef(m, 2, k0 - 1);
        if (forward(true)) {
          ebp = bp;
          exp_ = xp;
        }
      }
    }
  }
  // back substitution: rows k1-2 .. k0-1 of this run (the top run also row N),
  // entering from the run above
  const long ob = ((long)m * nl + l) * (N + 1);
  const bool top = k1 == N + 1 && (has || N == 0);
  double xin = 0.0;  // x of the row above this run's rows
  if (top) {
    xin = (1.0 / ebp) * exp_;
    out[ob + N] = xin;
  }
  auto backward = [&](double x, bool early) {
    int k = k1 - 2;
    if (k < k0 - 1) return;
    // each row's values loaded one row ahead
    double nrb = bk(f0, k), ncp = bk(f0 + 1, k), nxq = bk(f0 + 2, k), nso = early ? out[ob + k] : 0.0;
    for (; k >= k0 - 1; k--) {
      const double rb = nrb, c = ncp, xq = nxq, so = nso;
      if (k > k0 - 1) {
        nrb = bk(f0, k - 1);
        ncp = bk(f0 + 1, k - 1);
        nxq = bk(f0 + 2, k - 1);
        if (early) nso = out[ob + k - 1];
      }
      x = rb * (xq - c * x);
      if (early && __double_as_longlong(x) == __double_as_longlong(so)) return;
      out[ob + k] = x;
    }
  };
  double xs = xin;  // the entry the last pass started from
  if (has) backward(xin, false);
  for (int pass = 0;; pass++) {
    // the exit of a run is the x of its lowest row; an empty run passes its entry on
    const double mine = has ? out[ob + k0 - 1] : xs;
    const double above = __shfl(mine, lane + 1);
    const bool last = seg == segs - 1 || top;
    const bool changed = !last && __double_as_longlong(above) != __double_as_longlong(xs);
    if (!__any(changed) || pass > segs) break;
    if (changed) {
      xs = above;
      if (has) backward(above, true);
    }
  }
}

// The flow block chain of one system of one line (blocktridiagLU.c:159-301
// with the blocks inverted by _MatrixInvert_, matops.h:96-130: Gauss-Jordan
// without pivoting) on the four lanes of a quad: lane r < NF holds row r of
// every NF x NF block and row r of the vectors (x', the solution); lane 3 in
// 1D only takes part in the exchanges.  Rows move between the lanes by DPP
// quad permutations (no LDS), and each lane keeps whole copies of the blocks
// it multiplies by (the inverse, C_{k-1}).  Every entry goes through the
// reference's operations in its order: a step of the inverse updates only
// rows that the rows it reads do not depend on.  All four lanes execute every
// exchange (none sits under a lane-dependent branch).
template <int SRC>
__device__ __forceinline__ double quad_get(double v) {
  constexpr int ctrl = SRC | (SRC << 2) | (SRC << 4) | (SRC << 6);  // quad_perm [SRC, SRC, SRC, SRC]
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffLL), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)((unsigned long long)b >> 32), ctrl, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double qget(double v, int src) {  // src is a constant after unrolling
  switch (src) {
    case 0: return quad_get<0>(v);
    case 1: return quad_get<1>(v);
    case 2: return quad_get<2>(v);
    default: return quad_get<3>(v);
  }
}

// inverse of the block whose row ri this lane holds in bp; every lane gets the
// whole inverse.  Its divisions (elimination factors and back substitution)
// take wdiv: the IEEE quotient's bits for the blocks' operands, except that an
// exactly zero numerator -0 over a positive pivot gives +0; the inverse's
// entries are never divisors, so no nonzero value downstream changes.
#ifndef PIAFS_CW_FAST_INV
#define PIAFS_CW_FAST_INV 1
#endif
#if PIAFS_CW_FAST_INV
#define CW_DIV(n, d) wdiv(n, d)
#else
#define CW_DIV(n, d) ((n) / (d))
#endif
template <int NF>
__device__ __forceinline__ void cw_inv_quad(const double (&bp)[NF], int ri, double (&Bi)[NF][NF]) {
  double ac[NF], bb[NF];
#pragma unroll
  for (int j = 0; j < NF; j++) {
    ac[j] = bp[j];
    bb[j] = (ri == j) ? 1.0 : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NF - 1; i++) {
    double pa[NF], pb[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
      pa[k] = qget(ac[k], i);
      pb[k] = qget(bb[k], i);
    }
    if (ri > i) {
      const double f = CW_DIV(ac[i], pa[i]);
#pragma unroll
      for (int k = i + 1; k < NF; k++) ac[k] -= (f * pa[k]);
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (k < ri) bb[k] -= (f * pb[k]);
    }
  }
#pragma unroll
  for (int i = NF - 1; i >= 0; i--) {
    if (ri == i) {
#pragma unroll
      for (int k = 0; k < NF; k++) {
        double sum = 0.0;
#pragma unroll
        for (int j = i + 1; j < NF; j++) sum += (ac[j] * Bi[j][k]);
        bb[k] = CW_DIV(bb[k] - sum, ac[i]);
      }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) Bi[i][k] = qget(bb[k], i);
  }
}

// all four lanes of the quad hold p
__device__ __forceinline__ bool quad_all(bool p) {
  const double d = p ? 1.0 : 0.0;
  return quad_get<0>(d) * quad_get<1>(d) * quad_get<2>(d) * quad_get<3>(d) == 1.0;
}

template <int NF>
__device__ __forceinline__ void cwc_flow_quad(const CaseDev &C, const SweepGeom &S, int dir, int l, int sys, int e,
                                              int seg, int segs) {
  const int nv = C.G.nv, nl = S.nlines, N = S.N;
  const bool act = e < NF;
  const int ri = act ? e : NF - 1;  // the spare lane of a 1D quad reads row NF-1 and stores nothing
  const long vs = (long)nl * (N + 1);
  const long bs = 4L * nl * (N + 1), bb = (long)sys * (N + 1) * nl + l;
  const double *rows = C.cwr[dir];
  double *back = C.cwb[dir];
  double *out = (dir ? C.iy : C.ix)[sys];
  auto rf = [&](int f, int k) { return rows[(long)f * vs + (long)k * nl + l]; };
  auto bk = [&](int f, int k) -> double & { return back[(long)f * bs + bb + (long)k * nl]; };
  auto coef = [&](int which, int k) { return rf(NF * NF + (sys * nv + ri) * 4 + which, k); };
  const int R = cw_seg_rows(N, segs);
  const int k0 = 1 + seg * R, k1 = min(N + 1, k0 + R);  // forward rows of this run
  const bool has = k0 < k1;
  const int lane = threadIdx.x & 63;
  // this lane's row of the blocks of row k: A = a L, B = b L, C = c L, and F
  auto row = [&](int k, double (&A)[NF], double (&Bm)[NF], double (&Cm)[NF], double &F) {
    const double a = coef(0, k), b = coef(1, k), c = coef(2, k);
    F = coef(3, k);
#pragma unroll
    for (int j = 0; j < NF; j++) {
      const double Lj = rf(NF * ri + j, k);
      A[j] = a * Lj;
      Bm[j] = b * Lj;
      Cm[j] = c * Lj;
    }
  };
  // state after row k-1: row ri of B', row ri of C, component ri of x'
  double bp[NF], cprev[NF], xp;
  auto first = [&](int k) {  // row k as the first row of the chain
    double A[NF], Bm[NF], Cm[NF], F;
    row(k, A, Bm, Cm, F);
    xp = F;
#pragma unroll
    for (int j = 0; j < NF; j++) {
      bp[j] = Bm[j];
      cprev[j] = Cm[j];
    }
  };
  // one pass over rows k0..k1-1 (row k stores the inverse of B'_{k-1}, C_{k-1}
  // and x'_{k-1} at k-1); early: stop where a row repeats what is stored
  auto forward = [&](bool early) {
    // row k+1's coefficients and what the previous pass stored at k are loaded one row ahead
    double na = 0.0, nb = 0.0, nc = 0.0, nF = 0.0, nL[NF], sBi[NF], sxp = 0.0;
#pragma unroll
    for (int j = 0; j < NF; j++) nL[j] = sBi[j] = 0.0;
    auto load = [&](int k) {
      na = coef(0, k);
      nb = coef(1, k);
      nc = coef(2, k);
      nF = coef(3, k);
#pragma unroll
      for (int j = 0; j < NF; j++) nL[j] = rf(NF * ri + j, k);
    };
    auto load_stored = [&](int k) {
#pragma unroll
      for (int j = 0; j < NF; j++) sBi[j] = bk(NF * ri + j, k);
      sxp = bk(2 * NF * NF + ri, k);
    };
    if (has) load(k0);
    if (early && has) load_stored(k0 - 1);
    for (int k = k0; k < k1; k++) {
      double A[NF], Bm[NF], Cm[NF], cBi[NF];
      const double F = nF, cxp = sxp;
#pragma unroll
      for (int j = 0; j < NF; j++) {
        A[j] = na * nL[j];
        Bm[j] = nb * nL[j];
        Cm[j] = nc * nL[j];
        cBi[j] = sBi[j];
      }
      if (k + 1 < k1) {
        load(k + 1);
        if (early) load_stored(k);
      }
      double Bi[NF][NF];
      cw_inv_quad<NF>(bp, ri, Bi);
      if (early) {
        bool same = __double_as_longlong(xp) == __double_as_longlong(cxp);
#pragma unroll
        for (int j = 0; j < NF; j++) same = same && __double_as_longlong(Bi[ri][j]) == __double_as_longlong(cBi[j]);
        if (quad_all(same || !act)) return false;  // this row and every later one as stored
      }
      double fac[NF];
#pragma unroll
      for (int j = 0; j < NF; j++) {
        double acc = 0;
#pragma unroll
        for (int q = 0; q < NF; q++) acc += (A[q] * Bi[q][j]);
        fac[j] = acc;
      }
      if (act) {
#pragma unroll
        for (int j = 0; j < NF; j++) {
          bk(NF * ri + j, k - 1) = Bi[ri][j];
          bk(NF * NF + NF * ri + j, k - 1) = cprev[j];
        }
        bk(2 * NF * NF + ri, k - 1) = xp;
      }
      double cpf[NF][NF], xpf[NF];
#pragma unroll
      for (int q = 0; q < NF; q++) {
        xpf[q] = qget(xp, q);
#pragma unroll
        for (int j = 0; j < NF; j++) cpf[q][j] = qget(cprev[j], q);
      }
      double xn = F;
#pragma unroll
      for (int j = 0; j < NF; j++) {
        double bn = Bm[j];
#pragma unroll
        for (int q = 0; q < NF; q++) bn -= (fac[q] * cpf[q][j]);
        bp[j] = bn;
      }
#pragma unroll
      for (int q = 0; q < NF; q++) xn -= (fac[q] * xpf[q]);
      xp = xn;
#pragma unroll
      for (int j = 0; j < NF; j++) cprev[j] = Cm[j];
    }
    return true;
  };
  // run 0 starts at row 0 exactly; the others from their first row's predecessor taken as a first row
  first(seg == 0 || !has ? 0 : k0 - 1);
  double nbp[NF], nxp = xp;  // the entry this run last started from
#pragma unroll
  for (int j = 0; j < NF; j++) nbp[j] = bp[j];
  if (has) forward(false);
  double ebp[NF], exq = xp;  // exit state
#pragma unroll
  for (int j = 0; j < NF; j++) ebp[j] = bp[j];
  for (int pass = 0;; pass++) {
    double pbp[NF];
    bool changed = false;
#pragma unroll
    for (int j = 0; j < NF; j++) {
      pbp[j] = __shfl(ebp[j], lane - 4);
      changed = changed || __double_as_longlong(pbp[j]) != __double_as_longlong(nbp[j]);
    }
    const double pxp = __shfl(exq, lane - 4);
    changed = seg > 0 && (changed || __double_as_longlong(pxp) != __double_as_longlong(nxp));
    changed = !quad_all(!changed);  // a quad moves as one
    if (!__any(changed) || pass > segs) break;
    if (changed) {
#pragma unroll
      for (int j = 0; j < NF; j++) nbp[j] = bp[j] = pbp[j];
      nxp = xp = pxp;
      if (!has) {
#pragma unroll
        for (int j = 0; j < NF; j++) ebp[j] = pbp[j];
        exq = pxp;
      } else {
        double A[NF], Bm[NF], Cm[NF], F;
        row(k0 - 1, A, Bm, Cm, F);
#pragma unroll
        for (int j = 0; j < NF; j++) cprev[j] = Cm[j];
        if (forward(true)) {
#pragma unroll
          for (int j = 0; j < NF; j++) ebp[j] = bp[j];
          exq = xp;
        }
      }
    }
  }
  // back substitution x_k = binv_k (x'_k - C_k x_{k+1}): rows k1-2 .. k0-1 of
  // this run, the top run also row N, entering from the run above
  const long ob = ((long)ri * nl + l) * (N + 1);
  const bool top = k1 == N + 1 && (has || N == 0);
  double xin = 0.0;
  if (top) {
    double Bi[NF][NF], xpf[NF];
#pragma unroll
    for (int j = 0; j < NF; j++) bp[j] = ebp[j];
    cw_inv_quad<NF>(bp, ri, Bi);
#pragma unroll
    for (int q = 0; q < NF; q++) xpf[q] = qget(exq, q);
    double x = 0;
#pragma unroll
    for (int j = 0; j < NF; j++) x += (Bi[ri][j] * xpf[j]);
    if (act) out[ob + N] = x;
    xin = x;
  }
  auto backward = [&](double x, bool early) {
    int k = k1 - 2;
    if (k < k0 - 1) return;
    // each row's values loaded one row ahead
    double nbi[NF], ncp[NF], nxq = 0.0, nso = 0.0;
    auto load_back = [&](int k) {
#pragma unroll
      for (int j = 0; j < NF; j++) {
        nbi[j] = bk(NF * ri + j, k);
        ncp[j] = bk(NF * NF + NF * ri + j, k);
      }
      nxq = bk(2 * NF * NF + ri, k);
      if (early) nso = out[ob + k];
    };
    load_back(k);
    for (; k >= k0 - 1; k--) {
      double bi[NF], cp[NF], xf[NF], yf[NF];
      const double so = nso;
#pragma unroll
      for (int j = 0; j < NF; j++) {
        bi[j] = nbi[j];
        cp[j] = ncp[j];
        xf[j] = qget(x, j);
      }
      double y = nxq;
      if (k > k0 - 1) load_back(k - 1);
#pragma unroll
      for (int j = 0; j < NF; j++) y -= (cp[j] * xf[j]);
#pragma unroll
      for (int j = 0; j < NF; j++) yf[j] = qget(y, j);
      double xn = 0;
#pragma unroll
      for (int j = 0; j < NF; j++) xn += (bi[j] * yf[j]);
      x = xn;
      if (early && quad_all(!act || __double_as_longlong(x) == __double_as_longlong(so))) return;
      if (act) out[ob + k] = x;
    }
  };
  double xs = xin;  // the entry the last pass started from
  if (has) backward(xin, false);
  for (int pass = 0;; pass++) {
    const double mine = has ? out[ob + k0 - 1] : xs;  // this run's exit: the x of its lowest row
    const double above = __shfl(mine, lane + 4);
    const bool last = seg == segs - 1 || top;
    bool changed = !last && __double_as_longlong(above) != __double_as_longlong(xs);
    changed = !quad_all(!changed);
    if (!__any(changed) || pass > segs) break;
    if (changed) {
      xs = above;
      if (has) backward(above, true);
    }
  }
}

// Work items of a sweep: the flow chains first (a quad of lanes per run of a
// system; (run, system, line) with the run fastest), then one item per run of
// each species chain ((run, chain, system, line), the run fastest), so that
// the runs of a chain are neighbouring lanes of one wave.  The y sweep starts
// on a 64-item boundary, which keeps every quad of a flow chain on one quad
// of a wave.
__host__ __device__ inline int cw_flow_lanes(int) { return 4; }
__host__ __device__ inline long cw_items(const SweepGeom &S, int nv, int nf) {
  return 4L * S.nlines * cw_segs(S, nv, nf) * (cw_flow_lanes(nf) + nv - nf);
}
__host__ __device__ inline long cw_x_span(const CaseDev &c, int nf) {
  const long nX = cw_items(c.Sx, c.G.nv, nf);
  return c.G.ndims == 2 ? (nX + 63) & ~63L : nX;
}

#ifndef PIAFS_CWSOLVE_MINB
#define PIAFS_CWSOLVE_MINB 1
#endif
__global__ void __launch_bounds__(256, PIAFS_CWSOLVE_MINB) k_cw_solve(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const int nv = C.G.nv, nf = C.model == 1 ? 3 : 4;
  const long nX = cw_items(C.Sx, nv, nf), nXs = cw_x_span(C, nf);
  const long nY = (C.G.ndims == 2) ? cw_items(C.Sy, nv, nf) : 0;
  if (t >= nXs + nY || (t >= nX && t < nXs)) return;
  const int dir = t < nXs ? 0 : 1;
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const long u = dir ? t - nXs : t;
  const int segs = cw_segs(S, nv, nf);
  const long nl = S.nlines, nflow = 4L * nl * cw_flow_lanes(nf) * segs;
  if (u < nflow) {
    const long quad = u >> 2, grp = quad / segs;
    const int seg = (int)(quad % segs);
    if (nf == 4) cwc_flow_quad<4>(C, S, dir, (int)(grp % nl), (int)(grp / nl), (int)(u & 3), seg, segs);
    else cwc_flow_quad<3>(C, S, dir, (int)(grp % nl), (int)(grp / nl), (int)(u & 3), seg, segs);
    return;
  }
  const long v = u - nflow, per = 4L * nl;
  const long ch = v / segs;
  const int seg = (int)(v % segs);
  const int chain = 1 + (int)(ch / per), sys = (int)((ch % per) / nl), l = (int)(ch % nl);
  cwc_species(C, S, dir, l, sys, chain, nf, seg, segs);
}

// Blocks of a sweep for the team-based reconstruction (each block holds
// 256/team whole teams; a block never mixes sweeps).
__host__ __device__ inline int recon_sweep_blocks(const SweepGeom &S, int nv, int paired) {
  const long teams_per_block = 256 / S.team;
  const long nsys = recon_systems(S, nv, paired);
  return (int)((nsys + teams_per_block - 1) / teams_per_block);
}

// ---------------------------------------------------------------------------
// Component-wise reconstruction of lines longer than one team covers
// (SweepGeom::longl).  k_lrows assembles every row of every system, end rows
// included, one thread per (system, interface), into the sweep's scratch,
// system by system ([A|B|C|X][system][interface]).  k_lsolve gives each pair
// of systems (uL and uR, or fL and fR of one line and variable) a block, whose
// 256 lanes walk the line window by window (kMaxRows rows per lane, 4096 rows
// a window): the forward elimination of each window (team_forward) starts
// from the exact pivot, super-diagonal and x' of the previous window's last
// row and stores b' and x'; the back substitution (team_backward) then runs
// the windows from the last, each from the exact solution of the row after
// it.  Every row goes through the reference's arithmetic (tridiagLU.c:133-150,
// :243-255, over rows 0..N with identity end rows), so the result is the
// sequential sweep's bit for bit.  The block then writes the interface arrays
// as the team kernels do.
// ---------------------------------------------------------------------------
__host__ __device__ inline long long_rows(const SweepGeom &S, int nv) {
  return S.longl ? (long)(S.N + 1) * S.nlines * nv * 4 : 0;
}

template <int CR, int KIND>
__device__ __forceinline__ void lrow_at(const CaseDev &C, int dir, long t) {
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const int nv = C.G.nv, N = S.N;
  const long ns4 = (long)S.nlines * nv * 4;
  const long sys = t / (N + 1);
  const int k = (int)(t % (N + 1));
  const int r = (int)(sys % 4), v = (int)((sys / 4) % nv), l = (int)(sys / 4 / nv);
  const double *src = (r < 2 ? (dir ? C.UT : C.U) : (dir ? C.flux_y : C.flux_x)) + (long)v * S.plane + S.base +
                      (long)l * S.line_stride;
  double a, b, c, x;
  row_coeffs<CR, KIND>(k, N, !(r & 1), C.R, src[k - 3], src[k - 2], src[k - 1], src[k], src[k + 1], src[k + 2], a,
                       b, c, x);
  double *L = C.lrow[dir];
  const long f = (long)(N + 1) * ns4, o = sys * (N + 1) + k;
  L[o] = a;
  L[f + o] = b;
  L[2 * f + o] = c;
  L[3 * f + o] = x;
}

__global__ void __launch_bounds__(256) k_lrows(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const long nX = long_rows(C.Sx, C.G.nv), nY = C.G.ndims == 2 ? long_rows(C.Sy, C.G.nv) : 0;
  if (t >= nX + nY) return;
  const int dir = t < nX ? 0 : 1;
  const long tt = dir ? t - nX : t;
  const int cr = C.R.crweno;
  switch (C.R.kind) {
    case 1: cr ? lrow_at<1, 1>(C, dir, tt) : lrow_at<0, 1>(C, dir, tt); break;
    case 2: cr ? lrow_at<1, 2>(C, dir, tt) : lrow_at<0, 2>(C, dir, tt); break;
    case 3: cr ? lrow_at<1, 3>(C, dir, tt) : lrow_at<0, 3>(C, dir, tt); break;
    case 4: lrow_at<0, 4>(C, dir, tt); break;
    default: cr ? lrow_at<1, 0>(C, dir, tt) : lrow_at<0, 0>(C, dir, tt); break;
  }
}

// pairs of systems of the long sweeps, one block each
__host__ __device__ inline long long_pairs(const SweepGeom &S, int nv) {
  return S.longl ? (long)S.nlines * nv * 2 : 0;
}

// One system of a long line, rows 0..N of the scratch (A, B, C, X at
// interface k: L[arr*f + k]): rows 1..N-1 solved in windows of kLongWin
// rows, the identity rows 0 and N holding the end values.  On exit X holds
// the solution at every row.
constexpr int kLongTeam = 256, kLongWin = kLongTeam * kMaxRows;
__device__ void long_system(double *A, double *B, double *Cc, double *X, int N, double *red, double *carry,
                            bool plain_start, bool &zp) {
  const int lane = threadIdx.x;
  const int nwin = (N - 1 + kLongWin - 1) / kLongWin;  // interior rows 1..N-1
  auto window_rows = [&](int w, int &lo, int &n) {
    lo = 1 + w * kLongWin + lane * kMaxRows;
    n = N - lo;
    n = n < 0 ? 0 : (n > kMaxRows ? kMaxRows : n);
  };
  // forward: window w enters with the exact b', c and x' of the row before it
  double bfirst = 1.0, cfirst = 0.0, xfirst = X[0];  // row 0: identity, x' = its value
  for (int w = 0; w < nwin; w++) {
    int lo, n;
    window_rows(w, lo, n);
    double a[kMaxRows], b[kMaxRows], c[kMaxRows], x[kMaxRows], fk[kMaxRows];
#pragma unroll
    for (int i = 0; i < kMaxRows; i++) {
      const bool in = i < n;
      a[i] = in ? A[lo + i] : 0.0;
      b[i] = in ? B[lo + i] : 1.0;
      c[i] = in ? Cc[lo + i] : 0.0;
      x[i] = in ? X[lo + i] : 0.0;
    }
    team_forward<kMaxRows>(a, b, c, x, fk, n, lane, kLongTeam, bfirst, cfirst, xfirst, red, plain_start, zp);
#pragma unroll
    for (int i = 0; i < kMaxRows; i++)
      if (i < n) {
        B[lo + i] = b[i];
        X[lo + i] = x[i];
      }
    // the window's last row, from the lane that holds it, for the next window
    const int last = (w + 1) * kLongWin < N - 1 ? (w + 1) * kLongWin : N - 1;  // row index
    if (n > 0 && lo + n - 1 == last) {
      carry[0] = b[n - 1];
      carry[1] = c[n - 1];
      carry[2] = x[n - 1];
    }
    __syncthreads();
    bfirst = carry[0];
    cfirst = carry[1];
    xfirst = carry[2];
    __syncthreads();
  }
  // back substitution: window w ends at the row whose solution the window
  // after it (or row N, an identity row) has fixed
  double xlast = X[N];
  for (int w = nwin - 1; w >= 0; w--) {
    int lo, n;
    window_rows(w, lo, n);
    double b[kMaxRows], c[kMaxRows], x[kMaxRows], wk[kMaxRows];
#pragma unroll
    for (int i = 0; i < kMaxRows; i++) {
      const bool in = i < n;
      b[i] = in ? B[lo + i] : 1.0;
      c[i] = in ? Cc[lo + i] : 0.0;
      x[i] = in ? X[lo + i] : 0.0;
    }
    const int last = (w + 1) * kLongWin < N - 1 ? (w + 1) * kLongWin : N - 1;
    const bool is_last = n > 0 && lo + n - 1 == last;
    team_backward<kMaxRows>(b, c, x, wk, n, lane, kLongTeam, is_last, xlast, red, plain_start);
#pragma unroll
    for (int i = 0; i < kMaxRows; i++)
      if (i < n) X[lo + i] = x[i];
    if (lane == 0) carry[0] = x[0];  // the window's first row
    __syncthreads();
    xlast = carry[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kLongTeam) k_lsolve(CASE_ARGS) {
  CASE_PREAMBLE
  (void)t;
  if (stage >= C.nstages) return;
  __shared__ double red[16];
  __shared__ double carry[3];
  const int nv = C.G.nv;
  const long nX = long_pairs(C.Sx, nv);
  const long nY = C.G.ndims == 2 ? long_pairs(C.Sy, nv) : 0;
  const long pair = blockIdx.x - firstblock[cidx];
  if (pair >= nX + nY) return;  // uniform over the block
  const int dir = pair < nX ? 0 : 1;
  const long p = dir ? pair - nX : pair;  // (line*nv + var)*2 + q
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const int N = S.N, q = (int)(p % 2);
  const long lv = p / 2;  // line*nv + var
  const int v = (int)(lv % nv), l = (int)(lv / nv);
  const long ns4 = (long)S.nlines * nv * 4, f = (long)(N + 1) * ns4;
  double *L = C.lrow[dir];
  const long s0 = lv * 4 + 2 * q;  // the left-biased system; s0 + 1 the right-biased one
  if (C.R.crweno) {
    bool zp = false;
    for (int j = 0; j < 2; j++) {
      const long o = (s0 + j) * (N + 1);
      long_system(L + o, L + f + o, L + 2 * f + o, L + 3 * f + o, N, red, carry, C.R.plain_start != 0, zp);
    }
    if (zp) atomicOr(C.status, 1);
  }
  // interface arrays: split mode uL, uR (q 0) or fL, fR (q 1); paired mode
  // 0.5*(uR - uL) or 0.5*(fL + fR)
  double *const *io = dir ? C.iy : C.ix;
  const long line = ((long)v * S.nlines + l) * (N + 1);
  const double *xl = L + 3 * f + s0 * (N + 1), *xr = xl + (N + 1);
  for (int k = threadIdx.x; k <= N; k += kLongTeam) {
    if (C.paired) io[q][line + k] = q ? 0.5 * (xl[k] + xr[k]) : 0.5 * (xr[k] - xl[k]);
    else {
      io[2 * q][line + k] = xl[k];
      io[2 * q + 1][line + k] = xr[k];
    }
  }
}

__global__ void __launch_bounds__(256) k_recon_ends(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const long nsx = recon_systems(C.Sx, C.G.nv, C.paired);
  const long nsy = (C.G.ndims == 2) ? recon_systems(C.Sy, C.G.nv, C.paired) : 0;
  if (t >= 2 * (nsx + nsy)) return;
  const int dir = t < 2 * nsx ? 0 : 1;
  const long q = dir ? t - 2 * nsx : t;
  const long sys = q >> 1;
  const int e = (int)(q & 1);
  const int cr = C.R.crweno;
  switch (C.R.kind) {
    case 1: cr ? recon_end<1, 1>(C, dir, sys, e) : recon_end<0, 1>(C, dir, sys, e); break;
    case 2: cr ? recon_end<1, 2>(C, dir, sys, e) : recon_end<0, 2>(C, dir, sys, e); break;
    case 3: cr ? recon_end<1, 3>(C, dir, sys, e) : recon_end<0, 3>(C, dir, sys, e); break;
    case 4: recon_end<0, 4>(C, dir, sys, e); break;
    default: cr ? recon_end<1, 0>(C, dir, sys, e) : recon_end<0, 0>(C, dir, sys, e); break;
  }
}

// FULL: the kernel of class K_RECONF, for paired cases whose lanes all hold
// kMaxRows rows (recon_pair_lds<..., true>); the other cases run in K_RECON.
template <bool FULL>
__device__ __forceinline__ void recon_body(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  __shared__ double red[16];
  __shared__ double stash[kMaxRows * 256];
  const int bid = (int)(blockIdx.x - firstblock[cidx]);
  const int nbx = recon_sweep_blocks(C.Sx, C.G.nv, C.paired);
  const int dir = bid < nbx ? 0 : 1;
  const int lb = dir ? bid - nbx : bid;
  const SweepGeom &S = dir ? C.Sy : C.Sx;
  const int T = S.team;
  const int lane = threadIdx.x & (T - 1);
  const long sys = (long)lb * (256 / T) + threadIdx.x / T;
  const bool active = sys < recon_systems(S, C.G.nv, C.paired);
  const int cr = C.R.crweno;
#if PIAFS_RECON_LDS
  if (FULL || C.paired) {
    __shared__ double lines[kLineLds];
    switch (C.R.kind) {
      case 1: cr ? recon_pair_lds<1, 1, FULL>(C, dir, sys, active, lane, red, stash, lines) : recon_pair_lds<0, 1, FULL>(C, dir, sys, active, lane, red, stash, lines); break;
      case 2: cr ? recon_pair_lds<1, 2, FULL>(C, dir, sys, active, lane, red, stash, lines) : recon_pair_lds<0, 2, FULL>(C, dir, sys, active, lane, red, stash, lines); break;
      case 3: cr ? recon_pair_lds<1, 3, FULL>(C, dir, sys, active, lane, red, stash, lines) : recon_pair_lds<0, 3, FULL>(C, dir, sys, active, lane, red, stash, lines); break;
      case 4: recon_pair_lds<0, 4, FULL>(C, dir, sys, active, lane, red, stash, lines); break;
      default: cr ? recon_pair_lds<1, 0, FULL>(C, dir, sys, active, lane, red, stash, lines) : recon_pair_lds<0, 0, FULL>(C, dir, sys, active, lane, red, stash, lines); break;
    }
    return;
  }
#endif
  if (FULL) return;  // not reached: K_RECONF holds paired cases only
  switch (C.R.kind) {
    case 1: cr ? recon_team<1, 1>(C, dir, sys, active, lane, red, stash) : recon_team<0, 1>(C, dir, sys, active, lane, red, stash); break;
    case 2: cr ? recon_team<1, 2>(C, dir, sys, active, lane, red, stash) : recon_team<0, 2>(C, dir, sys, active, lane, red, stash); break;
    case 3: cr ? recon_team<1, 3>(C, dir, sys, active, lane, red, stash) : recon_team<0, 3>(C, dir, sys, active, lane, red, stash); break;
    case 4: recon_team<0, 4>(C, dir, sys, active, lane, red, stash); break;
    default: cr ? recon_team<1, 0>(C, dir, sys, active, lane, red, stash) : recon_team<0, 0>(C, dir, sys, active, lane, red, stash); break;
  }
}
__global__ void __launch_bounds__(256) k_recon(CASE_ARGS) { recon_body<false>(cases, blockcase, firstblock, stage); }
__global__ void __launch_bounds__(256, PIAFS_RECONF_MINB) k_recon_full(CASE_ARGS) { recon_body<true>(cases, blockcase, firstblock, stage); }

// ---------------------------------------------------------------------------
// Upwinding.
//  NS2D Roe: NavierStokes2DUpwind.c:26-92 with navierstokes2d.h:115-311 --
//   |A| = R |D| L built as the 4x4 flow block (the species block of the
//   reference's dense n x n product is diagonal: R = L = I, D = vn), Harten
//   entropy fix delta = 1e-6.
//  Euler1D Roe: Euler1DUpwind.c:27-93 (no entropy fix); rf-char :107-190.
// ---------------------------------------------------------------------------
// In paired mode io[0] holds 0.5*(uR - uL) and io[1] 0.5*(fL + fR) (the only
// combinations the Roe flux uses); in split mode io[0..3] hold uL uR fL fR.
#define UD(m) (PAIR ? io[0][(m) * vstride + ii] : 0.5 * (io[1][(m) * vstride + ii] - io[0][(m) * vstride + ii]))
#define FS(m) (PAIR ? io[1][(m) * vstride + ii] : 0.5 * (io[2][(m) * vstride + ii] + io[3][(m) * vstride + ii]))

template <int DIR, int PAIR>
__device__ __forceinline__ void upwind_ns2d_roe(const SweepGeom &S, int nv, double g, long t,
                                                const double *__restrict__ q, double *const *io,
                                                double *__restrict__ fI, bool species = true) {
  const int nl = S.nlines, N = S.N;
  const int k = (int)(t % (N + 1)), l = (int)(t / (N + 1));
  const long pl = S.plane;
  const long cL = S.base + (long)l * S.line_stride + (k - 1), cR = cL + 1;
  // Roe average of the cell-centred neighbours, then its eigenstructure
  const double uLc[4] = {q[cL], q[pl + cL], q[2 * pl + cL], q[3 * pl + cL]};
  const double uRc[4] = {q[cR], q[pl + cR], q[2 * pl + cR], q[3 * pl + cR]};
  double ua[4], D[4], L[4][4], R[4][4];
  ns2d_roeavg(g, uLc, uRc, ua);
  ns2d_eigvals<DIR>(g, ua, D);
  ns2d_eigvecs<DIR>(g, ua, L, R);
  const double vn = D[3];
  const double delta = 0.000001, delta2 = delta * delta;
  const double Dsp = (fabs(vn) < delta ? (vn * vn + delta2) / (2 * delta) : fabs(vn));
#pragma unroll
  for (int m = 0; m < 4; m++) D[m] = (fabs(D[m]) < delta ? (D[m] * D[m] + delta2) / (2 * delta) : fabs(D[m]));
  const long ii = (long)l * (N + 1) + k, vstride = (long)nl * (N + 1);
  double ud[4], fo[4];
#pragma unroll
  for (int m = 0; m < 4; m++) ud[m] = UD(m);
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
    fo[i] = FS(i) - us;
  }
#pragma unroll
  for (int m = 0; m < 4; m++) fI[m * vstride + ii] = fo[m];
  if (species)  // else the reconstruction has written them
    for (int m = 4; m < nv; m++) fI[m * vstride + ii] = FS(m) - Dsp * UD(m);
}

// mode: 0 Roe (Euler1DUpwind.c:27-93), 1 rf-char (:107-190), 2 llf-char
// (:199-266), 3 rusanov (:282-353; its wave speed is gamma P/rho + |u|, with
// no square root, as the reference computes it)
template <int PAIR>
__device__ __forceinline__ void upwind_e1d(const SweepGeom &S, int nv, double g, int mode, long t,
                                           const double *__restrict__ q, double *const *io, double *__restrict__ fI) {
  const double *uL = io[0], *uR = io[1], *fL = io[2], *fR = io[3];
  const int nl = S.nlines, N = S.N;
  const int k = (int)(t % (N + 1)), l = (int)(t / (N + 1));
  const long pl = S.plane;
  const long cL = S.base + (long)l * S.line_stride + (k - 1), cR = cL + 1;
  double a0, a1, a2;
  euler1d_roeavg(g, q[cL], q[pl + cL], q[2 * pl + cL], q[cR], q[pl + cR], q[2 * pl + cR], a0, a1, a2);
  const long ii = (long)l * (N + 1) + k, vs = (long)nl * (N + 1), vstride = vs;
  if (mode == 3) {
    auto speed = [&](double rho, double mom, double e) {
      const double uvel = mom / rho, P = (e - 0.5 * rho * uvel * uvel) * (g - 1.0);
      return g * P / rho + fabs(uvel);
    };
    const double aL = speed(q[cL], q[pl + cL], q[2 * pl + cL]), aR = speed(q[cR], q[pl + cR], q[2 * pl + cR]);
    const double aA = speed(a0, a1, a2);
    const double m1 = aL > aR ? aL : aR, m2 = aR > aA ? aR : aA;  // max3, math_ops.h:30
    const double alpha = m1 > m2 ? m1 : m2;
    for (int m = 0; m < nv; m++) fI[m * vs + ii] = FS(m) - alpha * UD(m);
    return;
  }
  Eig3 E;
  euler1d_eig(g, a0, a1, a2, E);
  if (mode == 0) {
    double ud[3], DL[3][3];
    for (int m = 0; m < 3; m++) ud[m] = UD(m);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) DL[i][j] = fabs(E.D[i]) * E.L[i][j];
    for (int i = 0; i < 3; i++) {
      double us = 0.0;
      for (int j = 0; j < 3; j++) {
        const double mA = E.R[i][0] * DL[0][j] + E.R[i][1] * DL[1][j] + E.R[i][2] * DL[2][j];
        us += mA * ud[j];
      }
      fI[i * vs + ii] = FS(i) - us;
    }
    const double Dsp = fabs(E.D[0]);
    for (int m = 3; m < nv; m++) fI[m * vs + ii] = FS(m) - Dsp * UD(m);
    return;
  }
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
    fc[m] = rf_pick(EL.D[m], E.D[m], ER.D[m], ucL, ucR, fcL, fcR, mode == 2);
  }
  for (int i = 0; i < 3; i++) fI[i * vs + ii] = E.R[i][0] * fc[0] + E.R[i][1] * fc[1] + E.R[i][2] * fc[2];
  // species rows: identity eigenvectors, eigenvalue v
  const double eLs = EL.D[0], eCs = E.D[0], eRs = ER.D[0];
  for (int m = 3; m < nv; m++)
    fI[m * vs + ii] =
        rf_pick(eLs, eCs, eRs, uL[m * vs + ii], uR[m * vs + ii], fL[m * vs + ii], fR[m * vs + ii], mode == 2);
}

// NavierStokes2DUpwindRF (NavierStokes2DUpwind.c:106-187): L and R from the
// Roe average of the two reconstructed interface states uL, uR (not of the
// neighbouring cells); a field is upwinded by the signs of its eigenvalue at
// uL, uR and the average.  Species fields have eigenvalue vn and identity
// eigenvectors.  Split mode only: uL, uR, fL, fR in io[0..3].
template <int DIR, bool LLF>
__device__ __forceinline__ void upwind_ns2d_rf(const SweepGeom &S, int nv, double g, long t, double *const *io,
                                               double *__restrict__ fI) {
  const int nl = S.nlines, N = S.N;
  const int k = (int)(t % (N + 1)), l = (int)(t / (N + 1));
  const long ii = (long)l * (N + 1) + k, vs = (long)nl * (N + 1);
  const double *uL = io[0] + ii, *uR = io[1] + ii, *fL = io[2] + ii, *fR = io[3] + ii;
  const double uLk[4] = {uL[0], uL[vs], uL[2 * vs], uL[3 * vs]};
  const double uRk[4] = {uR[0], uR[vs], uR[2 * vs], uR[3 * vs]};
  const double fLk[4] = {fL[0], fL[vs], fL[2 * vs], fL[3 * vs]};
  const double fRk[4] = {fR[0], fR[vs], fR[2 * vs], fR[3 * vs]};
  double ua[4], L[4][4], R[4][4], eL[4], eC[4], eR[4];
  ns2d_roeavg(g, uLk, uRk, ua);
  ns2d_eigvecs<DIR>(g, ua, L, R);
  ns2d_eigvals<DIR>(g, uLk, eL);
  ns2d_eigvals<DIR>(g, uRk, eR);
  ns2d_eigvals<DIR>(g, ua, eC);
  double fc[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    // MatVecMult (matmult_native.h:119-130): row sums in index order
    const double ucL = L[m][0] * uLk[0] + L[m][1] * uLk[1] + L[m][2] * uLk[2] + L[m][3] * uLk[3];
    const double ucR = L[m][0] * uRk[0] + L[m][1] * uRk[1] + L[m][2] * uRk[2] + L[m][3] * uRk[3];
    const double fcL = L[m][0] * fLk[0] + L[m][1] * fLk[1] + L[m][2] * fLk[2] + L[m][3] * fLk[3];
    const double fcR = L[m][0] * fRk[0] + L[m][1] * fRk[1] + L[m][2] * fRk[2] + L[m][3] * fRk[3];
    fc[m] = rf_pick(eL[m], eC[m], eR[m], ucL, ucR, fcL, fcR, LLF);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) fI[i * vs + ii] = R[i][0] * fc[0] + R[i][1] * fc[1] + R[i][2] * fc[2] + R[i][3] * fc[3];
  for (int m = 4; m < nv; m++)
    fI[m * vs + ii] = rf_pick(eL[3], eC[3], eR[3], uL[m * vs], uR[m * vs], fL[m * vs], fR[m * vs], LLF);
}

// NavierStokes2DUpwindRusanov (NavierStokes2DUpwind.c:271-332): the largest of
// c + |v_n| at the two neighbouring cells and their Roe average.
template <int DIR, int PAIR>
__device__ __forceinline__ void upwind_ns2d_rusanov(const SweepGeom &S, int nv, double g, long t,
                                                    const double *__restrict__ q, double *const *io,
                                                    double *__restrict__ fI) {
  const int nl = S.nlines, N = S.N;
  const int k = (int)(t % (N + 1)), l = (int)(t / (N + 1));
  const long pl = S.plane;
  const long cL = S.base + (long)l * S.line_stride + (k - 1), cR = cL + 1;
  const double uLc[4] = {q[cL], q[pl + cL], q[2 * pl + cL], q[3 * pl + cL]};
  const double uRc[4] = {q[cR], q[pl + cR], q[2 * pl + cR], q[3 * pl + cR]};
  double ua[4];
  ns2d_roeavg(g, uLc, uRc, ua);
  auto speed = [&](const double u[4]) {
    double rho, vx, vy, e, P;
    ns2d_flowvar(u, g, rho, vx, vy, e, P);
    return sqrt(g * P / rho) + fabs(DIR == 0 ? vx : vy);
  };
  const double aL = speed(uLc), aR = speed(uRc), aA = speed(ua);
  const double m1 = aL > aR ? aL : aR, m2 = aR > aA ? aR : aA;  // max3, math_ops.h:30
  const double alpha = m1 > m2 ? m1 : m2;
  const long ii = (long)l * (N + 1) + k, vstride = (long)nl * (N + 1);
  for (int m = 0; m < nv; m++) fI[m * vstride + ii] = FS(m) - alpha * UD(m);
}

#undef UD
#undef FS

__global__ void __launch_bounds__(256) k_upwind(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const long nX = (long)(C.Sx.N + 1) * C.Sx.nlines;
  const long nY = (C.G.ndims == 2) ? (long)(C.Sy.N + 1) * C.Sy.nlines : 0;
  if (t >= nX + nY) return;
  if (C.model == 1) {
    if (C.paired) upwind_e1d<1>(C.Sx, C.G.nv, C.P.gamma, C.upwind, t, C.U, C.ix, C.fIx);
    else upwind_e1d<0>(C.Sx, C.G.nv, C.P.gamma, C.upwind, t, C.U, C.ix, C.fIx);
  } else if (C.upwind == 1 || C.upwind == 2) {  // rf-char, llf-char
    const bool llf = C.upwind == 2;
    if (t < nX) {
      if (llf) upwind_ns2d_rf<0, true>(C.Sx, C.G.nv, C.P.gamma, t, C.ix, C.fIx);
      else upwind_ns2d_rf<0, false>(C.Sx, C.G.nv, C.P.gamma, t, C.ix, C.fIx);
    } else {
      if (llf) upwind_ns2d_rf<1, true>(C.Sy, C.G.nv, C.P.gamma, t - nX, C.iy, C.fIy);
      else upwind_ns2d_rf<1, false>(C.Sy, C.G.nv, C.P.gamma, t - nX, C.iy, C.fIy);
    }
  } else if (C.upwind == 3) {  // rusanov
    if (t < nX) {
      if (C.paired) upwind_ns2d_rusanov<0, 1>(C.Sx, C.G.nv, C.P.gamma, t, C.U, C.ix, C.fIx);
      else upwind_ns2d_rusanov<0, 0>(C.Sx, C.G.nv, C.P.gamma, t, C.U, C.ix, C.fIx);
    } else {
      if (C.paired) upwind_ns2d_rusanov<1, 1>(C.Sy, C.G.nv, C.P.gamma, t - nX, C.UT, C.iy, C.fIy);
      else upwind_ns2d_rusanov<1, 0>(C.Sy, C.G.nv, C.P.gamma, t - nX, C.UT, C.iy, C.fIy);
    }
  } else if (t < nX) {
    if (C.paired) upwind_ns2d_roe<0, 1>(C.Sx, C.G.nv, C.P.gamma, t, C.U, C.ix, C.fIx, !C.fuse_sp);
    else upwind_ns2d_roe<0, 0>(C.Sx, C.G.nv, C.P.gamma, t, C.U, C.ix, C.fIx);
  } else {
    if (C.paired) upwind_ns2d_roe<1, 1>(C.Sy, C.G.nv, C.P.gamma, t - nX, C.UT, C.iy, C.fIy, !C.fuse_sp);
    else upwind_ns2d_roe<1, 0>(C.Sy, C.G.nv, C.P.gamma, t - nX, C.UT, C.iy, C.fIy);
  }
}

// Conservation check, boundary part: the interface fluxes on the domain
// boundary summed over the lines in index order (HyperbolicFunction.c:104-110),
// one thread per variable; after the last stage the step's integral
// sum_s dt b_s (stage sums) (TimeRK.c:86-97), weighted by the face elements
// (BoundaryIntegral.c:37-58), is added to the case's total.
__global__ void __launch_bounds__(256) k_bflux(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const int nv = C.G.nv, nd = C.G.ndims;
  if (t >= nv) return;
  const int v = (int)t, nf = 2 * nd * nv;
  double *bf = C.bflux + (long)stage * nf;
  for (int d = 0; d < nd; d++) {
    const SweepGeom &S = d ? C.Sy : C.Sx;
    const double *fI = d ? C.fIy : C.fIx;
    const int N = S.N, nl = S.nlines;
    double lo = 0.0, hi = 0.0;
    for (int l = 0; l < nl; l++) {
      const long a = ((long)v * nl + l) * (N + 1);
      lo -= fI[a];
      hi += fI[a + N];
    }
    bf[(2 * d + 0) * nv + v] = lo;
    bf[(2 * d + 1) * nv + v] = hi;
  }
  if (stage != C.nstages - 1) return;
  double local = 0.0;
  for (int d = 0; d < nd; d++) {
    double step[2];
    for (int f = 0; f < 2; f++) {
      double sb = 0.0;
      for (int s = 0; s < C.nstages; s++) sb += (C.dt * C.b[s]) * C.bflux[(long)s * nf + (2 * d + f) * nv + v];
      step[f] = sb;
    }
    local += step[0] * C.dS[d];
    local += step[1] * C.dS[d];
  }
  C.tbi[v] += 1.0 * local;
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
__device__ __forceinline__ Prim prim_at(long plane, double gam, const double *U, long p) {
  const double rho = U[p];
  const double vx = (rho == 0) ? 0 : U[plane + p] / rho, vy = (rho == 0) ? 0 : U[2 * plane + p] / rho;
  const double e = U[3 * plane + p];
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
__device__ __forceinline__ void deriv_prim(long plane, int g, double gam, const double *U, long p0, long st, int i,
                                           int N, double out[3]) {
  auto P = [&](int k) { return prim_at(plane, gam, U, p0 + (long)k * st); };
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
// dxinv (NavierStokes2DParabolicFunction.c:96-106).  With the derivative
// exchange along a periodic dimension (xchg, MPIExchangeBoundariesnD.c:
// 109-167 as called at NavierStokes2DParabolicFunction.c:91-94), a ghost point
// of an interior line across that dimension takes the unscaled derivatives of
// its periodic image, scaled by its own dxinv; corners are not exchanged.
__global__ void __launch_bounds__(256) k_qderiv(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const Grid &G = C.G;
  if (t >= (long)G.nxg * G.nyg) return;
  const int i = (int)(t % G.nxg) - G.g, j = (int)(t / G.nxg) - G.g;
  const long pl = (long)G.nxg * G.nyg;
  const double gam = C.P.gamma;
  double dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0};
  const bool iin = i >= 0 && i < G.nx, jin = j >= 0 && j < G.ny;
  int si = i, sj = j;  // the point whose derivatives this point holds
  if (C.P.xchg[0] && jin && !iin) si = i < 0 ? i + G.nx : i - G.nx;
  if (C.P.xchg[1] && iin && !jin) sj = j < 0 ? j + G.ny : j - G.ny;
  if (sj >= 0 && sj < G.ny) {
    const long p0 = (long)(sj + G.g) * G.nxg + G.g;  // (0, sj)
    if (C.P.par4) deriv_prim<1>(G.plane, G.g, gam, C.U, p0, 1, si, G.nx, dx);
    else deriv_prim<0>(G.plane, G.g, gam, C.U, p0, 1, si, G.nx, dx);
    const double s = C.dxinv[G.g + i];
    dx[0] *= s; dx[1] *= s; dx[2] *= s;
  }
  if (si >= 0 && si < G.nx) {
    const long p0 = (long)G.g * G.nxg + (si + G.g);  // (si, 0)
    if (C.P.par4) deriv_prim<1>(G.plane, G.g, gam, C.U, p0, G.nxg, sj, G.ny, dy);
    else deriv_prim<0>(G.plane, G.g, gam, C.U, p0, G.nxg, sj, G.ny, dy);
    const double s = C.dxinv[(G.nx + 2 * G.g) + G.g + j];
    dy[0] *= s; dy[1] *= s; dy[2] *= s;
  }
  double *Qd = C.Qd;
  Qd[t] = dx[0]; Qd[pl + t] = dx[1]; Qd[2 * pl + t] = dx[2];
  Qd[3 * pl + t] = dy[0]; Qd[4 * pl + t] = dy[1]; Qd[5 * pl + t] = dy[2];
}

// Viscous fluxes (NavierStokes2DParabolicFunction.c:112-143 and :155-186),
// Sutherland viscosity and the conductivity law of navierstokes2d.h:315-331
// (raiseto(x,1.5) = exp(1.5 log x), include/math_ops.h:40).
// FV planes: 0..2 = x-flux components 1..3, 3..5 = y-flux components 1..3.
__global__ void __launch_bounds__(256) k_viscflux(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const Grid &G = C.G;
  const PhysParams &P = C.P;
  if (t >= (long)G.nxg * G.nyg) return;
  const int i = (int)(t % G.nxg) - G.g, j = (int)(t / G.nxg) - G.g;
  const bool ix = (i >= 0 && i < G.nx), jy = (j >= 0 && j < G.ny);
  if (!ix && !jy) return;  // corners are never read
  const long pl = (long)G.nxg * G.nyg;
  const double *Qd = C.Qd;
  double *FV = C.FV;
  const Prim q = prim_at(G.plane, P.gamma, C.U, t);
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
// (TimeRHSFunctionExplicit.c:55-74).  hyp = dxinv (fI[k+1] - fI[k]) summed
// over the dimensions in order (HyperbolicFunction.c:97-112): the x part from
// the x-sweep interfaces (contiguous along x), the y part from the y-sweep
// interfaces, which are contiguous along y and pass through an LDS tile;
// par: derivatives of the viscous fluxes; source: ChemistrySource
// (ChemistrySolve.c:18-115, chemistry.h:30-127) with the photon density of
// the Gaussian pulse at the stage time.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void viscous_par(const CaseDev &C, int i, int j, long p, double par[3]) {
  const Grid &G = C.G;
  const PhysParams &P = C.P;
  const long pl = (long)G.nxg * G.nyg;
  const double dxi = C.dxinv[G.g + i], dyi = C.dxinv[(G.nx + 2 * G.g) + G.g + j];
  for (int c = 0; c < 3; c++) {
    const double *Fx = C.FV + c * pl, *Fy = C.FV + (3 + c) * pl;
    double dfx, dfy;
    if (P.par4) {
      const double t12 = 1.0 / 12.0;
      dfx = (Fx[p - 2] - 8 * Fx[p - 1] + 8 * Fx[p + 1] - Fx[p + 2]) * t12;
      dfy = (Fy[p - 2 * G.nxg] - 8 * Fy[p - G.nxg] + 8 * Fy[p + G.nxg] - Fy[p + 2 * G.nxg]) * t12;
    } else {
      dfx = 0.0 * Fx[p + 1] - (-1.0) * Fx[p] + (-1.0) * Fx[p - 1];
      dfy = 0.0 * Fy[p + G.nxg] - (-1.0) * Fy[p] + (-1.0) * Fy[p - G.nxg];
    }
    par[c] = dxi * dfx;
    par[c] = par[c] + dyi * dfy;
  }
}

// Sponge zones act as a source (BCSponge.c:29-74, called by
// SourceFunction.c:44-58 after the physics source): zone by zone,
// s - sigma (U - uref) with sigma linear in the coordinate along the zone's dim.
__device__ __forceinline__ bool in_sponge(const CaseDev &C, int i, int j) {
  for (int z = 0; z < C.nzones; z++) {
    const Zone &Z = C.zones[z];
    if (Z.kind == Z_SPONGE && i >= Z.is[0] && i < Z.ie[0] && j >= Z.is[1] && j < Z.ie[1]) return true;
  }
  return false;
}

__device__ __forceinline__ double sponge(const CaseDev &C, int i, int j, int v, long p, double s) {
  const Grid &G = C.G;
  for (int z = 0; z < C.nzones; z++) {
    const Zone &Z = C.zones[z];
    if (Z.kind != Z_SPONGE || i < Z.is[0] || i >= Z.ie[0] || j < Z.is[1] || j >= Z.ie[1]) continue;
    const double x = Z.dim ? C.xc[(G.nx + 2 * G.g) + G.g + j] : C.xc[G.g + i];
    const double sigma = (Z.face > 0) ? (x - Z.xmin) / (Z.xmax - Z.xmin) : (x - Z.xmax) / (Z.xmin - Z.xmax);
    s -= (sigma * (C.U[(long)v * G.plane + p] - C.bcvals[Z.vals + v]));
  }
  return s;
}

// photon density reaching layer 0 at the stage time: the Gaussian pulse
// (ChemistrySolve.c:36-43) times the beam map
__device__ __forceinline__ double chem_nh0(const CaseDev &C, int stage, long p) {
  const PhysParams &P = C.P;
  const double time = *C.time + C.c[stage] * C.dt;
  double hs = 0.0;
  if (time > P.t_start_norm) {
    const double sigma = P.t_pulse_norm / 2.35482;
    const double tp = time - (P.t_start_norm + P.t_pulse_norm);
    hs = P.I0 * exp(-(tp * tp) / (2 * sigma * sigma));
  }
  return (hs * C.imap[p]) / P.hnu_den;
}

// sources of layer iz at point p (chemistry.h:30-127): its seven species s[]
// and the energy term, for photon density nh reaching the layer
__device__ __forceinline__ void chem_layer(const CaseDev &C, int iz, long p, double nh, double s[7], double &src_e) {
  const PhysParams &P = C.P;
  const double *U = C.U;
  const long pl = C.G.plane;
  const int nfv = P.nflow + P.z_stride * iz;
  const double nO2 = U[(long)(nfv + 0) * pl + p], nO3 = U[(long)(nfv + 1) * pl + p];
  const double n1D = U[(long)(nfv + 2) * pl + p], n1Dg = U[(long)(nfv + 3) * pl + p];
  const double n3Su = U[(long)(nfv + 4) * pl + p], n1Sg = U[(long)(nfv + 5) * pl + p];
  const double nCO2 = U[(long)(nfv + 6) * pl + p];
  s[0] = 0.0;
  s[1] = -(P.k0a + P.k0b) * nh * nO3 - (P.k2a + P.k2b) * nO3 * n1D - (P.k3a + P.k3b) * nO3 * n1Sg -
         P.k5 * nO3 * n3Su - P.k6 * n1Dg * nO3;
  s[2] = P.k0a * nh * nO3 - (P.k1a + P.k1b) * n1D * nO2 - (P.k2a + P.k2b) * n1D * nO3 - P.k4 * n1D * nCO2;
  s[3] = P.k0a * nh * nO3 + P.k5 * nO3 * n3Su - P.k6 * n1Dg * nO3;
  s[4] = P.k2a * nO3 * n1D - P.k5 * nO3 * n3Su;
  s[5] = P.k1a * n1D * nO2 - (P.k3a + P.k3b) * nO3 * n1Sg;
  s[6] = 0.0;
  src_e = ((P.q0a * P.k0a + P.q0b * P.k0b) * nh * nO3 + (P.q1a * P.k1a + P.q1b * P.k1b) * nO2 * n1D +
           (P.q2a * P.k2a + P.q2b * P.k2b) * nO3 * n1D + (P.q3a * P.k3a + P.q3b * P.k3b) * nO3 * n1Sg +
           P.q4 * P.k4 * n1D * nCO2 + P.q5 * P.k5 * nO3 * n3Su + P.q6 * P.k6 * n1Dg * nO3) /
          P.ch_gm1;
}

// rhs += chemistry source, layer by layer (ChemistrySolve.c:45-110); the
// energy takes the last layer's term
__device__ __forceinline__ void chem_source(const CaseDev &C, int stage, long p, double *udot, bool sp, int i,
                                            int j) {
  const Grid &G = C.G;
  const PhysParams &P = C.P;
  const long pl = G.plane;
  double nh = chem_nh0(C, stage, p);
  double src_e = 0.0;
  for (int iz = 0; iz < P.nlayers; iz++) {
    const int nfv = P.nflow + P.z_stride * iz;
    if (iz > 0) nh = nh * (1.0 - P.damp_coef * C.U[(long)(P.nflow + P.z_stride * (iz - 1) + 1) * pl + p]);
    double s[7];
    chem_layer(C, iz, p, nh, s, src_e);
    for (int m = 0; m < 7; m++) {
      double *o = udot + (long)(nfv + m) * pl + p;
      *o = *o + (sp ? sponge(C, i, j, nfv + m, p, s[m]) : s[m]);
    }
  }
  double *o = udot + (long)(P.nflow - 1) * pl + p;
  *o = *o + (sp ? sponge(C, i, j, P.nflow - 1, p, src_e) : src_e);
}

// The source term at interior point (i, j), ghosted offset p: chemistry,
// then sponge zones (SourceFunction.c:21-60); rhs += source.
__device__ __forceinline__ void add_source(const CaseDev &C, int stage, int i, int j, long p, double *udot) {
  const bool sp = C.sponge && in_sponge(C, i, j);
  if (C.P.chem) chem_source(C, stage, p, udot, sp, i, j);
  if (!sp) return;
  const int nplain = C.P.chem ? C.P.nflow - 1 : C.G.nv;  // variables chemistry leaves at zero
  for (int v = 0; v < nplain; v++) {
    double *o = udot + (long)v * C.G.plane + p;
    *o = *o + sponge(C, i, j, v, p, 0.0);
  }
}

__global__ void __launch_bounds__(256) k_assemble(CASE_ARGS) {
  CASE_PREAMBLE
  if (stage >= C.nstages) return;
  const Grid &G = C.G;
  const int nv = G.nv;
  double *udot = C.Udot[stage];
  if (G.ndims == 1) {
    if (t >= G.nx) return;
    const int i = (int)t;
    const long p = i + G.g;
    const double dxi = C.dxinv[G.g + i];
    for (int v = 0; v < nv; v++) {
      const long a = (long)v * (G.nx + 1) + i;
      const double hyp = dxi * (C.fIx[a + 1] - C.fIx[a]);
      udot[(long)v * G.plane + p] = (C.terms & 1) ? -hyp : 0.0;
    }
    if (C.terms & 4) add_source(C, stage, i, 0, p, udot);
    return;
  }
  // 2D: a 32 x 32 tile of points, four per thread
  __shared__ double tl[32][33];
  const int ntx = (G.nx + 31) / 32;
  const int nb = ntx * ((G.ny + 31) / 32);  // tiles; the case's blocks are padded to a multiple of eight
  const int id = xcd_order((int)(blockIdx.x - firstblock[cidx]), (nb + 7) & ~7, firstblock[cidx]);
  if (id >= nb) return;
  const int i0 = (id % ntx) * 32, j0 = (id / ntx) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i = i0 + tx;
  const bool visc = C.P.viscous && (C.terms & 2);
  double par[4][3];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int j = j0 + ty + 8 * q;
    par[q][0] = par[q][1] = par[q][2] = 0.0;
    if (visc && i < G.nx && j < G.ny) viscous_par(C, i, j, (long)(j + G.gy) * G.nxg + (i + G.g), par[q]);
  }
  const double dxi = i < G.nx ? C.dxinv[G.g + i] : 0.0;
#if PIAFS_ASM_PIPELINE
  // y-sweep interfaces of lines i0..i0+31, interfaces j0..j0+32, of variable
  // v: this thread's share (rows ty + 8q, interfaces tx and tx + 32), loaded
  // one variable ahead so that the loads are in flight during the previous
  // variable's work; the barriers order the LDS tile only, so that global
  // loads and stores stay in flight across them
  double fy[4][2];
  auto load_y = [&](int v) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int ii = i0 + ty + 8 * q;
      const double *row = C.fIy + ((long)v * G.nx + ii) * (G.ny + 1);
#pragma unroll
      for (int k = 0; k < 2; k++) {
        const int a = tx + 32 * k, j = j0 + a;
        fy[q][k] = (ii < G.nx && a < 33 && j <= G.ny) ? row[j] : 0.0;
      }
    }
  };
  load_y(0);
  for (int v = 0; v < nv; v++) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      tl[ty + 8 * q][tx] = fy[q][0];
      if (tx == 0) tl[ty + 8 * q][32] = fy[q][1];
    }
    lds_barrier();
    if (v + 1 < nv) load_y(v + 1);
#else
  for (int v = 0; v < nv; v++) {
    // y-sweep interfaces of lines i0..i0+31, interfaces j0..j0+32
    for (int b = ty; b < 32; b += 8) {
      const int ii = i0 + b;
      if (ii < G.nx) {
        const double *row = C.fIy + ((long)v * G.nx + ii) * (G.ny + 1);
        for (int a = tx; a < 33; a += 32) {
          const int j = j0 + a;
          if (j <= G.ny) tl[b][a] = row[j];
        }
      }
    }
    __syncthreads();
#endif
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int jl = ty + 8 * q, j = j0 + jl;
      if (i < G.nx && j < G.ny) {
        const long a = ((long)v * G.ny + j) * (G.nx + 1) + i;
        double hyp = dxi * (C.fIx[a + 1] - C.fIx[a]);
        hyp = hyp + C.dxinv[(G.nx + 2 * G.g) + G.g + j] * (tl[tx][jl + 1] - tl[tx][jl]);
        double r = (C.terms & 1) ? -hyp : 0.0;
        if (visc && v >= 1 && v <= 3) r = r + par[q][v - 1];
        udot[(long)v * G.plane + (long)(j + G.gy) * G.nxg + (i + G.g)] = r;
      }
    }
#if PIAFS_ASM_PIPELINE
    lds_barrier();
#else
    __syncthreads();
#endif
  }
  if ((C.P.chem || C.sponge) && (C.terms & 4)) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int j = j0 + ty + 8 * q;
      if (i < G.nx && j < G.ny) add_source(C, stage, i, j, (long)(j + G.gy) * G.nxg + (i + G.g), udot);
    }
  }
}

// u += (dt*b_s) * Udot_s stage by stage (TimeRK.c:86-99) at interior points;
// time += dt (TimePostStep.c:33).
__global__ void __launch_bounds__(256) k_update(CASE_ARGS) {
  CASE_PREAMBLE
  const Grid &G = C.G;
  const long nint = (long)G.nx * G.ny;
  if (t >= nint * G.nv) return;
  if (t == 0) *C.time = *C.time + C.dt;
  const int v = (int)(t / nint);
  const long q = t % nint;
  const int i = (int)(q % G.nx), j = (int)(q / G.nx);
  const long p = (long)v * G.plane + (long)(j + G.gy) * G.nxg + (i + G.g);
  double x = C.u[p];
  for (int s = 0; s < C.nstages; s++) x += (C.dt * C.b[s]) * C.Udot[s][p];
  C.u[p] = x;
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


// Batched tridiagonal solve for the unit-level entry point, one team per
// system (tridiagLU.c:86-277 semantics with a_0 = c_{n-1} = 0, as in the
// reference's serial use): systems stored [n][ns], rows distributed over the
// team's lanes as in the reconstruction.
__global__ void __launch_bounds__(TB) k_tridiag(const double *__restrict__ a, const double *__restrict__ b,
                                                const double *__restrict__ c, double *__restrict__ x, int n, int ns,
                                                int T, int m, int *__restrict__ flag) {
  __shared__ double red[16];
  const int lane = threadIdx.x & (T - 1);
  const long d = (long)blockIdx.x * (TB / T) + threadIdx.x / T;
  const bool active = d < ns;
  const int lo = lane * m;
  int nr = n - lo;
  nr = nr < 0 ? 0 : (nr > m ? m : nr);
  if (!active) nr = 0;
  double A[kMaxRows], B[kMaxRows], Cc[kMaxRows], X[kMaxRows];
#pragma unroll
  for (int i = 0; i < kMaxRows; i++) {
    if (i < nr) {
      const long q = (long)(lo + i) * ns + d;
      A[i] = (lo + i == 0) ? 0.0 : a[q];
      B[i] = b[q];
      Cc[i] = (lo + i == n - 1) ? 0.0 : c[q];
      X[i] = x[q];
    } else {
      A[i] = 0.0; B[i] = 1.0; Cc[i] = 0.0; X[i] = 0.0;
    }
  }
  team_solve<kMaxRows>(A, B, Cc, X, nr, lane, T, lane == (n - 1) / m, 0.0, 0.0, red, flag, false);
#pragma unroll
  for (int i = 0; i < kMaxRows; i++)
    if (i < nr) x[(long)(lo + i) * ns + d] = X[i];
}

// The same for systems longer than one team covers: one thread per system,
// the sequential sweep itself (tridiagLU.c:133-150 and :243-255, a_0 = c_{n-1}
// = 0).  b receives the pivots b', x receives x' and then the solution.
__global__ void __launch_bounds__(TB) k_tridiag_seq(const double *__restrict__ a, double *b,
                                                    const double *__restrict__ c, double *__restrict__ x, int n,
                                                    int ns, int *__restrict__ flag) {
  const long d = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ns) return;
  double bp = b[d], xp = x[d], cp = (n > 1) ? c[d] : 0.0;
  bool zp = bp == 0.0;
  for (int i = 1; i < n; i++) {
    const long q = (long)i * ns + d;
    const double fac = a[q] / bp;
    bp = b[q] - fac * cp;
    xp = x[q] - fac * xp;
    cp = (i == n - 1) ? 0.0 : c[q];
    zp = zp || bp == 0.0;
    b[q] = bp;
    x[q] = xp;
  }
  if (zp) atomicOr(flag, 1);
  double xn = x[(long)(n - 1) * ns + d] / b[(long)(n - 1) * ns + d];
  x[(long)(n - 1) * ns + d] = xn;
  for (int i = n - 2; i >= 0; i--) {
    const long q = (long)i * ns + d;
    xn = (x[q] - c[q] * xn) / b[q];
    x[q] = xn;
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

// ---------------------------------------------------------------------------
// Device-buffer operators on the reference's own array layouts, for callers
// that keep the reference's data structures on the device.  Ghosted arrays:
// every grid point with its ghost layers, x fastest, the nvars components of
// a point contiguous (include/hypar.h:106-111).  Interface arrays of
// dimension dir: the points of the grid with dim[dir]+1 interfaces along dir
// and no ghosts, same ordering (HyperbolicFunction.c:61-70).
// ---------------------------------------------------------------------------
// point offset of (i, j) (interior indices; ghosts negative) in a ghosted array
__device__ __forceinline__ long aos_point(const AosGrid &A, int i, int j) {
  return A.ndims == 2 ? (long)(j + A.g) * (A.nx + 2 * A.g) + (i + A.g) : (long)(i + A.g);
}
__device__ __forceinline__ long aos_points(const AosGrid &A) {
  return (long)(A.nx + 2 * A.g) * (A.ndims == 2 ? A.ny + 2 * A.g : 1);
}

// FluxFunction along dir at every point of the ghosted grid
__global__ void __launch_bounds__(TB) k_dev_flux(AosGrid A, int model, double gamma, int dir,
                                                 const double *__restrict__ u, double *__restrict__ f) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= aos_points(A)) return;
  const double *q = u + p * A.nv;
  double *o = f + p * A.nv;
  point_flux(model, dir, gamma, A.nv, [&](int m) { return q[m]; }, [&](int m, double x) { o[m] = x; });
}

// FirstDerivativePar along dir at every position (ghosts included) of every
// interior line across dir, one thread per (position, line, variable):
// FirstDerivativeFourthOrderCentral (FirstDerivativeFourthOrder.c:39-132) with
// par4, else FirstDerivativeFirstOrder (FirstDerivativeFirstOrder.c:20-95) with
// the given bias.  Points off the interior lines are left as they are.
__global__ void __launch_bounds__(TB) k_dev_fd(AosGrid A, int dir, int par4, int bias,
                                               const double *__restrict__ f, double *__restrict__ df) {
  const int N = dir ? A.ny : A.nx, g = A.g, nv = A.nv, P = N + 2 * g;
  const int nl = A.ndims == 2 ? (dir ? A.nx : A.ny) : 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)P * nl * nv) return;
  const int v = (int)(t % nv);
  const long r = t / nv;
  const int i = (int)(r % P) - g, l = (int)(r / P);
  auto F = [&](int k) { return f[aos_point(A, dir ? l : k, dir ? k : l) * nv + v]; };
  double d;
  if (par4) {
    const double t12 = 1.0 / 12.0;
    if (i == -g) d = (-25 * F(i) + 48 * F(i + 1) - 36 * F(i + 2) + 16 * F(i + 3) - 3 * F(i + 4)) * t12;
    else if (i == -g + 1) d = (-3 * F(i - 1) - 10 * F(i) + 18 * F(i + 1) - 6 * F(i + 2) + F(i + 3)) * t12;
    else if (i == N + g - 2) d = (-F(i - 3) + 6 * F(i - 2) - 18 * F(i - 1) + 10 * F(i) + 3 * F(i + 1)) * t12;
    else if (i == N + g - 1) d = (3 * F(i - 4) - 16 * F(i - 3) + 36 * F(i - 2) - 48 * F(i - 1) + 25 * F(i)) * t12;
    else d = (F(i - 2) - 8 * F(i - 1) + 8 * F(i + 1) - F(i + 2)) * t12;
  } else {
    const double mx = bias > 0 ? bias : 0, mn = bias < 0 ? bias : 0, bb = bias;
    if (i == -g) d = F(i + 1) - F(i);
    else if (i == N + g - 1) d = F(i) - F(i - 1);
    else d = mx * F(i + 1) - bb * F(i) + mn * F(i - 1);
  }
  df[aos_point(A, dir ? l : i, dir ? i : l) * nv + v] = d;
}

// SetInterpLimiterVar = WENOFifthOrderCalculateWeights{,Char}
// (WENOFifthOrderCalculateWeights.c:41-84, :112-237 and the M / Z / YC and
// characteristic variants): the nonlinear weights of every interface of dir
// on every interior line, for the left-biased flux, left-biased solution,
// right-biased flux and right-biased solution, in that order, set k at
// w + k*size.  Within a set, interface point p, variable v at p*nvars + v.
// Characteristic: the stencils are projected with the left eigenvectors of
// the Roe average of the two cells beside the interface (the flow block;
// species rows are the identity).  One thread per (interface, line).
template <int KIND>
__device__ __forceinline__ void dev_weights_at(const AosGrid &A, int model, double gam, int dir, int crweno,
                                               int characteristic, int no_limiting, double eps, int k, int l,
                                               const double *__restrict__ fC, const double *__restrict__ u,
                                               double *w1, double *w2, double *w3, long size) {
  const int nv = A.nv, N = dir ? A.ny : A.nx;
  const long p = dir ? (long)k * A.nx + l : (A.ndims == 2 ? (long)l * (A.nx + 1) + k : (long)k);
  auto at = [&](const double *a, int pos, int v) { return a[aos_point(A, dir ? l : pos, dir ? pos : l) * nv + v]; };
  double c1, c2, c3;
  if (crweno && k > 0 && k < N) { c1 = 0.2; c2 = 0.5; c3 = 0.3; }
  else { c1 = 0.1; c2 = 0.6; c3 = 0.3; }
  const int nf = characteristic ? (model == 1 ? 3 : 4) : 0;
  double L[4][4] = {{0}};
  if (nf == 3) {
    double a0, a1, a2;
    euler1d_roeavg(gam, at(u, k - 1, 0), at(u, k - 1, 1), at(u, k - 1, 2), at(u, k, 0), at(u, k, 1), at(u, k, 2),
                   a0, a1, a2);
    Eig3 E;
    euler1d_eig(gam, a0, a1, a2, E);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) L[a][b] = E.L[a][b];
  } else if (nf == 4) {
    const double uLc[4] = {at(u, k - 1, 0), at(u, k - 1, 1), at(u, k - 1, 2), at(u, k - 1, 3)};
    const double uRc[4] = {at(u, k, 0), at(u, k, 1), at(u, k, 2), at(u, k, 3)};
    double ua[4], R[4][4];
    ns2d_roeavg(gam, uLc, uRc, ua);
    if (dir) ns2d_eigvecs<1>(gam, ua, L, R);
    else ns2d_eigvecs<0>(gam, ua, L, R);
  }
  // stencil value of variable v at position pos (projected for the flow rows)
  auto sv = [&](const double *a, int pos, int v) {
    if (v >= nf) return at(a, pos, v);
    double y = L[v][0] * at(a, pos, 0);
    for (int b = 1; b < nf; b++) y = y + L[v][b] * at(a, pos, b);
    return y;
  };
  for (int set = 0; set < 4; set++) {  // LF, LU, RF, RU
    const double *a = (set & 1) ? u : fC;
    const bool left = set < 2;
    for (int v = 0; v < nv; v++) {
      double x1, x2, x3;
      if (no_limiting) {
        x1 = c1; x2 = c2; x3 = c3;
      } else if (left) {
        weno_w<KIND>(c1, c2, c3, sv(a, k - 3, v), sv(a, k - 2, v), sv(a, k - 1, v), sv(a, k, v), sv(a, k + 1, v),
                     eps, x1, x2, x3);
      } else {
        weno_w<KIND>(c1, c2, c3, sv(a, k + 2, v), sv(a, k + 1, v), sv(a, k, v), sv(a, k - 1, v), sv(a, k - 2, v),
                     eps, x1, x2, x3);
      }
      const long o = set * size + p * nv + v;
      w1[o] = x1;
      w2[o] = x2;
      w3[o] = x3;
    }
  }
}

__global__ void __launch_bounds__(TB) k_dev_weights(AosGrid A, int model, double gam, int dir, int kind,
                                                    int crweno, int characteristic, int no_limiting, double eps,
                                                    const double *__restrict__ fC, const double *__restrict__ u,
                                                    double *w1, double *w2, double *w3, long size) {
  const int N = dir ? A.ny : A.nx, nl = A.ndims == 2 ? (dir ? A.nx : A.ny) : 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)(N + 1) * nl) return;
  const int k = (int)(t % (N + 1)), l = (int)(t / (N + 1));
  switch (kind) {
    case 1: dev_weights_at<1>(A, model, gam, dir, crweno, characteristic, no_limiting, eps, k, l, fC, u, w1, w2, w3, size); break;
    case 2: dev_weights_at<2>(A, model, gam, dir, crweno, characteristic, no_limiting, eps, k, l, fC, u, w1, w2, w3, size); break;
    case 3: dev_weights_at<3>(A, model, gam, dir, crweno, characteristic, no_limiting, eps, k, l, fC, u, w1, w2, w3, size); break;
    default: dev_weights_at<0>(A, model, gam, dir, crweno, characteristic, no_limiting, eps, k, l, fC, u, w1, w2, w3, size); break;
  }
}

// Ghosted reference-layout array <-> the planar layouts of a case: the natural
// copy (x fastest) and, in 2D, the transposed copy (y fastest), either may be
// null.  One thread per (point, variable).  to_planar = 0 copies the other
// way, from the natural copy, at the ghost points only when ghosts_only.
__global__ void __launch_bounds__(TB) k_dev_planar(Grid G, const double *__restrict__ a, double *nat, double *tr,
                                                   int to_planar, int ghosts_only, double *out) {
  const long np = (long)G.nxg * G.nyg;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * G.nv) return;
  const int v = (int)(t % G.nv);
  const long p = t / G.nv;
  const int i = (int)(p % G.nxg) - G.g, j = (int)(p / G.nxg) - G.gy;
  const long pn = (long)v * G.plane + p;
  if (to_planar) {
    const double x = a[p * G.nv + v];
    if (nat) nat[pn] = x;
    if (tr) tr[(long)v * G.plane + (long)(i + G.g) * G.nyg + (j + G.gy)] = x;
    return;
  }
  const bool ghost = i < 0 || i >= G.nx || j < 0 || j >= G.ny;
  const bool corner = (i < 0 || i >= G.nx) && (j < 0 || j >= G.ny) && G.ndims == 2;
  if (ghosts_only && (!ghost || corner)) return;  // the reference never writes the corners
  out[p * G.nv + v] = nat[pn];
}

// Interface arrays of one sweep: reference layout <-> line-major
// a[(v*nlines + l)*(N+1) + k].  One thread per (interface, line, variable).
__global__ void __launch_bounds__(TB) k_dev_iface(int ndims, int dir, int nx, int N, int nl, int nv,
                                                  const double *__restrict__ src, double *__restrict__ dst,
                                                  int to_lines) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)(N + 1) * nl * nv) return;
  const int v = (int)(t % nv);
  const long r = t / nv;
  const int k = (int)(r % (N + 1)), l = (int)(r / (N + 1));
  const long p = dir ? (long)k * nx + l : (ndims == 2 ? (long)l * (N + 1) + k : (long)k);
  const long q = ((long)v * nl + l) * (N + 1) + k;
  if (to_lines) dst[q] = src[p * nv + v];
  else dst[p * nv + v] = src[q];
}


}  // namespace

// ===========================================================================
// host side: work decomposition and launchers
// ===========================================================================

static const char *kKNames[K_NCLASS] = {"stage_bc", "ghost",  "flux",    "recon_ends", "recon", "recon_full",
                                        "long_rows", "long_solve",
                                        "recon_char", "cw_rows", "cw_solve", "upwind", "bflux",
                                        "qderiv",   "viscflux", "assemble", "update"};
const char *kclass_name(int k) { return kKNames[k]; }

int cw_row_fields(int nflow, int nv) { return nflow * nflow + 16 * nv; }
int cw_back_fields(int nflow, int nv) { return 2 * nflow * nflow + nflow + 3 * (nv - nflow); }

bool team_shape(int N, int max_rows, int &team, int &rows) {
  if (max_rows < 1 || max_rows > kMaxRows) max_rows = kMaxRows;
  const int nrows = N - 1 > 1 ? N - 1 : 1;
  team = 1;
  while ((long)team * max_rows < nrows) team <<= 1;
  rows = (nrows + team - 1) / team;
  return team <= 256;
}

bool stage_tile_64(const CaseDev &c) { return c.G.ndims == 2 && (long)c.G.nx * c.G.ny >= 2048L * 2048L; }

#ifndef PIAFS_SP_TEAM_MIN
#define PIAFS_SP_TEAM_MIN 32
#endif
#ifndef PIAFS_SP_TEAM_MAX
#define PIAFS_SP_TEAM_MAX 64
#endif
bool species_in_recon(const CaseDev &c) {
  const bool two = (c.G.ndims == 2);
  return PIAFS_RECON_SPECIES && PIAFS_RECON_LDS && two && c.model == 2 && c.paired && c.upwind == 0 &&
         !c.R.characteristic && c.G.nv > 4 && c.dsp[0] && c.dsp[1] && c.Sx.team <= PIAFS_SP_TEAM_MAX &&
         c.Sy.team <= PIAFS_SP_TEAM_MAX && c.Sx.team >= PIAFS_SP_TEAM_MIN && c.Sy.team >= PIAFS_SP_TEAM_MIN &&
         long_rows(c.Sx, c.G.nv) == 0 && long_rows(c.Sy, c.G.nv) == 0;
}

void case_blocks(const CaseDev &c, int blocks[K_NCLASS]) {
  const Grid &G = c.G;
  const bool two = (G.ndims == 2);
  auto nb = [](long n) { return (int)((n + TB - 1) / TB); };
  const int st = c.stage_tile == 64 ? 64 : 32;
  // the tile kernels' counts are padded to a multiple of eight, so that every
  // case's tiles start on one and can be taken in xcd_order
  auto pad8 = [](int n) { return (n + 7) & ~7; };
  blocks[K_STAGE] = pad8(two ? ((G.nx + st - 1) / st) * ((G.ny + st - 1) / st) * G.nv : nb((long)G.nx * G.nv));
  blocks[K_GHOST] = nb(ghost_points(G) * G.nv);
  blocks[K_FLUX] = nb((long)(c.Sx.N + 2 * c.Sx.g) * c.Sx.nlines + (two ? (long)(c.Sy.N + 2 * c.Sy.g) * c.Sy.nlines : 0));
  blocks[K_CWROWS] = blocks[K_CWSOLVE] = blocks[K_LROWS] = blocks[K_LSOLVE] = 0;
  if (c.R.characteristic && c.R.crweno) {
    const int nf = c.model == 1 ? 3 : 4;
    blocks[K_ENDS] = blocks[K_RECON] = blocks[K_RECONF] = blocks[K_RCHAR] = 0;
    blocks[K_CWROWS] = nb((long)(c.Sx.N + 1) * c.Sx.nlines + (two ? (long)(c.Sy.N + 1) * c.Sy.nlines : 0));
    blocks[K_CWSOLVE] = nb(cw_x_span(c, nf) + (two ? cw_items(c.Sy, G.nv, nf) : 0));
  } else if (c.R.characteristic) {
    blocks[K_ENDS] = 0;
    blocks[K_RECON] = blocks[K_RECONF] = 0;
    blocks[K_RCHAR] = nb((long)(c.Sx.N + 1) * c.Sx.nlines + (two ? (long)(c.Sy.N + 1) * c.Sy.nlines : 0));
  } else {
    blocks[K_RCHAR] = 0;
    blocks[K_ENDS] = nb(2 * (recon_systems(c.Sx, G.nv, c.paired) + (two ? recon_systems(c.Sy, G.nv, c.paired) : 0)));
    const int nrec = recon_sweep_blocks(c.Sx, G.nv, c.paired) + (two ? recon_sweep_blocks(c.Sy, G.nv, c.paired) : 0);
    // full segments: both sweeps' lanes hold kMaxRows rows each
    const bool full = PIAFS_RECON_LDS && PIAFS_RECON_FULL && c.paired && c.Sx.rows == kMaxRows &&
                      (!two || c.Sy.rows == kMaxRows);
    blocks[K_RECON] = full ? 0 : nrec;
    blocks[K_RECONF] = full ? nrec : 0;
    blocks[K_LROWS] = nb(long_rows(c.Sx, G.nv) + (two ? long_rows(c.Sy, G.nv) : 0));
    blocks[K_LSOLVE] = (int)(long_pairs(c.Sx, G.nv) + (two ? long_pairs(c.Sy, G.nv) : 0));  // a block per pair
  }
  blocks[K_UPWIND] = nb((long)(c.Sx.N + 1) * c.Sx.nlines + (two ? (long)(c.Sy.N + 1) * c.Sy.nlines : 0));
  blocks[K_BFLUX] = c.cons ? 1 : 0;
  const bool visc = two && c.P.viscous;
  blocks[K_QDERIV] = visc ? nb((long)G.nxg * G.nyg) : 0;
  blocks[K_VISCFLUX] = visc ? nb((long)G.nxg * G.nyg) : 0;
  blocks[K_ASSEMBLE] = pad8(two ? ((G.nx + 31) / 32) * ((G.ny + 31) / 32) : nb(G.nx));
  blocks[K_UPDATE] = nb((long)G.nx * G.ny * G.nv);
}

typedef void (*KernFn)(const CaseDev *, const int *, const int *, int);

static void run(KernFn k, int cls, const CaseDev *d, const Plan &pl, int s, hipStream_t st, MarkFn mark,
                void *ctx) {
  if (pl.nblocks[cls] <= 0) return;
  if (mark) mark(ctx, cls, true, st);
  hipLaunchKernelGGL(k, dim3(pl.nblocks[cls]), dim3(TB), 0, st, d, pl.blockcase[cls], pl.firstblock[cls], s);
  if (mark) mark(ctx, cls, false, st);
}

// k_qderiv and k_viscflux read only the stage state (after k_ghost) and
// write only Qd and FV, which nothing before k_assemble reads; the
// reconstruction / upwinding chain neither reads nor writes Qd or FV.
void launch_stage(const CaseDev *d, const Plan &pl, int s, hipStream_t st, MarkFn mark, void *ctx,
                  const StageFork *fk, int last) {
  auto R = [&](KernFn k, int cls, hipStream_t on) {
    if (cls <= last) run(k, cls, d, pl, s, on, mark, ctx);
  };
  R(pl.stage64 ? k_stage64 : k_stage, K_STAGE, st);
  R(k_ghost, K_GHOST, st);
  R(k_flux, K_FLUX, st);
  R(k_recon_ends, K_ENDS, st);
  const bool fork = fk && fk->side && last >= K_ASSEMBLE &&
                    (pl.nblocks[K_QDERIV] > 0 || pl.nblocks[K_VISCFLUX] > 0);
  if (fork) {
    hipEventRecord(fk->fork, st);
    hipStreamWaitEvent(fk->side, fk->fork, 0);
    R(k_qderiv, K_QDERIV, fk->side);
    R(k_viscflux, K_VISCFLUX, fk->side);
    hipEventRecord(fk->join, fk->side);
  }
  R(k_recon, K_RECON, st);
  R(k_recon_full, K_RECONF, st);
  R(k_lrows, K_LROWS, st);
  R(k_lsolve, K_LSOLVE, st);
  R(k_recon_char, K_RCHAR, st);
  R(k_cw_rows, K_CWROWS, st);
  R(k_cw_solve, K_CWSOLVE, st);
  R(k_upwind, K_UPWIND, st);
  R(k_bflux, K_BFLUX, st);
  if (fork) {
    hipStreamWaitEvent(st, fk->join, 0);
  } else {
    R(k_qderiv, K_QDERIV, st);
    R(k_viscflux, K_VISCFLUX, st);
  }
  R(k_assemble, K_ASSEMBLE, st);
}

void launch_update(const CaseDev *d, const Plan &pl, hipStream_t st) { run(k_update, K_UPDATE, d, pl, 0, st, nullptr, nullptr); }

int cfl_norm_blocks(const Grid &G) {
  const long n = (long)G.nx * G.ny;
  long b = (n + TB - 1) / TB;
  return (int)(b < 1024 ? b : 1024);
}

void launch_cfl_norm(const Grid &G, int model, double gamma, double dt, const double *u, const double *u_prev,
                     const double *dxinv, double *partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_cfl_norm, dim3(nblocks), dim3(TB), 0, s, G, model, gamma, dt, u, u_prev, dxinv, partials);
}

bool launch_tridiag(const double *a, double *b, const double *c, double *x, int n, int ns, int *flag,
                    hipStream_t s) {
  int T, m;
  if (!team_shape(n + 1, kMaxRows, T, m)) {  // longer than one team: one thread per system
    hipLaunchKernelGGL(k_tridiag_seq, dim3(nblk(ns)), dim3(TB), 0, s, a, b, c, x, n, ns, flag);
    return true;
  }
  const long teams_per_block = TB / T;
  const long nblocks = (ns + teams_per_block - 1) / teams_per_block;
  hipLaunchKernelGGL(k_tridiag, dim3((unsigned)nblocks), dim3(TB), 0, s, a, b, c, x, n, ns, T, m, flag);
  return true;
}

void launch_dev_flux(const AosGrid &A, int model, double gamma, int dir, const double *u, double *f, hipStream_t s) {
  const long n = (long)(A.nx + 2 * A.g) * (A.ndims == 2 ? A.ny + 2 * A.g : 1);
  hipLaunchKernelGGL(k_dev_flux, dim3(nblk(n)), dim3(TB), 0, s, A, model, gamma, dir, u, f);
}

void launch_dev_fd(const AosGrid &A, int dir, int par4, int bias, const double *f, double *df, hipStream_t s) {
  const int N = dir ? A.ny : A.nx, nl = A.ndims == 2 ? (dir ? A.nx : A.ny) : 1;
  const long n = (long)(N + 2 * A.g) * nl * A.nv;
  hipLaunchKernelGGL(k_dev_fd, dim3(nblk(n)), dim3(TB), 0, s, A, dir, par4, bias, f, df);
}

void launch_dev_weights(const AosGrid &A, int model, double gamma, int dir, int kind, int crweno,
                        int characteristic, int no_limiting, double eps, const double *fC, const double *u,
                        double *w1, double *w2, double *w3, long size, hipStream_t s) {
  const int N = dir ? A.ny : A.nx, nl = A.ndims == 2 ? (dir ? A.nx : A.ny) : 1;
  hipLaunchKernelGGL(k_dev_weights, dim3(nblk((long)(N + 1) * nl)), dim3(TB), 0, s, A, model, gamma, dir, kind,
                     crweno, characteristic, no_limiting, eps, fC, u, w1, w2, w3, size);
}

void launch_dev_planar(const Grid &G, const double *a, double *nat, double *tr, int to_planar, int ghosts_only,
                       double *out, hipStream_t s) {
  const long n = (long)G.nxg * G.nyg * G.nv;
  hipLaunchKernelGGL(k_dev_planar, dim3(nblk(n)), dim3(TB), 0, s, G, a, nat, tr, to_planar, ghosts_only, out);
}

void launch_dev_iface(int ndims, int dir, int nx, int N, int nl, int nv, const double *src, double *dst,
                      int to_lines, hipStream_t s) {
  const long n = (long)(N + 1) * nl * nv;
  hipLaunchKernelGGL(k_dev_iface, dim3(nblk(n)), dim3(TB), 0, s, ndims, dir, nx, N, nl, nv, src, dst, to_lines);
}

void launch_class(const CaseDev *d, const Plan &pl, int cls, int s, hipStream_t st) {
  static const KernFn table[K_NCLASS] = {k_stage, k_ghost, k_flux, k_recon_ends, k_recon, k_recon_full, k_lrows,
                                         k_lsolve, k_recon_char,
                                         k_cw_rows, k_cw_solve, k_upwind, k_bflux, k_qderiv, k_viscflux,
                                         k_assemble, k_update};
  run(cls == K_STAGE && pl.stage64 ? k_stage64 : table[cls], cls, d, pl, s, st, nullptr, nullptr);
}

void launch_weno_weights(int kind, const double *copt, const double *st, int count, double eps, double *w,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_weno_weights, dim3(nblk(count)), dim3(TB), 0, s, kind, copt[0], copt[1], copt[2], st, count,
                     eps, w);
}

}  // namespace piafs
