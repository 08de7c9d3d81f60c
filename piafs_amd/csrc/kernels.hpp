// Device kernels of the PIAFS RHS / RK path and their host launchers.
//
// Layouts (all FP64, one plane per variable):
//   natural  : u[v*plane + (j+gy)*nxg + (i+g)]      x fastest (2D) / u[v*plane + i+g] (1D)
//   x-sweep  : uT[v*plane + (i+g)*nyg + (j+gy)]     transposed copy, y fastest
//   interface: a[(v*(N+1) + k)*nlines + l]          k = interface along the sweep,
//                                                    l = interior line index (fast)
// A "sweep" is one direction's hyperbolic term over all interior lines.  In
// both sweep layouts the sweep coordinate is the slow index and the line index
// the fast one, so one lane per line reads coalesced rows at every position.
#pragma once

#include <hip/hip_runtime.h>

namespace piafs {

constexpr int MAXV = 64;

struct Grid {
  int ndims, nv, g;
  int nx, ny;        // interior sizes (ny = 1 in 1D)
  int gy;            // ghosts along y (0 in 1D)
  int nxg, nyg;      // ghosted sizes (nyg = 1 in 1D)
  long plane;        // elements per variable plane (>= nxg*nyg)
};

struct SweepGeom {
  int N;             // interior points along the sweep
  int nlines;        // interior lines
  int pos_stride;    // elements between consecutive positions along the sweep
  int line_off;      // element offset of interior line 0 within a row
  int g;
  long plane;
};

// Boundary map for one face: type (0 periodic, 1 extrapolate, -1 none) and the
// covered interior range [lo, hi) in the other dimension.
struct FaceBC {
  int type, lo, hi;
};
struct BCMap {
  FaceBC f[2][2];  // [dim][face: 0 = left (face +1), 1 = right (face -1)]
};

struct PhysParams {
  double gamma;
  // viscous (navierstokes2d)
  double inv_Re, inv_Pr, inv_gm1, Tref, T0, TS, TA, TB, two_third;
  double kappa_num;   // T0 + TA*exp(-TB/T0), constant numerator factor
  int viscous, par4;
  // chemistry
  int chem, nflow, nlayers, z_stride;
  double k0a, k0b, k1a, k1b, k2a, k2b, k3a, k3b, k4, k5, k6;
  double q0a, q0b, q1a, q1b, q2a, q2b, q3a, q3b, q4, q5, q6;
  double ch_gm1, hnu_den, damp_coef;   // (gamma_chem - 1), c*h*nu*n_O2, dz*sO3*n_O2
};

struct ReconParams {
  int crweno, kind, no_limiting;
  double eps;
};

// ---- launchers (stream-ordered, no synchronisation) ----
// Stage state with boundary conditions: U = u + sum_t a[t]*Udot[t] at every
// ghosted point, each ghost taken from its boundary source point.  Writes the
// natural layout and, in 2D, the transposed copy for the x-sweep.
void launch_stage_bc(const Grid &G, const BCMap &bc, const double *u, const double *const *udot,
                     const double *a, int nprev, double *U, double *UT, hipStream_t s);
void launch_flux(const Grid &G, const SweepGeom &S, int model, int dir, double gamma, const double *q,
                 double *f, hipStream_t s);
void launch_recon_components(const Grid &G, const SweepGeom &S, const ReconParams &R, const double *q,
                             const double *f, double *uL, double *uR, double *fL, double *fR,
                             double *scr_b, double *scr_c, hipStream_t s);
void launch_recon_char_euler1d(const Grid &G, const SweepGeom &S, const ReconParams &R, double gamma,
                               const double *q, const double *f, double *uL, double *uR, double *fL,
                               double *fR, hipStream_t s);
void launch_upwind(const Grid &G, const SweepGeom &S, int model, int upw, int dir, double gamma,
                   const double *q, const double *uL, const double *uR, const double *fL, const double *fR,
                   double *fI, hipStream_t s);
void launch_divergence(const Grid &G, const SweepGeom &S, const double *dxinv_dir, const double *fI,
                       double *h, hipStream_t s);
void launch_transpose(int nv, int rows, int cols, const double *src, double *dst, hipStream_t s);
void launch_qderiv(const Grid &G, const PhysParams &P, const double *U, const double *dxinv, double *Qd,
                   hipStream_t s);
void launch_viscflux(const Grid &G, const PhysParams &P, const double *U, const double *Qd, double *FV,
                     hipStream_t s);
void launch_assemble(const Grid &G, const PhysParams &P, const double *U, const double *hx,
                     const double *fIy, const double *FV, const double *dxinv, const double *imap,
                     double hnu_scale, double *udot, hipStream_t s);
void launch_update(const Grid &G, double *u, const double *const *udot, const double *bcoef, int nst,
                   hipStream_t s);
void launch_cfl_norm(const Grid &G, int model, double gamma, double dt, const double *u, const double *u_prev,
                     const double *dxinv, double *partials, int nblocks, hipStream_t s);
int cfl_norm_blocks(const Grid &G);

// per-operator kernels used by the unit-level entry points
void launch_tridiag(const double *a, const double *b, const double *c, double *x, int n, int ns, int *flag,
                    double *scr_b, hipStream_t s);
void launch_weno_weights(int kind, const double *copt, const double *st, int count, double eps, double *w,
                         hipStream_t s);

}  // namespace piafs
