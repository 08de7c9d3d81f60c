// Device kernels of the PIAFS RHS / RK path and their host launchers.
//
// Every kernel covers a whole ensemble of cases in one launch.  Each case's
// work items are padded to whole blocks so that a block belongs to exactly
// one case: block b runs case blockcase[b], local item
// (b - firstblock[case]) * blockDim + threadIdx.  A single simulation is an
// ensemble of one.
//
// Layouts (all FP64, one plane per variable):
//   natural   : u[v*plane + (j+gy)*nxg + (i+g)]     x fastest (2D) / u[v*plane + i+g] (1D)
//   transposed: uT[v*plane + (i+g)*nyg + (j+gy)]    y fastest (2D only)
//   interface : a[(v*nlines + l)*(N+1) + k]         k = interface along the sweep (fast),
//                                                    l = interior line index
// Every sweep reads lines that are contiguous in memory: the x-sweep the
// natural state, the y-sweep the transposed copy written by the stage kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace piafs {

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
  int g;
  long plane;
  long base;         // element offset of (line 0, position 0); positions run -g .. N+g-1
  int line_stride;   // elements between consecutive lines
  int team;          // lanes sharing one line system (power of two, <= 256)
  int rows;          // interior rows (1 .. N-1) held by each lane, <= kMaxRows
  int longl;         // lines longer than one team covers (team_shape fails): their
                     // component-wise systems go through k_lrows / k_lsolve instead
};

// Longest segment of a line system one lane keeps in registers (the
// register arrays of the reconstruction kernel are sized by it).
#ifndef PIAFS_MAX_ROWS
#define PIAFS_MAX_ROWS 16
#endif
constexpr int kMaxRows = PIAFS_MAX_ROWS;
// Doubles in the allocation-free entry point's device scratch array (32 MiB):
// piafs_tridiag_host_noalloc takes n*ns <= (capacity - 1) / 4.
#ifndef PIAFS_NOALLOC_SCRATCH_DOUBLES
#define PIAFS_NOALLOC_SCRATCH_DOUBLES ((size_t)1 << 22)
#endif
// Paired reconstruction with each team's line staged in LDS (1) or read by
// every lane from global memory (0).
#ifndef PIAFS_RECON_LDS
#define PIAFS_RECON_LDS 1
#endif
// Paired reconstruction of cases whose lanes all hold kMaxRows rows in a
// kernel of its own (class K_RECONF), whose line solves need no per-row
// bounds (rows past the system's end made identity rows).
#ifndef PIAFS_RECON_FULL
#define PIAFS_RECON_FULL 1
#endif
// NS2D paired Roe: the reconstruction writes the species' interface fluxes
// f_I = 0.5 (fL + fR) - |v_n| 0.5 (uR - uL) itself, with |v_n| per interface
// from k_flux, and k_upwind does the four flow variables only (1); or every
// variable goes through k_upwind (0).
#ifndef PIAFS_RECON_SPECIES
#define PIAFS_RECON_SPECIES 1
#endif
// Lane exchanges of the line-solve scans and relays by DPP (1) or by LDS
// permutes (0).
#ifndef PIAFS_SCAN_DPP
#define PIAFS_SCAN_DPP 1
#endif
// k_assemble loads each variable's y-interface tile one variable ahead and
// synchronises on the LDS tile only (1), or loads and waits variable by
// variable (0).
#ifndef PIAFS_ASM_PIPELINE
#define PIAFS_ASM_PIPELINE 1
#endif
// the tile kernels (k_assemble, k_stage with 32-point tiles) give each XCD a
// contiguous band of tiles (1) or take tiles in workgroup order (0)
#ifndef PIAFS_XCD_ORDER
#define PIAFS_XCD_ORDER 1
#endif
// blocks per CU that the register allocation of k_recon_full aims at (2: two
// waves per SIMD, which the DPP exchanges need a few registers' spill for)
#ifndef PIAFS_RECONF_MINB
#define PIAFS_RECONF_MINB 2
#endif
// Team shape for lines of N points: the fewest lanes (a power of two) whose
// segments of at most max_rows rows cover interior rows 1..N-1.  Returns
// false when a line is too long for one block.
bool team_shape(int N, int max_rows, int &team, int &rows);

// Boundary zones of boundary.inp, applied in list order: a later zone
// overwrites an earlier one where both set a variable at a ghost point.
enum ZoneKind {
  Z_PERIODIC = 0, Z_EXTRAPOLATE, Z_DIRICHLET, Z_REFLECT, Z_SPONGE, Z_NOSLIP_WALL, Z_SLIP_WALL,
  Z_SUBSONIC_INFLOW, Z_SUBSONIC_OUTFLOW, Z_SUBSONIC_AMBIVALENT, Z_SUPERSONIC_INFLOW, Z_SUPERSONIC_OUTFLOW
};
constexpr int kMaxZones = 16;
struct Zone {
  int kind, dim, face;  // face +1: the low end of dim, -1: the high end (boundary.inp)
  int is[2], ie[2];     // index box [is, ie) per dimension: the ghost layer along dim and the
                        // covered interior range across it; a sponge: a box of interior points
  int vals;             // offset of the zone's values in CaseDev::bcvals (dirichlet / sponge:
                        // nvars values; inflow: the passive scalars), -1 if none
  double rho, vel[2], p, gamma;
  double xmin, xmax;    // sponge: the zone's extent along dim
};

struct PhysParams {
  double gamma;
  // viscous (navierstokes2d)
  double inv_Re, inv_Pr, inv_gm1, Tref, T0, TS, TA, TB, two_third;
  double kappa_num;   // T0 + TA*exp(-TB/T0), constant numerator factor
  int viscous, par4;
  // derivative ghosts exchanged along dimension d (0: x, 1: y): the MPI
  // exchange of QDerivX / QDerivY that a run with more than one rank along a
  // periodic dimension performs (a no-op in a serial run)
  int xchg[2];
  // chemistry
  int chem, nflow, nlayers, z_stride;
  double k0a, k0b, k1a, k1b, k2a, k2b, k3a, k3b, k4, k5, k6;
  double q0a, q0b, q1a, q1b, q2a, q2b, q3a, q3b, q4, q5, q6;
  double ch_gm1, hnu_den, damp_coef;   // (gamma_chem - 1), c*h*nu*n_O2, dz*sO3*n_O2
  double I0, t_start_norm, t_pulse_norm;
};

struct ReconParams {
  int crweno, kind, no_limiting, characteristic;
  int plain_start;  // start the line-solve relay from trivial estimates (testing)
  double eps;
};

// Everything a kernel needs to know about one case (device-resident table).
struct CaseDev {
  Grid G;
  SweepGeom Sx, Sy;
  int nzones, sponge;  // sponge: the case has a sponge zone
  Zone zones[kMaxZones];
  PhysParams P;
  ReconParams R;
  int model, upwind, nstages;  // upwind: 0 Roe, 1 rf-char, 2 llf-char, 3 rusanov
  int paired;    // reconstruction writes 0.5*(uR-uL), 0.5*(fL+fR) only (Roe with component-wise interpolation)
  int fuse_sp;   // species interface fluxes written by the reconstruction (set by the plan; see species_in_recon)
  double *dsp[2];  // per sweep and interface: Roe-average |v_n| with the entropy fix (k_flux, when fuse_sp)
  int stage_tile; // side of k_stage's transpose tiles, 32 or 64 (set by the plan)
  int terms;     // terms k_assemble adds: bit 0 hyperbolic, bit 1 parabolic, bit 2 source (7 in a step)
  double dt, A[16], b[4], c[4];
  double *u, *U, *UT, *Udot[4];
  double *flux_x, *flux_y, *ix[4], *iy[4], *fIx, *fIy;
  double *Qd, *FV, *dxinv, *imap;
  double *xc;      // grid coordinates with ghosts, one dimension after the other
  double *bcvals;  // the zones' values (Zone::vals)
  double *ends;  // paired mode: raw end-interface values of both systems of every pair [dir][pair][side][0|N]
  double *time;  // current simulation time of the case (one double)
  // conservation check (conservation_check yes): boundary flux sums of each
  // stage [stage][(2 d + face) nv + v], the total boundary integral [nv], and
  // dS[d], the face element of dimension d (BoundaryIntegral.c:37-44)
  int cons;
  double dS[2];
  double *bflux, *tbi;
  // characteristic CRWENO5: block rows of each sweep [field][line][k] and the
  // back-substitution data of its four systems [field][system][line][k]
  double *cwr[2], *cwb[2];
  // sweeps of long lines: rows A, B, C and right-hand side of every system
  // [field][interface k][system], system = (line*nvars + var)*4 + r
  // (r: uL, uR, fL, fR), written by k_lrows and solved in place by k_lsolve
  double *lrow[2];
  int *status;   // bit 0: a zero pivot met in a line solve (tridiagLU.c:139-142 returns -1)
};

// Kernel classes launched per RK stage (plus the final update).
enum KClass {
  K_STAGE = 0, K_GHOST, K_FLUX, K_ENDS, K_RECON, K_RECONF, K_LROWS, K_LSOLVE, K_RCHAR, K_CWROWS, K_CWSOLVE, K_UPWIND,
  K_BFLUX, K_QDERIV, K_VISCFLUX, K_ASSEMBLE, K_UPDATE,
  K_NCLASS
};
const char *kclass_name(int k);

// Block -> case tables of every kernel class for one ensemble.
struct Plan {
  int ncases = 0;
  bool stage64 = false;  // k_stage with 64 x 64 tiles (every case at least 2048 x 2048)
  int nblocks[K_NCLASS] = {0};
  int *blockcase[K_NCLASS] = {nullptr};   // device arrays, one int per block
  int *firstblock[K_NCLASS] = {nullptr};  // device arrays, one int per case
};

// Host side: blocks each case needs for each kernel class (0 = class unused).
void case_blocks(const CaseDev &c, int blocks[K_NCLASS]);
// Whether the reconstruction can write the species' Roe interface fluxes
// itself (NS2D, paired Roe, teams of 32 or 64 lanes -- lines of 258 to 1025
// rows -- no line longer than a team, |v_n| buffers present).
bool species_in_recon(const CaseDev &c);
// Whether the case is large enough for k_stage's 64 x 64 tiles.
bool stage_tile_64(const CaseDev &c);
// doubles per interface of the characteristic CRWENO5 scratch of one sweep:
// block rows (the flow block of L, then a, b, c coefficients and the right-hand
// side of every variable of the four systems), back-substitution data (four
// systems: flow block inverse, super-diagonal block and x'; per species the
// reciprocal pivot, c and x')
int cw_row_fields(int nflow, int nv);
int cw_back_fields(int nflow, int nv);

// A second stream for the branch of a stage that does not depend on the
// reconstruction: the viscous derivatives and fluxes (HBM-bound) run beside
// the reconstruction (FP64-VALU-bound).  The branch forks from the main
// stream before the reconstruction and joins it before the assembly.
struct StageFork {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// Stage s of the RK step for every case in the plan: stage state + BCs,
// fluxes, reconstruction, upwinding, divergence, viscous terms, assembly of
// Udot[s].  Times: stage time = *time + c[s]*dt of each case.  With `fk`
// the viscous branch runs on fk->side.  `mark` (per-class timing) is called
// before and after each launch with the stream the launch goes to.
typedef void (*MarkFn)(void *ctx, int cls, bool begin, hipStream_t st);
// With `last` < K_NCLASS only the classes up to `last` run (in launch order,
// which is the order of KClass), for the per-operator entry points.
void launch_stage(const CaseDev *d_cases, const Plan &plan, int s, hipStream_t st, MarkFn mark, void *mark_ctx,
                  const StageFork *fk = nullptr, int last = K_NCLASS);
// u += dt*b_s*Udot_s for every stage, time += dt.
void launch_update(const CaseDev *d_cases, const Plan &plan, hipStream_t st);

// diagnostics (single case): per-block partial max CFL and sum of squared
// step differences of the interior state
void launch_cfl_norm(const Grid &G, int model, double gamma, double dt, const double *u, const double *u_prev,
                     const double *dxinv, double *partials, int nblocks, hipStream_t s);
int cfl_norm_blocks(const Grid &G);

// Sizes of an array in the reference's layout: every point with its ghosts,
// x fastest, the nv components of a point contiguous (ny = 1 in 1D).
struct AosGrid {
  int ndims, nv, g, nx, ny;
};
// device-buffer operators on reference-layout arrays (see kernels.hip)
void launch_dev_flux(const AosGrid &A, int model, double gamma, int dir, const double *u, double *f, hipStream_t s);
void launch_dev_fd(const AosGrid &A, int dir, int par4, int bias, const double *f, double *df, hipStream_t s);
void launch_dev_weights(const AosGrid &A, int model, double gamma, int dir, int kind, int crweno,
                        int characteristic, int no_limiting, double eps, const double *fC, const double *u,
                        double *w1, double *w2, double *w3, long size, hipStream_t s);
// ghosted reference-layout array <-> planar natural / transposed copies
void launch_dev_planar(const Grid &G, const double *a, double *nat, double *tr, int to_planar, int ghosts_only,
                       double *out, hipStream_t s);
// one sweep's interface array: reference layout <-> line-major
void launch_dev_iface(int ndims, int dir, int nx, int N, int nl, int nv, const double *src, double *dst,
                      int to_lines, hipStream_t s);
// one kernel class of stage s over the plan's cases
void launch_class(const CaseDev *d_cases, const Plan &plan, int cls, int s, hipStream_t st);

// doubles of the long-line scratch of one sweep (SweepGeom::longl)
inline size_t lrow_doubles(const SweepGeom &S, int nv) { return (size_t)4 * (S.N + 1) * S.nlines * nv * 4; }

// per-operator kernels used by the unit-level entry points
// (systems longer than one team are solved one thread per system)
// b may be overwritten (with the pivots, for systems longer than one team)
bool launch_tridiag(const double *a, double *b, const double *c, double *x, int n, int ns, int *flag,
                    hipStream_t s);
void launch_weno_weights(int kind, const double *copt, const double *st, int count, double eps, double *w,
                         hipStream_t s);

}  // namespace piafs
