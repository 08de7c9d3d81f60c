/*
 * piafs_oracle.c -- CPU restatement of the PIAFS RHS / explicit RK path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Plain C99,
 * single threaded, reference AoS layout.  Every function names the reference
 * file:line it restates; the arithmetic is written in the same association
 * order as the reference so that gcc -O2 builds agree bit for bit.
 *
 * Supported (the hot path): models euler1d / navierstokes2d (+ photochemistry),
 * hyp_space_scheme weno5 / crweno5 / upw5, hyp_interp_type components and
 * characteristic (both models; characteristic CRWENO5 with the block solver), WENO weights JS / mapped / borges(Z) /
 * yc and no_limiting, upwinding roe / rf-char / llf-char / rusanov (both
 * models), every boundary type of boundary.inp (periodic, extrapolate,
 * dirichlet, reflect, sponge, slip and no-slip walls, subsonic inflow /
 * outflow / ambivalent, supersonic inflow / outflow), par_space_scheme 2 / 4,
 * RK 1fe/22/33/44/ssprk3.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "piafs_oracle.h"

#define MAXS 256
#define NMAX 64

typedef struct {
  char   type[MAXS];
  int    dim, face;
  double xmin[3], xmax[3];
  int    is[3], ie[3];
  int    on;
  /* type-specific data, InitializeBoundaries.c:95-150 */
  double rho, vel[3], p, gamma;
  double vals[NMAX];  /* dirichlet / sponge values, or inflow passive scalars */
} obc;

struct orc_case {
  /* solver.inp (ReadInputs.c:200-430) */
  int    ndims, nvars, ghosts, dim[3], n_iter, restart_iter, screen_op_iter;
  char   time_scheme[MAXS], time_scheme_type[MAXS], hyp[MAXS], interp[MAXS];
  char   par[MAXS], model[MAXS], ip_file_type[MAXS];
  double dt;
  int    nbc;
  obc   *bc;
  /* derived sizes */
  int    npts;      /* ghosted points */
  int    npts_int;  /* interior points */
  int    stride[3];
  double *x, *dxinv, *u;
  /* physics */
  int    model_id;  /* 1 euler1d, 2 navierstokes2d */
  double gamma, Re, Pr, Tref, T0, TS, TA, TB;
  char   upw[MAXS];
  int    include_chem;
  /* weno.inp (WENOInitialize.c:46-123) */
  int    mapped, borges, yc, no_limiting;
  double eps;
  int    is_crweno, is_upw5, is_char, par4;
  /* RK tableau (TimeExplicitRKInitialize.c:20-95) */
  int    nstages;
  double A[16], b[4], c[4];
  /* chemistry (ChemistryInitialize.c:85-477) */
  int    nflow, z_i, nz_layers, z_stride, grid_stride, n_react;
  double ch_gamma, ch_R, Ti, n_O2, sO3, dz, I0, c_light, h_planck, nu;
  double t_start_norm, t_pulse_norm;
  double k0a, k0b, k1a, k1b, k2a, k2b, k3a, k3b, k4, k5, k6;
  double q0a, q0b, q1a, q1b, q2a, q2b, q3a, q3b, q4, q5, q6;
  double *imap, *nv_hnu;
  /* time */
  double waqt;
  /* work arrays */
  double *U[4], *Udot[4], *hyp_a, *par_a, *src_a, *fluxC, *fluxI;
  double *uL, *uR, *fL, *fR;
  double *Q, *QdX, *QdY, *FV, *FD;
  /* optional capture of the interface fluxes of every line, per dimension
   * ([line][interface][var]), set by orc_rhs_parts */
  double *fi_cap[2];
  /* optional capture, for dimension cap_dir (-1: none), of the interface
   * values uL, uR, fL, fR of every line ([line][interface][var]) and of the
   * nonlinear weights they were made with (four sets LF, LU, RF, RU at
   * w + k*w_size, interface point p, variable v at p*nvars + v), set by
   * orc_rhs_interfaces; wcap_p0 + k*wcap_pk is the interface point of
   * interface k of the line being reconstructed */
  double *if_cap[4], *w_cap[3];
  long   w_size, wcap_p0, wcap_pk;
  int    cap_dir, wcap_on;
  /* derivative ghosts of the parabolic term exchanged along periodic
   * dimensions, as a run with more than one rank along them does
   * (orc_set_derivative_exchange) */
  int    deriv_exchange;
};

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */

static int read_word(FILE *f, char *w) { return fscanf(f, "%255s", w) == 1; }

/* index of a ghosted point: x fastest (arrayfunctions.h:43-50) */
static int gidx(const orc_case *c, const int *i)
{
  int d, p = i[c->ndims - 1] + c->ghosts;
  for (d = c->ndims - 2; d >= 0; d--) p = p * (c->dim[d] + 2 * c->ghosts) + (i[d] + c->ghosts);
  return p;
}

/* ------------------------------------------------------------------ */
/* input decks                                                          */
/* ------------------------------------------------------------------ */

/* solver.inp -- ReadInputs.c:200-480 (keywords, defaults) */
static int read_solver(orc_case *c, const char *dir)
{
  char fn[1024], w[MAXS];
  FILE *f;
  c->ndims = 1; c->nvars = 1; c->ghosts = 1; c->dt = 0; c->n_iter = 0; c->restart_iter = 0;
  c->screen_op_iter = 1;
  strcpy(c->time_scheme, "euler"); strcpy(c->time_scheme_type, " ");
  strcpy(c->hyp, "1"); strcpy(c->par, "2"); strcpy(c->interp, "characteristic");
  strcpy(c->ip_file_type, "ascii"); strcpy(c->model, "none");
  snprintf(fn, sizeof fn, "%s/solver.inp", dir);
  f = fopen(fn, "r");
  if (!f) return 1;
  if (!read_word(f, w) || strcmp(w, "begin")) { fclose(f); return 1; }
  while (read_word(f, w) && strcmp(w, "end")) {
    int ok = 1;
    if      (!strcmp(w, "ndims"))  ok = fscanf(f, "%d", &c->ndims) == 1;
    else if (!strcmp(w, "nvars"))  ok = fscanf(f, "%d", &c->nvars) == 1;
    else if (!strcmp(w, "size"))   { int d; for (d = 0; d < c->ndims; d++) ok &= fscanf(f, "%d", &c->dim[d]) == 1; }
    else if (!strcmp(w, "ghost"))  ok = fscanf(f, "%d", &c->ghosts) == 1;
    else if (!strcmp(w, "n_iter")) ok = fscanf(f, "%d", &c->n_iter) == 1;
    else if (!strcmp(w, "restart_iter")) ok = fscanf(f, "%d", &c->restart_iter) == 1;
    else if (!strcmp(w, "time_scheme")) ok = read_word(f, c->time_scheme);
    else if (!strcmp(w, "time_scheme_type")) ok = read_word(f, c->time_scheme_type);
    else if (!strcmp(w, "hyp_space_scheme")) ok = read_word(f, c->hyp);
    else if (!strcmp(w, "hyp_interp_type")) ok = read_word(f, c->interp);
    else if (!strcmp(w, "par_space_scheme")) ok = read_word(f, c->par);
    else if (!strcmp(w, "dt")) ok = fscanf(f, "%lf", &c->dt) == 1;
    else if (!strcmp(w, "screen_op_iter")) ok = fscanf(f, "%d", &c->screen_op_iter) == 1;
    else if (!strcmp(w, "ip_file_type")) ok = read_word(f, c->ip_file_type);
    else if (!strcmp(w, "model")) ok = read_word(f, c->model);
    else if (!strcmp(w, "input_mode") || !strcmp(w, "output_mode")) {
      char v[MAXS]; ok = read_word(f, v);
      if (ok && strcmp(v, "serial")) { int n; ok = fscanf(f, "%d", &n) == 1; }
    } else { char v[MAXS]; ok = read_word(f, v); } /* other keywords: skip one token */
    if (!ok) { fclose(f); return 1; }
  }
  fclose(f);
  if (c->screen_op_iter <= 0) c->screen_op_iter = 1;
  return 0;
}

/* keyword file reader for physics.inp / weno.inp / chemistry.inp */
typedef struct { const char *key; char type; void *dst; } kw;

static int read_kw_file(const char *dir, const char *name, const kw *tab, int ntab)
{
  char fn[1024], w[MAXS];
  FILE *f;
  snprintf(fn, sizeof fn, "%s/%s", dir, name);
  f = fopen(fn, "r");
  if (!f) return 0; /* optional: defaults */
  if (!read_word(f, w) || strcmp(w, "begin")) { fclose(f); return 1; }
  while (read_word(f, w) && strcmp(w, "end")) {
    int k, hit = 0, ok = 1;
    for (k = 0; k < ntab; k++) {
      if (strcmp(w, tab[k].key)) continue;
      hit = 1;
      if (tab[k].type == 'd') ok = fscanf(f, "%lf", (double *)tab[k].dst) == 1;
      else if (tab[k].type == 'i') ok = fscanf(f, "%d", (int *)tab[k].dst) == 1;
      else ok = read_word(f, (char *)tab[k].dst);
      break;
    }
    if (!hit) { char v[MAXS]; ok = read_word(f, v); }
    if (!ok) { fclose(f); return 1; }
  }
  fclose(f);
  return 0;
}

/* boundary.inp -- InitializeBoundaries.c:60-200.  Type-specific values are
 * read as the reference reads them (a missing value leaves zero). */
static const char *bc_types[] = {"periodic", "extrapolate", "dirichlet", "reflect", "sponge", "noslip-wall",
                                 "slip-wall", "subsonic-inflow", "subsonic-outflow", "subsonic-ambivalent",
                                 "supersonic-inflow", "supersonic-outflow"};
static int read_boundary(orc_case *c, const char *dir)
{
  char fn[1024];
  FILE *f;
  int n, d, v, known;
  snprintf(fn, sizeof fn, "%s/boundary.inp", dir);
  f = fopen(fn, "r");
  if (!f) return 1;
  if (fscanf(f, "%d", &c->nbc) != 1) { fclose(f); return 1; }
  c->bc = (obc *)calloc((size_t)c->nbc, sizeof(obc));
  for (n = 0; n < c->nbc; n++) {
    obc *b = &c->bc[n];
    if (!read_word(f, b->type)) { fclose(f); return 1; }
    if (fscanf(f, "%d %d", &b->dim, &b->face) != 2) { fclose(f); return 1; }
    for (d = 0; d < c->ndims; d++)
      if (fscanf(f, "%lf %lf", &b->xmin[d], &b->xmax[d]) != 2) { fclose(f); return 1; }
    for (known = 0, v = 0; v < 12; v++) known |= !strcmp(b->type, bc_types[v]);
    if (!known) {
      fprintf(stderr, "oracle: boundary type %s not supported\n", b->type);
      fclose(f);
      return 1;
    }
    if (!strcmp(b->type, "dirichlet") || !strcmp(b->type, "sponge"))
      for (v = 0; v < c->nvars; v++) if (fscanf(f, "%lf", &b->vals[v]) != 1) break;
    if (!strcmp(b->type, "slip-wall") || !strcmp(b->type, "noslip-wall"))
      for (d = 0; d < c->ndims; d++) if (fscanf(f, "%lf", &b->vel[d]) != 1) break;
    if (!strcmp(b->type, "subsonic-inflow")) {
      if (fscanf(f, "%lf", &b->rho) != 1) {}
      for (d = 0; d < c->ndims; d++) if (fscanf(f, "%lf", &b->vel[d]) != 1) break;
      for (v = 0; v < c->nvars - (2 + c->ndims); v++) if (fscanf(f, "%lf", &b->vals[v]) != 1) break;
    }
    if (!strcmp(b->type, "subsonic-outflow")) { if (fscanf(f, "%lf", &b->p) != 1) {} }
    if (!strcmp(b->type, "subsonic-ambivalent")) {
      if (fscanf(f, "%lf", &b->rho) != 1) {}
      for (d = 0; d < c->ndims; d++) if (fscanf(f, "%lf", &b->vel[d]) != 1) break;
      if (fscanf(f, "%lf", &b->p) != 1) {}
    }
    if (!strcmp(b->type, "supersonic-inflow")) {
      if (fscanf(f, "%lf", &b->rho) != 1) {}
      for (d = 0; d < c->ndims; d++) if (fscanf(f, "%lf", &b->vel[d]) != 1) break;
      if (fscanf(f, "%lf", &b->p) != 1) {}
      for (v = 0; v < c->nvars - (2 + c->ndims); v++) if (fscanf(f, "%lf", &b->vals[v]) != 1) break;
    }
  }
  fclose(f);
  return 0;
}

/* FindInterval.c:22-45 */
static void find_interval(double a, double b, const double *x, int N, int *imin, int *imax)
{
  int i;
  double min_dx = x[1] - x[0], tol;
  *imax = -1; *imin = N;
  for (i = 2; i < N; i++) { double dx = x[i] - x[i - 1]; if (dx < min_dx) min_dx = dx; }
  tol = 1e-10 * min_dx;
  for (i = 0; i < N; i++) if (x[i] <= (b + tol)) *imax = i + 1;
  for (i = N - 1; i > -1; i--) if (x[i] >= (a - tol)) *imin = i;
}

/* CalculateLocalExtent, InitializeBoundaries.c:296-377 (serial: ip == 0 ==
 * iproc-1 so both faces are on this process) */
static void boundary_extents(orc_case *c)
{
  int n, d;
  for (n = 0; n < c->nbc; n++) {
    obc *b = &c->bc[n];
    int off = 0;
    b->on = 1;
    for (d = 0; d < c->ndims; d++) {
      if (!strcmp(b->type, "sponge")) {  /* the zone is interior points in every dimension */
        find_interval(b->xmin[d], b->xmax[d], &c->x[off + c->ghosts], c->dim[d], &b->is[d], &b->ie[d]);
        if (b->ie[d] - b->is[d] <= 0) b->on = 0;
      } else if (d == b->dim) {
        if (b->face == 1) { b->is[d] = -c->ghosts; b->ie[d] = 0; }
        else { b->is[d] = c->dim[d]; b->ie[d] = c->dim[d] + c->ghosts; }
      } else {
        find_interval(b->xmin[d], b->xmax[d], &c->x[off + c->ghosts], c->dim[d], &b->is[d], &b->ie[d]);
        if (b->ie[d] - b->is[d] <= 0) b->on = 0;
      }
      off += c->dim[d] + 2 * c->ghosts;
    }
  }
}

/* ReadArraySerial (ReadArray.c:153-300): ascii (var-major) or binary
 * (point-major).  x may be NULL.  Fills the interior of u (ghosted). */
static int read_array(orc_case *c, const char *dir, const char *root, int nv, double *x, double *u, int *found)
{
  char fn[1024];
  FILE *f;
  int d, i, size = 1, xs = 0;
  double *ug, *xg;
  int bin = (!strcmp(c->ip_file_type, "bin") || !strcmp(c->ip_file_type, "binary"));
  *found = 0;
  snprintf(fn, sizeof fn, "%s/%s.inp", dir, root);
  f = fopen(fn, bin ? "rb" : "r");
  if (!f) return 0;
  *found = 1;
  for (d = 0; d < c->ndims; d++) { size *= c->dim[d]; xs += c->dim[d]; }
  ug = (double *)calloc((size_t)size * nv, sizeof(double));
  xg = (double *)calloc((size_t)xs, sizeof(double));
  if (bin) {
    if ((int)fread(xg, sizeof(double), (size_t)xs, f) != xs) { fclose(f); return 1; }
    if ((int)fread(ug, sizeof(double), (size_t)size * nv, f) != size * nv) { fclose(f); return 1; }
  } else {
    for (i = 0; i < xs; i++) if (fscanf(f, "%lf", &xg[i]) != 1) { fclose(f); return 1; }
    for (d = 0; d < nv; d++)
      for (i = 0; i < size; i++) if (fscanf(f, "%lf", &ug[i * nv + d]) != 1) { fclose(f); return 1; }
  }
  fclose(f);
  /* scatter into the ghosted array */
  {
    int idx[3] = {0, 0, 0}, p, q;
    for (q = 0; q < size; q++) {
      int r = q;
      for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
      p = gidx(c, idx);
      for (i = 0; i < nv; i++) u[p * nv + i] = ug[q * nv + i];
    }
  }
  if (x) {
    int off = 0, og = 0, g = c->ghosts;
    for (d = 0; d < c->ndims; d++) {
      double *X = &x[off];
      int N = c->dim[d];
      for (i = 0; i < N; i++) X[i + g] = xg[og + i];
      /* ghost coordinates, ReadArray.c:71-91 */
      for (i = 0; i < g; i++) { int delta = g - i; X[i] = X[g] + ((double)delta) * (X[g] - X[g + 1]); }
      for (i = N + g; i < N + 2 * g; i++) {
        int delta = i - (N + g - 1);
        X[i] = X[N + g - 1] + ((double)delta) * (X[N + g - 1] - X[N + g - 2]);
      }
      off += N + 2 * g;
      og += N;
    }
  }
  free(ug); free(xg);
  return 0;
}

/* ------------------------------------------------------------------ */
/* initialisation                                                       */
/* ------------------------------------------------------------------ */

/* TimeExplicitRKInitialize.c:20-95 */
static int rk_init(orc_case *c)
{
  memset(c->A, 0, sizeof c->A); memset(c->b, 0, sizeof c->b); memset(c->c, 0, sizeof c->c);
  if (!strcmp(c->time_scheme, "euler")) { c->nstages = 1; c->b[0] = 1.0; return 0; }
  if (strcmp(c->time_scheme, "rk")) return 1;
  if (!strcmp(c->time_scheme_type, "1fe")) { c->nstages = 1; c->b[0] = 1.0; }
  else if (!strcmp(c->time_scheme_type, "22")) {
    c->nstages = 2; c->A[2] = 1.0; c->c[1] = 1.0; c->b[0] = c->b[1] = 0.5;
  } else if (!strcmp(c->time_scheme_type, "33")) {
    c->nstages = 3;
    c->A[3] = 2.0 / 3.0; c->A[6] = 2.0 / 3.0 - 1.0 / 4.0; c->A[7] = 1.0 / 4.0;
    c->c[1] = 2.0 / 3.0; c->c[2] = 2.0 / 3.0;
    c->b[0] = 1.0 / 4.0; c->b[1] = -1.0 / 4.0; c->b[2] = 1.0;
  } else if (!strcmp(c->time_scheme_type, "44")) {
    c->nstages = 4; c->A[4] = 0.5; c->A[9] = 0.5; c->A[14] = 1.0;
    c->c[1] = c->c[2] = 0.5; c->c[3] = 1.0;
    c->b[0] = 1.0 / 6.0; c->b[1] = 1.0 / 3.0; c->b[2] = 1.0 / 3.0; c->b[3] = 1.0 / 6.0;
  } else if (!strcmp(c->time_scheme_type, "ssprk3") || !strcmp(c->time_scheme_type, "tvdrk3")) {
    c->nstages = 3; c->A[3] = 1.0; c->A[6] = 0.25; c->A[7] = 0.25;
    c->c[1] = 1.0; c->c[2] = 0.5;
    c->b[0] = c->b[1] = 1.0 / 6.0; c->b[2] = 2.0 / 3.0;
  } else return 1;
  return 0;
}

/* ChemistryInitialize.c:85-353 -- constants and normalisation */
static int chem_init(orc_case *c, const char *dir)
{
  double pi = 4.0 * atan(1.0), NA = 6.02214076e23, kB = 1.380649e-23, e = 1.60217663e-19;
  double M_O2 = 0.032, M_CO2 = 0.044, Cp_O2 = 918.45, Cv_O2 = 650.0, Cp_CO2 = 842.86, Cv_CO2 = 654.5;
  double mu0_O2 = 1.99e-5, kappa0_O2 = 2.70e-2, mu0_CO2 = 1.45e-5, kappa0_CO2 = 1.58e-2;
  double lambda_UV = 2.48e-7, theta = 0.17 * pi / 180, f_CO2 = 0.0, f_O3 = 0.005, Ptot = 101325.0;
  double Ti = 288, t_start = 0.0, t_pulse = 10 * 1e-9, Lz = 0.03, z_mm = 0.0;
  int nz = 22;
  double k0a = 0.9 * 3.3e-13, k0b = 0.1 * 3.3e-13, k1a = 0.8 * 3.95e-17, k1b = 0.2 * 3.95e-17;
  double k2a = 1.2e-16, k2b = 1.2e-16, k3a = 1.2e-17, k3b = 1.0e-17, k4 = 1.1e-16, k5 = 2.0e-16;
  double k6 = 0.2 * 3.0e-17, F0 = 2000, sO3 = 1.1e-21, IA = 1.0, IB = 1.0, IC = 0.0;
  char wz[MAXS] = "yes";
  kw tab[] = {
    {"lambda_UV", 'd', &lambda_UV}, {"t_start", 'd', &t_start}, {"t_pulse", 'd', &t_pulse},
    {"theta", 'd', &theta}, {"f_CO2", 'd', &f_CO2}, {"f_O3", 'd', &f_O3}, {"Ptot", 'd', &Ptot},
    {"Ti", 'd', &Ti}, {"Lz", 'd', &Lz}, {"z_mm", 'd', &z_mm}, {"nz", 'i', &nz},
    {"k0a", 'd', &k0a}, {"k1a", 'd', &k1a}, {"k2a", 'd', &k2a}, {"k3a", 'd', &k3a},
    {"k0b", 'd', &k0b}, {"k1b", 'd', &k1b}, {"k2b", 'd', &k2b}, {"k3b", 'd', &k3b},
    {"k4", 'd', &k4}, {"k5", 'd', &k5}, {"k6", 'd', &k6}, {"F0", 'd', &F0}, {"sO3", 'd', &sO3},
    {"IA", 'd', &IA}, {"IB", 'd', &IB}, {"IC", 'd', &IC}, {"write_all_zlocs", 's', wz},
  };
  double f_O2, R, Cp, Cv, gam, mu0, kappa0, kUV, kg, n_O2, cs, L_ref, v_ref, t_ref, rho_ref, P_ref;
  double rate, Ei;
  int found, d, p, N;
  if (read_kw_file(dir, "chemistry.inp", tab, (int)(sizeof tab / sizeof tab[0]))) return 1;
  if (f_CO2 > 1.0 || f_CO2 < 0.0 || f_O3 > 1.0 || f_O3 < 0.0) return 1;
  f_O2 = 1.0 - f_CO2;
  R = NA * kB / (f_O2 * M_O2 + f_CO2 * M_CO2);
  Cp = f_O2 * Cp_O2 + f_CO2 * Cp_CO2;
  Cv = f_O2 * Cv_O2 + f_CO2 * Cv_CO2;
  gam = Cp / Cv;
  mu0 = f_O2 * mu0_O2 + f_CO2 * mu0_CO2;
  kappa0 = f_O2 * kappa0_O2 + f_CO2 * kappa0_CO2;
  kUV = 2 * pi / lambda_UV;
  kg = 2 * kUV * sin(theta);
  n_O2 = f_O2 * Ptot / (kB * Ti);
  cs = sqrt(gam * R * Ti);
  L_ref = 1.0 / kg; v_ref = cs; t_ref = L_ref / v_ref;
  rho_ref = Ptot / (R * Ti); P_ref = rho_ref * v_ref * v_ref;
  c->t_start_norm = t_start / t_ref;
  c->t_pulse_norm = t_pulse / t_ref;
  rate = kg * cs / n_O2;
  c->k0a = k0a / rate; c->k1a = k1a / rate; c->k2a = k2a / rate; c->k3a = k3a / rate;
  c->k0b = k0b / rate; c->k1b = k1b / rate; c->k2b = k2b / rate; c->k3b = k3b / rate;
  c->k4 = k4 / rate; c->k5 = k5 / rate; c->k6 = k6 / rate;
  c->I0 = F0 / t_pulse;
  c->c_light = 3.0e8; c->h_planck = 6.62607015e-34;
  c->nu = c->c_light / lambda_UV;
  Ei = P_ref / ((gam - 1.0) * n_O2);
  c->q0a = 0.73 * e / Ei; c->q0b = 2.8 * e / Ei; c->q1a = 0.29 * e / Ei; c->q1b = 0.84 * e / Ei;
  c->q2a = 4.3 * e / Ei; c->q2b = 0.81 * e / Ei; c->q3a = 0.42 * e / Ei; c->q3b = 1.16 * e / Ei;
  c->q4 = 1.23 * e / Ei; c->q5 = 1.81 * e / Ei; c->q6 = 0.42 * e / Ei;
  c->dz = Lz / nz;
  c->z_i = (int)(ceil(z_mm * 0.001 * nz / Lz));
  c->ch_gamma = gam; c->ch_R = R; c->Ti = Ti; c->n_O2 = n_O2; c->sO3 = sO3;
  c->nz_layers = c->z_i + 1;
  c->n_react = 7 * c->nz_layers;
  c->z_stride = 7;
  c->grid_stride = c->nflow + c->n_react;
  /* imap (ChemistryInitialize.c:444-477) */
  c->imap = (double *)calloc((size_t)c->npts, sizeof(double));
  c->nv_hnu = (double *)calloc((size_t)c->npts * c->nz_layers, sizeof(double));
  if (read_array(c, dir, "imap", 1, NULL, c->imap, &found)) return 1;
  if (!found) {
    double x0 = c->x[c->ghosts];
    int idx[3] = {0, 0, 0}, q;
    N = c->npts_int;
    for (q = 0; q < N; q++) {
      int r = q; double xx;
      for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
      p = gidx(c, idx);
      xx = c->x[c->ghosts + idx[0]];
      c->imap[p] = IA + IB * cos(kg * L_ref * (xx - x0) * (1.0 - IC * (xx - x0)));
    }
  }
  /* overrides in the flow models: NavierStokes2DInitialize.c:188-216,
   * Euler1DInitialize.c:150-158 */
  c->gamma = gam;
  if (c->model_id == 2) {
    c->Tref = Ti;
    if (c->Re > 0) {
      double mu_ref = mu0 * exp(1.5 * log(Ti / c->T0)) * ((c->T0 + c->TS) / (Ti + c->TS));
      double kappa_ref = kappa0 * exp(1.5 * log(Ti / c->T0))
                       * ((c->T0 + c->TA * exp(-c->TB / c->T0)) / (Ti + c->TA * exp(-c->TB / Ti)));
      c->Re = rho_ref * v_ref * L_ref / mu_ref;
      c->Pr = gam * R / (gam - 1) * mu_ref / kappa_ref;
    }
  }
  return 0;
}

static void *xcalloc(size_t n) { return calloc(n, sizeof(double)); }

int orc_load(const char *dir, orc_case **out)
{
  orc_case *c = (orc_case *)calloc(1, sizeof(orc_case));
  int d, i, found, s;
  char include_chem[MAXS] = "no";
  *out = NULL;
  c->cap_dir = -1;
  if (read_solver(c, dir)) { orc_free(c); return 1; }
  if (c->ndims < 1 || c->ndims > 2 || c->ghosts < 3) { orc_free(c); return 2; }
  /* sizes (Initialize.c:121-175) */
  c->npts = 1; c->npts_int = 1;
  for (d = 0; d < c->ndims; d++) { c->npts *= c->dim[d] + 2 * c->ghosts; c->npts_int *= c->dim[d]; }
  c->stride[0] = 1;
  for (d = 1; d < c->ndims; d++) c->stride[d] = c->stride[d - 1] * (c->dim[d - 1] + 2 * c->ghosts);
  {
    int xs = 0; for (d = 0; d < c->ndims; d++) xs += c->dim[d] + 2 * c->ghosts;
    c->x = (double *)xcalloc((size_t)xs);
    c->dxinv = (double *)xcalloc((size_t)xs);
  }
  c->u = (double *)xcalloc((size_t)c->npts * c->nvars);
  /* InitialSolution.c:24-114 */
  if (read_array(c, dir, "initial", c->nvars, c->x, c->u, &found) || !found) { orc_free(c); return 3; }
  {
    int off = 0, g = c->ghosts;
    for (d = 0; d < c->ndims; d++) {
      for (i = 0; i < c->dim[d]; i++)
        c->dxinv[i + off + g] = 2.0 / (c->x[i + 1 + off + g] - c->x[i - 1 + off + g]);
      for (i = 0; i < g; i++) c->dxinv[off + i] = c->dxinv[off + g];
      for (i = c->dim[d] + g; i < c->dim[d] + 2 * g; i++) c->dxinv[off + i] = c->dxinv[off + c->dim[d] + g - 1];
      off += c->dim[d] + 2 * g;
    }
  }
  if (read_boundary(c, dir)) { orc_free(c); return 4; }
  boundary_extents(c);
  /* InitializeSolvers.c:57-212 */
  if (!strcmp(c->par, "4")) c->par4 = 1;
  else if (!strcmp(c->par, "2")) c->par4 = 0;
  else { orc_free(c); return 5; }
  c->is_upw5 = 0;
  if (!strcmp(c->hyp, "weno5")) c->is_crweno = 0;
  else if (!strcmp(c->hyp, "crweno5")) c->is_crweno = 1;
  else if (!strcmp(c->hyp, "upw5")) { c->is_crweno = 0; c->is_upw5 = 1; }
  else { orc_free(c); return 5; }
  if (strcmp(c->interp, "characteristic") && strcmp(c->interp, "components")) { orc_free(c); return 5; }
  c->is_char = (c->nvars > 1 && !strcmp(c->interp, "characteristic"));
  c->mapped = c->borges = c->yc = c->no_limiting = 0; c->eps = 1e-6;
  {
    double p = 2.0, rc = 0.3, xi = 0.001, tol = 1e-16;
    kw tab[] = {{"mapped", 'i', &c->mapped}, {"borges", 'i', &c->borges}, {"yc", 'i', &c->yc},
                {"no_limiting", 'i', &c->no_limiting}, {"epsilon", 'd', &c->eps}, {"p", 'd', &p},
                {"rc", 'd', &rc}, {"xi", 'd', &xi}, {"tol", 'd', &tol}};
    /* upw5 reads no weno.inp (InitializeSolvers.c:141-148) */
    if (!c->is_upw5 && read_kw_file(dir, "weno.inp", tab, 9)) { orc_free(c); return 6; }
  }
  if (rk_init(c)) { orc_free(c); return 7; }
  /* physics (Euler1DInitialize.c:58-170, NavierStokes2DInitialize.c:68-235) */
  c->gamma = 1.4; c->Pr = 0.72; c->Re = -1; c->Tref = 273.15; c->T0 = 275.0; c->TS = 110.4;
  c->TA = 245.4; c->TB = 27.6;
  strcpy(c->upw, "roe");
  if (!strcmp(c->model, "euler1d")) {
    kw tab[] = {{"gamma", 'd', &c->gamma}, {"upwinding", 's', c->upw}, {"include_chemistry", 's', include_chem}};
    c->model_id = 1; c->nflow = 3;
    if (c->ndims != 1) { orc_free(c); return 8; }
    if (read_kw_file(dir, "physics.inp", tab, 3)) { orc_free(c); return 8; }
  } else if (!strcmp(c->model, "navierstokes2d")) {
    char wo[MAXS];
    kw tab[] = {{"gamma", 'd', &c->gamma}, {"upwinding", 's', c->upw}, {"Pr", 'd', &c->Pr},
                {"Re", 'd', &c->Re}, {"T_ref", 'd', &c->Tref}, {"T_0", 'd', &c->T0},
                {"T_S", 'd', &c->TS}, {"T_A", 'd', &c->TA}, {"T_B", 'd', &c->TB},
                {"include_chemistry", 's', include_chem}, {"write_output", 's', wo}};
    c->model_id = 2; c->nflow = 4;
    if (c->ndims != 2) { orc_free(c); return 8; }
    if (read_kw_file(dir, "physics.inp", tab, 11)) { orc_free(c); return 8; }
  } else { orc_free(c); return 8; }
  if (strcmp(c->upw, "roe") && strcmp(c->upw, "rf-char") && strcmp(c->upw, "llf-char")
      && strcmp(c->upw, "rusanov")) { orc_free(c); return 8; }
  /* the zones' gamma: NavierStokes2DInitialize.c:184 sets the physics.inp value
     before chemistry replaces it; Euler1DInitialize never sets it */
  for (s = 0; s < c->nbc; s++) c->bc[s].gamma = (c->model_id == 2) ? c->gamma : 0.0;
  c->include_chem = !strcmp(include_chem, "yes");
  if (c->include_chem) { if (chem_init(c, dir)) { orc_free(c); return 9; } }
  if (c->nvars != c->nflow + (c->include_chem ? c->n_react : 0)) { orc_free(c); return 10; }
  /* time (TimeInitialize.c:51) and work arrays */
  c->waqt = (double)c->restart_iter * c->dt;
  {
    size_t n = (size_t)c->npts * c->nvars;
    for (s = 0; s < c->nstages; s++) { c->U[s] = (double *)xcalloc(n); c->Udot[s] = (double *)xcalloc(n); }
    c->hyp_a = (double *)xcalloc(n); c->par_a = (double *)xcalloc(n); c->src_a = (double *)xcalloc(n);
    c->fluxC = (double *)xcalloc(n);
    c->fluxI = (double *)xcalloc(n * 2); c->uL = (double *)xcalloc(n * 2); c->uR = (double *)xcalloc(n * 2);
    c->fL = (double *)xcalloc(n * 2); c->fR = (double *)xcalloc(n * 2);
    c->Q = (double *)xcalloc(n); c->QdX = (double *)xcalloc(n); c->QdY = (double *)xcalloc(n);
    c->FV = (double *)xcalloc(n); c->FD = (double *)xcalloc(n);
  }
  *out = c;
  return 0;
}

void orc_free(orc_case *c)
{
  int s;
  if (!c) return;
  free(c->bc); free(c->x); free(c->dxinv); free(c->u); free(c->imap); free(c->nv_hnu);
  for (s = 0; s < 4; s++) { free(c->U[s]); free(c->Udot[s]); }
  free(c->hyp_a); free(c->par_a); free(c->src_a); free(c->fluxC); free(c->fluxI);
  free(c->uL); free(c->uR); free(c->fL); free(c->fR);
  free(c->Q); free(c->QdX); free(c->QdY); free(c->FV); free(c->FD);
  free(c);
}

int orc_ndims(const orc_case *c) { return c->ndims; }
int orc_nvars(const orc_case *c) { return c->nvars; }
int orc_dim(const orc_case *c, int d) { return c->dim[d]; }
int orc_ghosts(const orc_case *c) { return c->ghosts; }
double orc_time(const orc_case *c) { return c->waqt; }
double orc_dt(const orc_case *c) { return c->dt; }
int orc_n_iter(const orc_case *c) { return c->n_iter; }
double orc_param(const orc_case *c, const char *name)
{
  if (!strcmp(name, "gamma")) return c->gamma;
  if (!strcmp(name, "Re")) return c->Re;
  if (!strcmp(name, "Pr")) return c->Pr;
  if (!strcmp(name, "Tref")) return c->Tref;
  if (!strcmp(name, "t_pulse_norm")) return c->t_pulse_norm;
  if (!strcmp(name, "k0a")) return c->k0a;
  if (!strcmp(name, "q0a")) return c->q0a;
  if (!strcmp(name, "I0")) return c->I0;
  return NAN;
}

/* ------------------------------------------------------------------ */
/* boundary conditions, applied zone by zone in list order            */
/* (ApplyBoundaryConditions.c:20-40): BCPeriodic.c:22-62,              */
/* BCExtrapolate.c:16-51, BCDirichlet.c:16-45, BCReflect.c:17-50,      */
/* BCSlipWall.c:24-135, BCNoslipWall.c:22-80, BCSubsonicInflow.c:22-90,*/
/* BCSubsonicOutflow.c:21-78, BCSubsonicAmbivalent.c:30-125,           */
/* BCSupersonicInflow.c:26-80, BCSupersonicOutflow.c:25-80.  A sponge  */
/* zone acts through the source term (BCSponge.c:29-74).               */
/* ------------------------------------------------------------------ */

/* _NavierStokes2DGetFlowVar_, navierstokes2d.h:70-79 */
static void ns2d_flowvar(const double *u, double gam, double *rho, double *vx, double *vy, double *e, double *P)
{
  double vsq;
  *rho = u[0];
  *vx = (*rho == 0) ? 0 : u[1] / *rho;
  *vy = (*rho == 0) ? 0 : u[2] / *rho;
  *e = u[3];
  vsq = (*vx * *vx) + (*vy * *vy);
  *P = (*e - 0.5 * *rho * vsq) * (gam - 1.0);
}

static void apply_bcs(orc_case *c, double *phi)
{
  int n, nv = c->nvars, g = c->ghosts;
  for (n = 0; n < c->nbc; n++) {
    obc *b = &c->bc[n];
    const char *t = b->type;
    int d = b->dim, bounds[3], i1[3], i2[3], k, done = 0;
    int generic = !strcmp(t, "periodic") || !strcmp(t, "extrapolate") || !strcmp(t, "dirichlet")
                  || !strcmp(t, "reflect");
    if (!b->on || !strcmp(t, "sponge")) continue;
    /* the flow BCs exist for ndims 2 only, slip-wall also for ndims 1 */
    if (!generic && c->ndims != 2 && !(c->ndims == 1 && !strcmp(t, "slip-wall"))) continue;
    for (k = 0; k < c->ndims; k++) { bounds[k] = b->ie[k] - b->is[k]; i1[k] = 0; }
    while (!done) {
      int p1, p2, v, tmp[3];
      for (k = 0; k < c->ndims; k++) tmp[k] = i1[k] + b->is[k];
      p1 = gidx(c, tmp);
      if (!strcmp(t, "periodic")) {
        /* the source index is taken without the zone offset (BCPeriodic.c:44-52) */
        for (k = 0; k < c->ndims; k++) i2[k] = i1[k];
        if (b->face == 1) i2[d] = i1[d] + c->dim[d] - g;
        p2 = gidx(c, i2);
        for (v = 0; v < nv; v++) phi[nv * p1 + v] = phi[nv * p2 + v];
      } else if (!strcmp(t, "dirichlet")) {
        for (v = 0; v < nv; v++) phi[nv * p1 + v] = b->vals[v];
      } else {
        for (k = 0; k < c->ndims; k++) i2[k] = tmp[k];
        if (b->face == 1) i2[d] = g - 1 - i1[d];
        else i2[d] = c->dim[d] - i1[d] - 1;
        p2 = gidx(c, i2);
        if (!strcmp(t, "extrapolate")) {
          for (v = 0; v < nv; v++) phi[nv * p1 + v] = phi[nv * p2 + v];
        } else if (!strcmp(t, "reflect")) {
          for (v = 0; v < nv; v++) phi[nv * p1 + v] = (-1.0) * phi[nv * p2 + v];
        } else if (c->ndims == 1) {
          /* slip-wall, 1D (Euler1D never sets the zone's gamma: it stays 0) */
          double gam = b->gamma, igm1 = 1.0 / (gam - 1.0);
          double rho = phi[nv * p2], uvel = phi[nv * p2 + 1] / rho, energy = phi[nv * p2 + 2];
          double pressure = (energy - 0.5 * rho * uvel * uvel) * (gam - 1.0);
          double rg = rho, pg = pressure, ug = 2.0 * b->vel[0] - uvel;
          double eg = igm1 * pg + 0.5 * rg * ug * ug;
          phi[nv * p1 + 0] = rg;
          phi[nv * p1 + 1] = rg * ug;
          phi[nv * p1 + 2] = eg;
        } else {
          double gam = b->gamma, igm1 = 1.0 / (gam - 1.0);
          double rho, uvel, vvel, energy, pressure, rg, ug, vg, pg, eg;
          int scalars = 0;
          if (!strcmp(t, "subsonic-ambivalent")) {
            double nx = (d == 0) ? 1.0 : 0.0, ny = (d == 1) ? 1.0 : 0.0, u1, v1, u2, v2, vn;
            int ia[3], ib[3];
            nx *= (double)b->face; ny *= (double)b->face;
            for (k = 0; k < c->ndims; k++) ia[k] = ib[k] = tmp[k];
            ia[d] = (b->face == 1) ? 0 : c->dim[d] - 1;
            ib[d] = (b->face == 1) ? ia[d] + 1 : ia[d] - 1;
            ns2d_flowvar(phi + nv * gidx(c, ia), gam, &rho, &u1, &v1, &energy, &pressure);
            ns2d_flowvar(phi + nv * gidx(c, ib), gam, &rho, &u2, &v2, &energy, &pressure);
            vn = (1.5 * u1 - 0.5 * u2) * nx + (1.5 * v1 - 0.5 * v2) * ny;
            ns2d_flowvar(phi + nv * p2, gam, &rho, &uvel, &vvel, &energy, &pressure);
            if (vn > 0) { rg = b->rho; pg = pressure; ug = b->vel[0]; vg = b->vel[1]; }
            else { rg = rho; pg = b->p; ug = uvel; vg = vvel; }
          } else {
            ns2d_flowvar(phi + nv * p2, gam, &rho, &uvel, &vvel, &energy, &pressure);
            rg = rho; pg = pressure; ug = uvel; vg = vvel;  /* supersonic-outflow */
            if (!strcmp(t, "slip-wall")) {
              if (d == 0) ug = 2.0 * b->vel[0] - uvel;
              else vg = 2.0 * b->vel[1] - vvel;
            } else if (!strcmp(t, "noslip-wall")) {
              ug = 2.0 * b->vel[0] - uvel;
              vg = 2.0 * b->vel[1] - vvel;
            } else if (!strcmp(t, "subsonic-inflow")) {
              rg = b->rho; ug = b->vel[0]; vg = b->vel[1]; scalars = 1;
            } else if (!strcmp(t, "subsonic-outflow")) {
              pg = b->p;
            } else if (!strcmp(t, "supersonic-inflow")) {
              rg = b->rho; pg = b->p; ug = b->vel[0]; vg = b->vel[1]; scalars = 1;
            }
          }
          eg = igm1 * pg + 0.5 * rg * (ug * ug + vg * vg);
          phi[nv * p1 + 0] = rg;
          phi[nv * p1 + 1] = rg * ug;
          phi[nv * p1 + 2] = rg * vg;
          phi[nv * p1 + 3] = eg;
          if (scalars) for (v = 4; v < nv; v++) phi[nv * p1 + v] = b->vals[v - 4];
        }
      }
      for (k = 0; k < c->ndims; k++) {
        if (i1[k] == bounds[k] - 1) i1[k] = 0;
        else { i1[k]++; break; }
      }
      if (k == c->ndims) done = 1;
    }
  }
}

/* BCSpongeSource (BCSponge.c:29-74), called by SourceFunction.c:44-58
 * after the physics source, zone by zone */
static void sponge_source(orc_case *c, const double *u, double *src)
{
  int n, nv = c->nvars, g = c->ghosts;
  for (n = 0; n < c->nbc; n++) {
    obc *b = &c->bc[n];
    int d = b->dim, bounds[3], i1[3], k, v, done = 0, off = 0;
    if (!b->on || strcmp(b->type, "sponge")) continue;
    for (k = 0; k < d; k++) off += c->dim[k] + 2 * g;
    for (k = 0; k < c->ndims; k++) { bounds[k] = b->ie[k] - b->is[k]; i1[k] = 0; }
    while (!done) {
      int tmp[3], p;
      double x, sigma;
      for (k = 0; k < c->ndims; k++) tmp[k] = i1[k] + b->is[k];
      x = c->x[off + g + tmp[d]];
      if (b->face > 0) sigma = (x - b->xmin[d]) / (b->xmax[d] - b->xmin[d]);
      else sigma = (x - b->xmax[d]) / (b->xmin[d] - b->xmax[d]);
      p = gidx(c, tmp);
      for (v = 0; v < nv; v++) src[nv * p + v] -= (sigma * (u[nv * p + v] - b->vals[v]));
      for (k = 0; k < c->ndims; k++) {
        if (i1[k] == bounds[k] - 1) i1[k] = 0;
        else { i1[k]++; break; }
      }
      if (k == c->ndims) done = 1;
    }
  }
}

/* ------------------------------------------------------------------ */
/* physics: flux, Roe average, eigenstructure                          */
/* ------------------------------------------------------------------ */

/* Euler1DFlux.c:19 via euler1d.h:61-85 */
static void flux_euler1d(const orc_case *c, const double *u, double *f)
{
  double g = c->gamma, rho = u[0], v = u[1] / rho, e = u[2];
  double P = (e - 0.5 * rho * v * v) * (g - 1.0);
  int m;
  f[0] = rho * v;
  f[1] = rho * v * v + P;
  f[2] = (e + P) * v;
  for (m = 3; m < c->nvars; m++) f[m] = v * u[m];
}

/* NavierStokes2DFlux.c:24-52 via navierstokes2d.h:70-110 */
static void flux_ns2d(const orc_case *c, const double *u, double *f, int dir)
{
  double g = c->gamma, rho = u[0];
  double vx = (rho == 0) ? 0 : u[1] / rho, vy = (rho == 0) ? 0 : u[2] / rho, e = u[3];
  double vsq = (vx * vx) + (vy * vy);
  double P = (e - 0.5 * rho * vsq) * (g - 1.0);
  int m;
  if (dir == 0) {
    f[0] = rho * vx; f[1] = rho * vx * vx + P; f[2] = rho * vx * vy; f[3] = (e + P) * vx;
    for (m = 4; m < c->nvars; m++) f[m] = vx * u[m];
  } else {
    f[0] = rho * vy; f[1] = rho * vy * vx; f[2] = rho * vy * vy + P; f[3] = (e + P) * vy;
    for (m = 4; m < c->nvars; m++) f[m] = vy * u[m];
  }
}

/* euler1d.h:91-124 */
static void roeavg_euler1d(const orc_case *c, double *ua, const double *uL, const double *uR)
{
  double g = c->gamma;
  double rhoL = uL[0], vL = uL[1] / rhoL, eL = uL[2];
  double PL = (eL - 0.5 * rhoL * vL * vL) * (g - 1.0), cLsq = g * PL / rhoL;
  double HL = 0.5 * vL * vL + cLsq / (g - 1.0);
  double rhoR = uR[0], vR = uR[1] / rhoR, eR = uR[2];
  double PR = (eR - 0.5 * rhoR * vR * vR) * (g - 1.0), cRsq = g * PR / rhoR;
  double HR = 0.5 * vR * vR + cRsq / (g - 1.0);
  double tL = sqrt(rhoL), tR = sqrt(rhoR);
  double rho = tL * tR, v = (tL * vL + tR * vR) / (tL + tR), H = (tL * HL + tR * HR) / (tL + tR);
  double csq = (g - 1.0) * (H - 0.5 * v * v), P = csq * rho / g;
  double e = P / (g - 1.0) + 0.5 * rho * v * v;
  int m;
  ua[0] = rho; ua[1] = rho * v; ua[2] = e;
  for (m = 3; m < c->nvars; m++) ua[m] = sqrt(uL[m]) * sqrt(uR[m]);
}

/* euler1d.h:131-193: eigenvalues (diagonal D), L, R (row-major n x n) */
static void eig_euler1d(const orc_case *c, const double *u, double *D, double *L, double *R)
{
  int n = c->nvars, m;
  double g = c->gamma, rho = u[0], v = u[1] / rho, e = u[2];
  double P = (e - 0.5 * rho * v * v) * (g - 1.0), cs = sqrt(g * P / rho);
  if (D) {
    memset(D, 0, sizeof(double) * n * n);
    D[0] = v; D[1 * n + 1] = (v - cs); D[2 * n + 2] = (v + cs);
    for (m = 3; m < n; m++) D[m * n + m] = v;
  }
  if (L) {
    double k = (g - 1) / (rho * cs);
    memset(L, 0, sizeof(double) * n * n);
    L[1 * n + 0] = k * (-(v * v) / 2 - cs * v / (g - 1));
    L[1 * n + 1] = k * (v + cs / (g - 1));
    L[1 * n + 2] = k * (-1);
    L[0 * n + 0] = k * (rho * (-(v * v) / 2 + cs * cs / (g - 1)) / cs);
    L[0 * n + 1] = k * (rho * v / cs);
    L[0 * n + 2] = k * (-rho / cs);
    L[2 * n + 0] = k * ((v * v) / 2 - cs * v / (g - 1));
    L[2 * n + 1] = k * (-v + cs / (g - 1));
    L[2 * n + 2] = k * (1);
    for (m = 3; m < n; m++) L[m * n + m] = 1.0;
  }
  if (R) {
    memset(R, 0, sizeof(double) * n * n);
    R[0 * n + 1] = -rho / (2 * cs);
    R[1 * n + 1] = -rho * (v - cs) / (2 * cs);
    R[2 * n + 1] = -rho * ((v * v) / 2 + (cs * cs) / (g - 1) - cs * v) / (2 * cs);
    R[0 * n + 0] = 1; R[1 * n + 0] = v; R[2 * n + 0] = v * v / 2;
    R[0 * n + 2] = rho / (2 * cs);
    R[1 * n + 2] = rho * (v + cs) / (2 * cs);
    R[2 * n + 2] = rho * ((v * v) / 2 + (cs * cs) / (g - 1) + cs * v) / (2 * cs);
    for (m = 3; m < n; m++) R[m * n + m] = 1.0;
  }
}

/* navierstokes2d.h:115-160 */
static void roeavg_ns2d(const orc_case *c, double *ua, const double *uL, const double *uR)
{
  double g = c->gamma;
  double rhoL = uL[0], vxL = uL[1] / rhoL, vyL = uL[2] / rhoL, eL = uL[3];
  double vsqL = (vxL * vxL) + (vyL * vyL), PL = (eL - 0.5 * rhoL * vsqL) * (g - 1.0);
  double cLsq = g * PL / rhoL, HL = 0.5 * (vxL * vxL + vyL * vyL) + cLsq / (g - 1.0);
  double rhoR = uR[0], vxR = uR[1] / rhoR, vyR = uR[2] / rhoR, eR = uR[3];
  double vsqR = (vxR * vxR) + (vyR * vyR), PR = (eR - 0.5 * rhoR * vsqR) * (g - 1.0);
  double cRsq = g * PR / rhoR, HR = 0.5 * (vxR * vxR + vyR * vyR) + cRsq / (g - 1.0);
  double tL = sqrt(rhoL), tR = sqrt(rhoR);
  double rho = tL * tR;
  double vx = (tL * vxL + tR * vxR) / (tL + tR), vy = (tL * vyL + tR * vyR) / (tL + tR);
  double H = (tL * HL + tR * HR) / (tL + tR), vsq = vx * vx + vy * vy;
  double csq = (g - 1.0) * (H - 0.5 * vsq), P = csq * rho / g, e = P / (g - 1.0) + 0.5 * rho * vsq;
  int m;
  ua[0] = rho; ua[1] = rho * vx; ua[2] = rho * vy; ua[3] = e;
  for (m = 4; m < c->nvars; m++) ua[m] = sqrt(uL[m]) * sqrt(uR[m]);
}

/* navierstokes2d.h:166-311 */
static void eig_ns2d(const orc_case *c, const double *u, int dir, double *D, double *L, double *R)
{
  int n = c->nvars, m;
  double ga = c->gamma, gm1 = ga - 1.0, rho = u[0];
  double vx = (rho == 0) ? 0 : u[1] / rho, vy = (rho == 0) ? 0 : u[2] / rho, e = u[3];
  double vsq = (vx * vx) + (vy * vy), P = (e - 0.5 * rho * vsq) * (ga - 1.0);
  double a = sqrt(ga * P / rho), ek = 0.5 * (vx * vx + vy * vy), un, nx = 0, ny = 0;
  if (D) {
    double vn = (dir == 0) ? vx : vy;
    memset(D, 0, sizeof(double) * n * n);
    D[0] = vn - a;
    if (dir == 0) { D[1 * n + 1] = vn + a; D[2 * n + 2] = vn; }
    else { D[1 * n + 1] = vn; D[2 * n + 2] = vn + a; }
    D[3 * n + 3] = vn;
    for (m = 4; m < n; m++) D[m * n + m] = vn;
  }
  if (L) {
    memset(L, 0, sizeof(double) * n * n);
    if (dir == 0) {
      un = vx; nx = 1.0;
      L[0 * n + 0] = (gm1 * ek + a * un) / (2 * a * a);
      L[0 * n + 1] = ((-gm1) * vx - a * nx) / (2 * a * a);
      L[0 * n + 2] = ((-gm1) * vy - a * ny) / (2 * a * a);
      L[0 * n + 3] = gm1 / (2 * a * a);
      L[3 * n + 0] = (a * a - gm1 * ek) / (a * a);
      L[3 * n + 1] = (gm1 * vx) / (a * a);
      L[3 * n + 2] = (gm1 * vy) / (a * a);
      L[3 * n + 3] = (-gm1) / (a * a);
      L[1 * n + 0] = (gm1 * ek - a * un) / (2 * a * a);
      L[1 * n + 1] = ((-gm1) * vx + a * nx) / (2 * a * a);
      L[1 * n + 2] = ((-gm1) * vy + a * ny) / (2 * a * a);
      L[1 * n + 3] = gm1 / (2 * a * a);
      L[2 * n + 0] = (vy - un * ny) / nx;
      L[2 * n + 1] = ny;
      L[2 * n + 2] = (ny * ny - 1.0) / nx;
      L[2 * n + 3] = 0.0;
    } else {
      un = vy; ny = 1.0;
      L[0 * n + 0] = (gm1 * ek + a * un) / (2 * a * a);
      L[0 * n + 1] = ((1.0 - ga) * vx - a * nx) / (2 * a * a);
      L[0 * n + 2] = ((1.0 - ga) * vy - a * ny) / (2 * a * a);
      L[0 * n + 3] = gm1 / (2 * a * a);
      L[3 * n + 0] = (a * a - gm1 * ek) / (a * a);
      L[3 * n + 1] = gm1 * vx / (a * a);
      L[3 * n + 2] = gm1 * vy / (a * a);
      L[3 * n + 3] = (1.0 - ga) / (a * a);
      L[2 * n + 0] = (gm1 * ek - a * un) / (2 * a * a);
      L[2 * n + 1] = ((1.0 - ga) * vx + a * nx) / (2 * a * a);
      L[2 * n + 2] = ((1.0 - ga) * vy + a * ny) / (2 * a * a);
      L[2 * n + 3] = gm1 / (2 * a * a);
      L[1 * n + 0] = (un * nx - vx) / ny;
      L[1 * n + 1] = (1.0 - nx * nx) / ny;
      L[1 * n + 2] = -nx;
      L[1 * n + 3] = 0;
    }
    for (m = 4; m < n; m++) L[m * n + m] = 1.0;
  }
  if (R) {
    double h0 = a * a / gm1 + ek;
    memset(R, 0, sizeof(double) * n * n);
    if (dir == 0) { un = vx; nx = 1.0; ny = 0.0; } else { un = vy; ny = 1.0; nx = 0.0; }
    R[0 * n + 0] = 1.0; R[1 * n + 0] = vx - a * nx; R[2 * n + 0] = vy - a * ny; R[3 * n + 0] = h0 - a * un;
    R[0 * n + 3] = 1.0; R[1 * n + 3] = vx; R[2 * n + 3] = vy; R[3 * n + 3] = ek;
    if (dir == 0) {
      R[0 * n + 1] = 1.0; R[1 * n + 1] = vx + a * nx; R[2 * n + 1] = vy + a * ny; R[3 * n + 1] = h0 + a * un;
      R[0 * n + 2] = 0.0; R[1 * n + 2] = ny; R[2 * n + 2] = -nx; R[3 * n + 2] = vx * ny - vy * nx;
    } else {
      R[0 * n + 2] = 1.0; R[1 * n + 2] = vx + a * nx; R[2 * n + 2] = vy + a * ny; R[3 * n + 2] = h0 + a * un;
      R[0 * n + 1] = 0; R[1 * n + 1] = ny; R[2 * n + 1] = -nx; R[3 * n + 1] = vx * ny - vy * nx;
    }
    for (m = 4; m < n; m++) R[m * n + m] = 1.0;
  }
}

/* matmult_native.h:21-34 (MatMult) and :119-130 (MatVecMult) */
static void matmult(int n, double *A, const double *X, const double *Y)
{
  int i, j, k;
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) {
      A[i * n + j] = 0.0;
      for (k = 0; k < n; k++) A[i * n + j] += X[i * n + k] * Y[k * n + j];
    }
}
static void matvec(int n, double *y, const double *A, const double *x)
{
  int i, j;
  for (i = 0; i < n; i++) { y[i] = 0; for (j = 0; j < n; j++) y[i] += A[i * n + j] * x[j]; }
}

/* ------------------------------------------------------------------ */
/* WENO weights: interpolation.h:255-490                               */
/* ------------------------------------------------------------------ */

void orc_weno_weights(int kind, double c1, double c2, double c3,
                      double m3, double m2, double m1, double p1, double p2, double eps, double *w)
{
  static const double t1312 = 13.0 / 12.0, q14 = 1.0 / 4.0;
  double b1 = t1312 * (m3 - 2 * m2 + m1) * (m3 - 2 * m2 + m1) + q14 * (m3 - 4 * m2 + 3 * m1) * (m3 - 4 * m2 + 3 * m1);
  double b2 = t1312 * (m2 - 2 * m1 + p1) * (m2 - 2 * m1 + p1) + q14 * (m2 - p1) * (m2 - p1);
  double b3 = t1312 * (m1 - 2 * p1 + p2) * (m1 - 2 * p1 + p2) + q14 * (3 * m1 - 4 * p1 + p2) * (3 * m1 - 4 * p1 + p2);
  double a1, a2, a3, s;
  if (kind == 2 || kind == 3) {
    double tau;
    if (kind == 2) tau = fabs(b3 - b1);
    else tau = (m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2) * (m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2);
    a1 = c1 * (1.0 + (tau / (b1 + eps)) * (tau / (b1 + eps)));
    a2 = c2 * (1.0 + (tau / (b2 + eps)) * (tau / (b2 + eps)));
    a3 = c3 * (1.0 + (tau / (b3 + eps)) * (tau / (b3 + eps)));
    s = 1.0 / (a1 + a2 + a3);
    w[0] = a1 * s; w[1] = a2 * s; w[2] = a3 * s;
    return;
  }
  a1 = c1 / ((b1 + eps) * (b1 + eps));
  a2 = c2 / ((b2 + eps) * (b2 + eps));
  a3 = c3 / ((b3 + eps) * (b3 + eps));
  s = 1.0 / (a1 + a2 + a3);
  w[0] = a1 * s; w[1] = a2 * s; w[2] = a3 * s;
  if (kind == 1) {
    a1 = w[0] * (c1 + c1 * c1 - 3 * c1 * w[0] + w[0] * w[0]) / (c1 * c1 + w[0] * (1.0 - 2.0 * c1));
    a2 = w[1] * (c2 + c2 * c2 - 3 * c2 * w[1] + w[1] * w[1]) / (c2 * c2 + w[1] * (1.0 - 2.0 * c2));
    a3 = w[2] * (c3 + c3 * c3 - 3 * c3 * w[2] + w[2] * w[2]) / (c3 * c3 + w[2] * (1.0 - 2.0 * c3));
    s = 1.0 / (a1 + a2 + a3);
    w[0] = a1 * s; w[1] = a2 * s; w[2] = a3 * s;
  }
}

static int weight_kind(const orc_case *c)
{
  /* WENOFifthOrderCalculateWeights.c:56-59: yc > borges > mapped > JS */
  if (c->yc) return 3;
  if (c->borges) return 2;
  if (c->mapped) return 1;
  return 0;
}

/* fifth-order linear upwind value, Interp1PrimFifthOrderUpwind.c:92-96 and
 * :138-142 (the same formula on characteristic values in
 * Interp1PrimFifthOrderUpwindChar.c:174-182) */
static double upw5(double m3, double m2, double m1, double p1, double p2)
{
  static const double one_by_thirty = 1.0 / 30.0, thirteen_by_sixty = 13.0 / 60.0,
                      fortyseven_by_sixty = 47.0 / 60.0, twentyseven_by_sixty = 27.0 / 60.0,
                      one_by_twenty = 1.0 / 20.0;
  return one_by_thirty * m3 - thirteen_by_sixty * m2 + fortyseven_by_sixty * m1 + twentyseven_by_sixty * p1
         - one_by_twenty * p2;
}

/* optimal weights (WENOFifthOrderCalculateWeights.c:202-220) */
static void optimal_weights(const orc_case *c, int k, int N, double *cw)
{
  if (c->is_crweno && !(k == 0 || k == N)) { cw[0] = 0.2; cw[1] = 0.5; cw[2] = 0.3; }
  else { cw[0] = 0.1; cw[1] = 0.6; cw[2] = 0.3; }
}

/* record the weights of interface k, variable v of reconstruction `which`
 * (0 uL, 1 uR, 2 fL, 3 fR) while a capture is on: the set order of
 * WENOFifthOrderCalculateWeights.c:144-155 is LF, LU, RF, RU */
static void cap_weights(const orc_case *c, int which, int k, int v, const double *w)
{
  static const int set_of[4] = {1, 3, 0, 2};
  size_t o;
  if (!c->w_cap[0] || !c->wcap_on) return;
  o = (size_t)set_of[which] * c->w_size + (size_t)(c->wcap_p0 + k * c->wcap_pk) * c->nvars + v;
  c->w_cap[0][o] = w[0];
  c->w_cap[1][o] = w[1];
  c->w_cap[2][o] = w[2];
}

/* ------------------------------------------------------------------ */
/* tridiagonal solver: tridiagLU.c:86-277 (serial stages 1 and 4)      */
/* ------------------------------------------------------------------ */

int orc_tridiag(double *a, double *b, double *c, double *x, int n, int ns)
{
  int i, d;
  if (ns == 0 || n == 0) return 0;
  for (i = 1; i < n; i++)
    for (d = 0; d < ns; d++) {
      double f;
      if (b[(i - 1) * ns + d] == 0) return -1;
      f = a[i * ns + d] / b[(i - 1) * ns + d];
      b[i * ns + d] -= f * c[(i - 1) * ns + d];
      a[i * ns + d] = -f * a[(i - 1) * ns + d];
      x[i * ns + d] -= f * x[(i - 1) * ns + d];
    }
  for (d = 0; d < ns; d++) {
    if (b[(n - 1) * ns + d] == 0) return -1;
    x[(n - 1) * ns + d] = (x[(n - 1) * ns + d] - a[(n - 1) * ns + d] * x[d] - c[(n - 1) * ns + d] * 0.0)
                          / b[(n - 1) * ns + d];
  }
  for (i = n - 2; i >= 0; i--)
    for (d = 0; d < ns; d++) {
      if (b[i * ns + d] == 0) return -1;
      x[i * ns + d] = (x[i * ns + d] - c[i * ns + d] * x[(i + 1) * ns + d] - a[i * ns + d] * x[d]) / b[i * ns + d];
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* reconstruction along one line                                        */
/* ------------------------------------------------------------------ */

/* Gather pointers along a grid line: line point k (-g..N+g-1) of the line
 * through `base` (ghosted index of point 0 on the line) with stride st. */
#define LP(arr, k) ((arr) + (size_t)(base + (k) * st) * nv)

/* Component-wise WENO5 / CRWENO5 for all four interface arrays of one line.
 * WENOFifthOrderCalculateWeightsJS/M/Z/YC (WENOFifthOrderCalculateWeights.c:112-237 ...),
 * Interp1PrimFifthOrderWENO.c:77-171, Interp1PrimFifthOrderCRWENO.c:81-240.
 * Output interface arrays are indexed [k*ostride + v] for interface k. */
static void recon_line_components(const orc_case *c, const double *fC, const double *uC, int base, int st,
                                  int N, double *uLo, double *uRo, double *fLo, double *fRo, size_t ostride)
{
  int nv = c->nvars, k, v, kind = weight_kind(c);
  static const double s6 = 1.0 / 6.0, t3 = 1.0 / 3.0;
  double *A = (double *)malloc(sizeof(double) * (N + 1) * 4);
  double *B = A + (N + 1), *C = B + (N + 1), *X = C + (N + 1);
  int which;
  for (v = 0; v < nv; v++) {
    for (which = 0; which < 4; which++) { /* 0:uL 1:uR 2:fL 3:fR */
      const double *src = (which < 2) ? uC : fC;
      int upw = (which == 0 || which == 2) ? 1 : -1;
      double *out = (which == 0) ? uLo : (which == 1) ? uRo : (which == 2) ? fLo : fRo;
      for (k = 0; k <= N; k++) {
        double m3, m2, m1, p1, p2, w[3], cw[3], f1, f2, f3;
        int bnd = (k == 0 || k == N);
        if (upw > 0) {
          m3 = LP(src, k - 3)[v]; m2 = LP(src, k - 2)[v]; m1 = LP(src, k - 1)[v];
          p1 = LP(src, k)[v]; p2 = LP(src, k + 1)[v];
        } else {
          m3 = LP(src, k + 2)[v]; m2 = LP(src, k + 1)[v]; m1 = LP(src, k)[v];
          p1 = LP(src, k - 1)[v]; p2 = LP(src, k - 2)[v];
        }
        if (c->is_upw5) {
          out[k * ostride + v] = upw5(m3, m2, m1, p1, p2);
          continue;
        }
        optimal_weights(c, k, N, cw);
        if (c->no_limiting) { w[0] = cw[0]; w[1] = cw[1]; w[2] = cw[2]; }
        else orc_weno_weights(kind, cw[0], cw[1], cw[2], m3, m2, m1, p1, p2, c->eps, w);
        cap_weights(c, which, k, v, w);
        if (!c->is_crweno) {
          f1 = (2 * s6) * m3 + (-7 * s6) * m2 + (11 * s6) * m1;
          f2 = (-s6) * m2 + (5 * s6) * m1 + (2 * s6) * p1;
          f3 = (2 * s6) * m1 + (5 * s6) * p1 + (-s6) * p2;
          out[k * ostride + v] = w[0] * f1 + w[1] * f2 + w[2] * f3;
        } else {
          if (bnd) {
            f1 = (2 * s6) * m3 + (-7 * s6) * m2 + (11 * s6) * m1;
            f2 = (-s6) * m2 + (5 * s6) * m1 + (2 * s6) * p1;
            f3 = (2 * s6) * m1 + (5 * s6) * p1 + (-s6) * p2;
            A[k] = 0.0; B[k] = 1.0; C[k] = 0.0;
          } else {
            f1 = (s6) * m2 + (5 * s6) * m1;
            f2 = (5 * s6) * m1 + (s6) * p1;
            f3 = (s6) * m1 + (5 * s6) * p1;
            if (upw > 0) {
              A[k] = (2 * t3) * w[0] + (t3) * w[1];
              B[k] = (t3) * w[0] + (2 * t3) * w[1] + (2 * t3) * w[2];
              C[k] = (t3) * w[2];
            } else {
              C[k] = (2 * t3) * w[0] + (t3) * w[1];
              B[k] = (t3) * w[0] + (2 * t3) * w[1] + (2 * t3) * w[2];
              A[k] = (t3) * w[2];
            }
          }
          X[k] = w[0] * f1 + w[1] * f2 + w[2] * f3;
        }
      }
      if (c->is_crweno) {
        orc_tridiag(A, B, C, X, N + 1, 1);
        for (k = 0; k <= N; k++) out[k * ostride + v] = X[k];
      }
    }
  }
  free(A);
}

/* Characteristic WENO5 for one line:
 * WENOFifthOrderCalculateWeightsChar* (WENOFifthOrderCalculateWeights.c:761-930),
 * Interp1PrimFifthOrderWENOChar.c:89-201. */
static void recon_line_char(const orc_case *c, const double *fC, const double *uC, int base, int st,
                            int N, int dir, double *uLo, double *uRo, double *fLo, double *fRo, size_t ostride)
{
  int nv = c->nvars, k, v, j, kind = weight_kind(c), which;
  static const double s6 = 1.0 / 6.0;
  double L[NMAX * NMAX], R[NMAX * NMAX], ua[NMAX], fch[NMAX];
  for (k = 0; k <= N; k++) {
    /* L from the Roe average of the two cells adjacent to interface k (the
       physics' AveragingFunction: Euler1DRoeAverage / NavierStokes2DRoeAverage) */
    if (c->model_id == 1) {
      roeavg_euler1d(c, ua, LP(uC, k - 1), LP(uC, k));
      eig_euler1d(c, ua, NULL, L, R);
    } else {
      roeavg_ns2d(c, ua, LP(uC, k - 1), LP(uC, k));
      eig_ns2d(c, ua, dir, NULL, L, R);
    }
    for (which = 0; which < 4; which++) {
      const double *src = (which < 2) ? uC : fC;
      int upw = (which == 0 || which == 2) ? 1 : -1;
      double *out = (which == 0) ? uLo : (which == 1) ? uRo : (which == 2) ? fLo : fRo;
      int q[5]; /* m3 m2 m1 p1 p2 */
      if (upw > 0) { q[0] = k - 3; q[1] = k - 2; q[2] = k - 1; q[3] = k; q[4] = k + 1; }
      else { q[0] = k + 2; q[1] = k + 1; q[2] = k; q[3] = k - 1; q[4] = k - 2; }
      for (v = 0; v < nv; v++) {
        double s[5], w[3], cw[3], f1, f2, f3;
        int t;
        for (t = 0; t < 5; t++) {
          double acc = 0;
          for (j = 0; j < nv; j++) acc += L[v * nv + j] * LP(src, q[t])[j];
          s[t] = acc;
        }
        if (c->is_upw5) {
          fch[v] = upw5(s[0], s[1], s[2], s[3], s[4]);
          continue;
        }
        optimal_weights(c, k, N, cw);
        if (c->no_limiting) { w[0] = cw[0]; w[1] = cw[1]; w[2] = cw[2]; }
        else orc_weno_weights(kind, cw[0], cw[1], cw[2], s[0], s[1], s[2], s[3], s[4], c->eps, w);
        cap_weights(c, which, k, v, w);
        f1 = (2 * s6) * s[0] - (7.0 * s6) * s[1] + (11.0 * s6) * s[2];
        f2 = (-s6) * s[1] + (5.0 * s6) * s[2] + (2 * s6) * s[3];
        f3 = (2 * s6) * s[2] + (5 * s6) * s[3] - (s6) * s[4];
        fch[v] = w[0] * f1 + w[1] * f2 + w[2] * f3;
      }
      matvec(nv, out + k * ostride, R, fch);
    }
  }
}

/* ------------------------------------------------------------------ */
/* characteristic CRWENO5: Interp1PrimFifthOrderCRWENOChar.c:96-280     */
/* with the serial block solver blocktridiagLU.c:110-322 and the dense  */
/* matrix macros of matops.h:16-130, all n x n (n = nvars)              */
/* ------------------------------------------------------------------ */

/* _MatrixInvert_ (matops.h:96-130): Gauss-Jordan without pivoting */
static void bt_inv(int n, const double *A, double *B)
{
  double Ac[NMAX * NMAX], f, sum;
  int i, j, k;
  for (i = 0; i < n * n; i++) Ac[i] = A[i];
  for (i = 0; i < n * n; i++) B[i] = 0.0;
  for (i = 0; i < n; i++) B[i * n + i] = 1.0;
  for (i = 0; i < n - 1; i++)
    for (j = i + 1; j < n; j++) {
      f = Ac[j * n + i] / Ac[i * n + i];
      for (k = i + 1; k < n; k++) Ac[j * n + k] -= (f * Ac[i * n + k]);
      for (k = 0; k < j; k++) B[j * n + k] -= (f * B[i * n + k]);
    }
  for (i = n - 1; i >= 0; i--)
    for (k = 0; k < n; k++) {
      sum = 0.0;
      for (j = i + 1; j < n; j++) sum += (Ac[i * n + j] * B[j * n + k]);
      B[i * n + k] = (B[i * n + k] - sum) / Ac[i * n + i];
    }
}
/* _MatrixMultiply_: C = A B */
static void bt_mul(int n, const double *A, const double *B, double *C)
{
  int i, j, k;
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) {
      C[i * n + j] = 0;
      for (k = 0; k < n; k++) C[i * n + j] += (A[i * n + k] * B[k * n + j]);
    }
}
/* _MatrixMultiplySubtract_: C -= A B */
static void bt_mulsub(int n, double *C, const double *A, const double *B)
{
  int i, j, k;
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++)
      for (k = 0; k < n; k++) C[i * n + j] -= (A[i * n + k] * B[k * n + j]);
}
/* _MatVecMultiplySubtract_: y -= A x */
static void bt_mvsub(int n, double *y, const double *A, const double *x)
{
  int i, j;
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) y[i] -= (A[i * n + j] * x[j]);
}
/* _MatVecMultiply_: y = A x */
static void bt_mv(int n, const double *A, const double *x, double *y)
{
  int i, j;
  for (i = 0; i < n; i++) {
    y[i] = 0;
    for (j = 0; j < n; j++) y[i] += (A[i * n + j] * x[j]);
  }
}

/* blocktridiagLU on one process (blocktridiagLU.c:159-180 stage 1, :279-301
 * stage 4), one system of m block rows; a, b, c, x are overwritten */
static void bt_solve(int m, int n, double *a, double *b, double *c, double *x)
{
  int i, j, n2 = n * n;
  double binv[NMAX * NMAX], factor[NMAX * NMAX], xt[NMAX], xp1[NMAX];
  for (j = 0; j < n; j++) xp1[j] = 0.0;
  for (i = 1; i < m; i++) {
    bt_inv(n, b + (i - 1) * n2, binv);
    bt_mul(n, a + i * n2, binv, factor);
    bt_mulsub(n, b + i * n2, factor, c + (i - 1) * n2);
    for (j = 0; j < n2; j++) a[i * n2 + j] = 0.0;
    bt_mulsub(n, a + i * n2, factor, a + (i - 1) * n2);
    bt_mvsub(n, x + i * n, factor, x + (i - 1) * n);
  }
  bt_inv(n, b + (m - 1) * n2, binv);
  bt_mvsub(n, x + (m - 1) * n, a + (m - 1) * n2, x);
  bt_mvsub(n, x + (m - 1) * n, c + (m - 1) * n2, xp1);
  bt_mv(n, binv, x + (m - 1) * n, xt);
  for (j = 0; j < n; j++) x[(m - 1) * n + j] = xt[j];
  for (i = m - 2; i > -1; i--) {
    bt_inv(n, b + i * n2, binv);
    bt_mvsub(n, x + i * n, c + i * n2, x + (i + 1) * n);
    bt_mvsub(n, x + i * n, a + i * n2, x);
    bt_mv(n, binv, x + i * n, xt);
    for (j = 0; j < n; j++) x[i * n + j] = xt[j];
  }
}

/* One line: the four block systems (uL, uR, fL, fR), each row the CRWENO5
 * relation of every characteristic field at interface k, written in the
 * physical variables through L of the Roe average at k; WENO5 rows
 * (B = L) at the two ends of the line. */
static void recon_line_crweno_char(const orc_case *c, const double *fC, const double *uC, int base, int st,
                                   int N, int dir, double *uLo, double *uRo, double *fLo, double *fRo,
                                   size_t ostride)
{
  int nv = c->nvars, n2 = nv * nv, k, v, j, which, kind = weight_kind(c);
  static const double t3 = 1.0 / 3.0, s6 = 1.0 / 6.0;
  double *A = (double *)malloc(sizeof(double) * (size_t)(N + 1) * (3 * n2 + nv));
  double *B = A + (size_t)(N + 1) * n2, *C = B + (size_t)(N + 1) * n2, *F = C + (size_t)(N + 1) * n2;
  double L[NMAX * NMAX], R[NMAX * NMAX], ua[NMAX];
  for (which = 0; which < 4; which++) { /* 0:uL 1:uR 2:fL 3:fR */
    const double *src = (which < 2) ? uC : fC;
    int upw = (which == 0 || which == 2) ? 1 : -1;
    double *out = (which == 0) ? uLo : (which == 1) ? uRo : (which == 2) ? fLo : fRo;
    for (k = 0; k <= N; k++) {
      int q[5], bnd = (k == 0 || k == N);
      if (upw > 0) { q[0] = k - 3; q[1] = k - 2; q[2] = k - 1; q[3] = k; q[4] = k + 1; }
      else { q[0] = k + 2; q[1] = k + 1; q[2] = k; q[3] = k - 1; q[4] = k - 2; }
      /* Roe average of the cells qm1 and qp1 (swapped for the right bias) */
      if (c->model_id == 1) {
        roeavg_euler1d(c, ua, LP(uC, q[2]), LP(uC, q[3]));
        eig_euler1d(c, ua, NULL, L, R);
      } else {
        roeavg_ns2d(c, ua, LP(uC, q[2]), LP(uC, q[3]));
        eig_ns2d(c, ua, dir, NULL, L, R);
      }
      for (v = 0; v < nv; v++) {
        double s[5], w[3], cw[3], f1, f2, f3, ca, cb, cc;
        int t;
        for (t = 0; t < 5; t++) {
          double acc = 0;
          for (j = 0; j < nv; j++) acc += L[v * nv + j] * LP(src, q[t])[j];
          s[t] = acc;
        }
        optimal_weights(c, k, N, cw);
        if (c->no_limiting) { w[0] = cw[0]; w[1] = cw[1]; w[2] = cw[2]; }
        else orc_weno_weights(kind, cw[0], cw[1], cw[2], s[0], s[1], s[2], s[3], s[4], c->eps, w);
        cap_weights(c, which, k, v, w);
        if (bnd) {
          f1 = (2 * s6) * s[0] - (7.0 * s6) * s[1] + (11.0 * s6) * s[2];
          f2 = (-s6) * s[1] + (5.0 * s6) * s[2] + (2 * s6) * s[3];
          f3 = (2 * s6) * s[2] + (5 * s6) * s[3] - (s6) * s[4];
          for (j = 0; j < nv; j++) {
            A[k * n2 + v * nv + j] = 0.0;
            C[k * n2 + v * nv + j] = 0.0;
            B[k * n2 + v * nv + j] = L[v * nv + j];
          }
        } else {
          f1 = (s6) * (s[1] + 5 * s[2]);
          f2 = (s6) * (5 * s[2] + s[3]);
          f3 = (s6) * (s[2] + 5 * s[3]);
          ca = ((2 * t3) * w[0] + (t3) * w[1]);
          cb = ((t3) * w[0] + (2 * t3) * (w[1] + w[2]));
          cc = ((t3) * w[2]);
          for (j = 0; j < nv; j++) {
            if (upw > 0) {
              A[k * n2 + v * nv + j] = ca * L[v * nv + j];
              B[k * n2 + v * nv + j] = cb * L[v * nv + j];
              C[k * n2 + v * nv + j] = cc * L[v * nv + j];
            } else {
              C[k * n2 + v * nv + j] = ca * L[v * nv + j];
              B[k * n2 + v * nv + j] = cb * L[v * nv + j];
              A[k * n2 + v * nv + j] = cc * L[v * nv + j];
            }
          }
        }
        F[k * nv + v] = w[0] * f1 + w[1] * f2 + w[2] * f3;
      }
    }
    bt_solve(N + 1, nv, A, B, C, F);
    for (k = 0; k <= N; k++)
      for (v = 0; v < nv; v++) out[k * ostride + v] = F[k * nv + v];
  }
  free(A);
}

/* Upwinding at the interfaces of one line.
 * NavierStokes2DUpwindRoe (NavierStokes2DUpwind.c:26-92),
 * Euler1DUpwindRoe (Euler1DUpwind.c:27-93), Euler1DUpwindRF (:107-190). */
static void upwind_line(const orc_case *c, const double *uC, int base, int st, int N, int dir,
                        const double *uLi, const double *uRi, const double *fLi, const double *fRi,
                        double *fI, size_t ostride)
{
  int nv = c->nvars, k, m;
  double D[NMAX * NMAX], L[NMAX * NMAX], R[NMAX * NMAX], DL[NMAX * NMAX], mA[NMAX * NMAX];
  double ua[NMAX], ud[NMAX], us[NMAX];
  int rf = !strcmp(c->upw, "rf-char"), llf = !strcmp(c->upw, "llf-char"), rus = !strcmp(c->upw, "rusanov");
  for (k = 0; k <= N; k++) {
    const double *uLk = uLi + k * ostride, *uRk = uRi + k * ostride;
    const double *fLk = fLi + k * ostride, *fRk = fRi + k * ostride;
    double *fo = fI + k * ostride;
    if (rus) {
      /* Euler1DUpwindRusanov (Euler1DUpwind.c:282-353; its wave speed is
         gamma*P/rho + |u|, without a square root) and NavierStokes2DUpwindRusanov
         (NavierStokes2DUpwind.c:271-332) */
      double a[3], alpha;
      const double *pts[3];
      int q;
      if (c->model_id == 1) roeavg_euler1d(c, ua, LP(uC, k - 1), LP(uC, k));
      else roeavg_ns2d(c, ua, LP(uC, k - 1), LP(uC, k));
      pts[0] = LP(uC, k - 1); pts[1] = LP(uC, k); pts[2] = ua;
      for (q = 0; q < 3; q++) {
        if (c->model_id == 1) {
          double rho = pts[q][0], uvel = pts[q][1] / rho, E = pts[q][2];
          double P = (E - 0.5 * rho * uvel * uvel) * (c->gamma - 1.0);
          a[q] = c->gamma * P / rho + fabs(uvel);
        } else {
          double rho, vx, vy, E, P;
          ns2d_flowvar(pts[q], c->gamma, &rho, &vx, &vy, &E, &P);
          a[q] = sqrt(c->gamma * P / rho) + fabs(dir == 0 ? vx : vy);
        }
      }
      {  /* max3(alphaL, alphaR, alphaavg), math_ops.h:30 */
        double m1 = (a[0] > a[1]) ? a[0] : a[1], m2 = (a[1] > a[2]) ? a[1] : a[2];
        alpha = (m1 > m2) ? m1 : m2;
      }
      for (m = 0; m < nv; m++) fo[m] = 0.5 * (fLk[m] + fRk[m]) - alpha * (0.5 * (uRk[m] - uLk[m]));
      continue;
    }
    if (llf) {
      /* Euler1DUpwindLLF (Euler1DUpwind.c:199-266): Roe average and eigenvalues
         of the neighbouring cells; NavierStokes2DUpwindLLF
         (NavierStokes2DUpwind.c:190-258): of the interface states */
      double ucL[NMAX], ucR[NMAX], fcL[NMAX], fcR[NMAX], fc[NMAX], eL[NMAX], eR[NMAX], eC[NMAX];
      if (c->model_id == 1) {
        roeavg_euler1d(c, ua, LP(uC, k - 1), LP(uC, k));
        eig_euler1d(c, ua, D, L, R);
      } else {
        roeavg_ns2d(c, ua, uLk, uRk);
        eig_ns2d(c, ua, dir, D, L, R);
      }
      matvec(nv, ucL, L, uLk); matvec(nv, ucR, L, uRk);
      matvec(nv, fcL, L, fLk); matvec(nv, fcR, L, fRk);
      if (c->model_id == 1) {
        eig_euler1d(c, LP(uC, k - 1), D, NULL, NULL);
        for (m = 0; m < nv; m++) eL[m] = D[m * nv + m];
        eig_euler1d(c, LP(uC, k), D, NULL, NULL);
        for (m = 0; m < nv; m++) eR[m] = D[m * nv + m];
        eig_euler1d(c, ua, D, NULL, NULL);
      } else {
        eig_ns2d(c, uLk, dir, D, NULL, NULL);
        for (m = 0; m < nv; m++) eL[m] = D[m * nv + m];
        eig_ns2d(c, uRk, dir, D, NULL, NULL);
        for (m = 0; m < nv; m++) eR[m] = D[m * nv + m];
        eig_ns2d(c, ua, dir, D, NULL, NULL);
      }
      for (m = 0; m < nv; m++) eC[m] = D[m * nv + m];
      for (m = 0; m < nv; m++) {
        double aL = fabs(eL[m]), aC = fabs(eC[m]), aR = fabs(eR[m]);
        double m1 = (aL > aC) ? aL : aC, m2 = (aC > aR) ? aC : aR;
        double alpha = (m1 > m2) ? m1 : m2;
        fc[m] = 0.5 * (fcL[m] + fcR[m] + alpha * (ucL[m] - ucR[m]));
      }
      matvec(nv, fo, R, fc);
      continue;
    }
    if (c->model_id == 1 && rf) {
      double ucL[NMAX], ucR[NMAX], fcL[NMAX], fcR[NMAX], fc[NMAX], Dl[NMAX * NMAX];
      roeavg_euler1d(c, ua, LP(uC, k - 1), LP(uC, k));
      eig_euler1d(c, ua, D, L, R);
      matvec(nv, ucL, L, uLk); matvec(nv, ucR, L, uRk);
      matvec(nv, fcL, L, fLk); matvec(nv, fcR, L, fRk);
      for (m = 0; m < nv; m++) {
        double eL, eC, eR;
        eig_euler1d(c, LP(uC, k - 1), Dl, NULL, NULL); eL = Dl[m * nv + m];
        eig_euler1d(c, LP(uC, k), Dl, NULL, NULL); eR = Dl[m * nv + m];
        eig_euler1d(c, ua, Dl, NULL, NULL); eC = Dl[m * nv + m];
        if (eL > 0 && eC > 0 && eR > 0) fc[m] = fcL[m];
        else if (eL < 0 && eC < 0 && eR < 0) fc[m] = fcR[m];
        else {
          double aL = fabs(eL), aC = fabs(eC), aR = fabs(eR);
          double m1 = (aL > aC) ? aL : aC, m2 = (aC > aR) ? aC : aR;
          double alpha = (m1 > m2) ? m1 : m2;
          fc[m] = 0.5 * (fcL[m] + fcR[m] + alpha * (ucL[m] - ucR[m]));
        }
      }
      matvec(nv, fo, R, fc);
      continue;
    }
    if (c->model_id == 2 && rf) {
      /* NavierStokes2DUpwindRF (NavierStokes2DUpwind.c:106-187): Roe average of
         the two interface states; eigenvalues of uL, uR and the average */
      double ucL[NMAX], ucR[NMAX], fcL[NMAX], fcR[NMAX], fc[NMAX], eL[NMAX], eR[NMAX], eC[NMAX];
      roeavg_ns2d(c, ua, uLk, uRk);
      eig_ns2d(c, ua, dir, D, L, R);
      matvec(nv, ucL, L, uLk); matvec(nv, ucR, L, uRk);
      matvec(nv, fcL, L, fLk); matvec(nv, fcR, L, fRk);
      eig_ns2d(c, uLk, dir, D, NULL, NULL);
      for (m = 0; m < nv; m++) eL[m] = D[m * nv + m];
      eig_ns2d(c, uRk, dir, D, NULL, NULL);
      for (m = 0; m < nv; m++) eR[m] = D[m * nv + m];
      eig_ns2d(c, ua, dir, D, NULL, NULL);
      for (m = 0; m < nv; m++) eC[m] = D[m * nv + m];
      for (m = 0; m < nv; m++) {
        if (eL[m] > 0 && eC[m] > 0 && eR[m] > 0) fc[m] = fcL[m];
        else if (eL[m] < 0 && eC[m] < 0 && eR[m] < 0) fc[m] = fcR[m];
        else {
          double aL = fabs(eL[m]), aC = fabs(eC[m]), aR = fabs(eR[m]);
          double m1 = (aL > aC) ? aL : aC, m2 = (aC > aR) ? aC : aR;
          double alpha = (m1 > m2) ? m1 : m2;
          fc[m] = 0.5 * (fcL[m] + fcR[m] + alpha * (ucL[m] - ucR[m]));
        }
      }
      matvec(nv, fo, R, fc);
      continue;
    }
    for (m = 0; m < nv; m++) ud[m] = 0.5 * (uRk[m] - uLk[m]);
    if (c->model_id == 1) {
      roeavg_euler1d(c, ua, LP(uC, k - 1), LP(uC, k));
      eig_euler1d(c, ua, D, L, R);
      for (m = 0; m < nv; m++) D[m * nv + m] = fabs(D[m * nv + m]);
    } else {
      double delta = 0.000001, delta2 = delta * delta;
      roeavg_ns2d(c, ua, LP(uC, k - 1), LP(uC, k));
      eig_ns2d(c, ua, dir, D, L, R);
      for (m = 0; m < nv; m++) {
        double dd = D[m * nv + m];
        D[m * nv + m] = (fabs(dd) < delta ? (dd * dd + delta2) / (2 * delta) : fabs(dd));
      }
    }
    matmult(nv, DL, D, L);
    matmult(nv, mA, R, DL);
    matvec(nv, us, mA, ud);
    for (m = 0; m < nv; m++) fo[m] = 0.5 * (fLk[m] + fRk[m]) - us[m];
  }
}

/* ------------------------------------------------------------------ */
/* hyperbolic term: HyperbolicFunction.c:34-127                        */
/* ------------------------------------------------------------------ */

static void hyperbolic(orc_case *c, const double *u, double *hyp)
{
  int nv = c->nvars, g = c->ghosts, d, v;
  size_t n = (size_t)c->npts * nv;
  int xoff = 0;
  memset(hyp, 0, n * sizeof(double));
  for (d = 0; d < c->ndims; d++) {
    int N = c->dim[d], st = c->stride[d], p, nl, l;
    int other = (c->ndims == 2) ? 1 - d : -1;
    /* cell-centred flux at all points including ghosts */
    for (p = 0; p < c->npts; p++) {
      if (c->model_id == 1) flux_euler1d(c, u + (size_t)p * nv, c->fluxC + (size_t)p * nv);
      else flux_ns2d(c, u + (size_t)p * nv, c->fluxC + (size_t)p * nv, d);
    }
    nl = (other >= 0) ? c->dim[other] : 1;
    for (l = 0; l < nl; l++) {
      int idx[3] = {0, 0, 0}, base, k;
      double *uLl = c->uL, *uRl = c->uR, *fLl = c->fL, *fRl = c->fR, *fIl = c->fluxI;
      if (other >= 0) idx[other] = l;
      idx[d] = 0;
      base = gidx(c, idx);
      c->wcap_on = (d == c->cap_dir);
      if (d == c->cap_dir) {
        c->wcap_p0 = (c->ndims == 1) ? 0 : (d == 0 ? (long)l * (N + 1) : l);
        c->wcap_pk = (c->ndims == 1 || d == 0) ? 1 : c->dim[0];
      }
      if (c->is_char && c->is_crweno)
        recon_line_crweno_char(c, c->fluxC, u, base, st, N, d, uLl, uRl, fLl, fRl, (size_t)nv);
      else if (c->is_char) recon_line_char(c, c->fluxC, u, base, st, N, d, uLl, uRl, fLl, fRl, (size_t)nv);
      else recon_line_components(c, c->fluxC, u, base, st, N, uLl, uRl, fLl, fRl, (size_t)nv);
      if (d == c->cap_dir) {
        double *src4[4] = {uLl, uRl, fLl, fRl};
        int r;
        for (r = 0; r < 4; r++)
          if (c->if_cap[r]) memcpy(c->if_cap[r] + (size_t)l * (N + 1) * nv, src4[r], (size_t)(N + 1) * nv * sizeof(double));
      }
      upwind_line(c, u, base, st, N, d, uLl, uRl, fLl, fRl, fIl, (size_t)nv);
      if (c->fi_cap[d]) memcpy(c->fi_cap[d] + (size_t)l * (N + 1) * nv, fIl, (size_t)(N + 1) * nv * sizeof(double));
      /* hyp += dxinv * (fI[k+1] - fI[k]) at the N interior points of the line */
      for (k = 0; k < N; k++) {
        int pp = base + k * st;
        double dxi = c->dxinv[xoff + g + k];
        for (v = 0; v < nv; v++) hyp[(size_t)pp * nv + v] += dxi * (fIl[(k + 1) * nv + v] - fIl[k * nv + v]);
      }
    }
    xoff += N + 2 * g;
  }
  c->wcap_on = 0;
}

/* ------------------------------------------------------------------ */
/* parabolic term (navierstokes2d): NavierStokes2DParabolicFunction.c:41-204 */
/* ------------------------------------------------------------------ */

/* FirstDerivativeFourthOrder.c:39-132 along one line (N interior, g ghosts),
 * f/df accessed with stride (in points) st and nv components. */
static void fd4_line(const double *f, double *df, int base, int st, int N, int g, int nv)
{
  static const double t12 = 1.0 / 12.0;
  int i, v;
#define F(k) f[(size_t)(base + (k) * st) * nv + v]
#define DF(k) df[(size_t)(base + (k) * st) * nv + v]
  for (v = 0; v < nv; v++) {
    i = -g;
    DF(i) = (-25 * F(i) + 48 * F(i + 1) - 36 * F(i + 2) + 16 * F(i + 3) - 3 * F(i + 4)) * t12;
    i = -g + 1;
    DF(i) = (-3 * F(i - 1) - 10 * F(i) + 18 * F(i + 1) - 6 * F(i + 2) + F(i + 3)) * t12;
    for (i = -g + 2; i < N + g - 2; i++) DF(i) = (F(i - 2) - 8 * F(i - 1) + 8 * F(i + 1) - F(i + 2)) * t12;
    i = N + g - 2;
    DF(i) = (-F(i - 3) + 6 * F(i - 2) - 18 * F(i - 1) + 10 * F(i) + 3 * F(i + 1)) * t12;
    i = N + g - 1;
    DF(i) = (3 * F(i - 4) - 16 * F(i - 3) + 36 * F(i - 2) - 48 * F(i - 1) + 25 * F(i)) * t12;
  }
}

/* FirstDerivativeFirstOrder.c:20-95 */
static void fd1_line(const double *f, double *df, int base, int st, int N, int g, int nv, int bias)
{
  int i, v;
  double mx = bias > 0 ? bias : 0, mn = bias < 0 ? bias : 0, bb = bias;
  for (v = 0; v < nv; v++) {
    i = -g; DF(i) = F(i + 1) - F(i);
    for (i = -g + 1; i < N + g - 1; i++) DF(i) = mx * F(i + 1) - bb * F(i) + mn * F(i - 1);
    i = N + g - 1; DF(i) = F(i) - F(i - 1);
  }
#undef F
#undef DF
}

void orc_fd4_line(const double *f, double *df, int N, int g) { fd4_line(f, df, g, 1, N, g, 1); }

/* derivative along dir over the interior lines of the other dimension */
static void first_derivative(const orc_case *c, double *Df, const double *f, int dir, int bias)
{
  int g = c->ghosts, nv = c->nvars, other = 1 - dir, l;
  for (l = 0; l < c->dim[other]; l++) {
    int idx[3] = {0, 0, 0}, base;
    idx[other] = l; idx[dir] = 0;
    base = gidx(c, idx);
    if (c->par4) fd4_line(f, Df, base, c->stride[dir], c->dim[dir], g, nv);
    else fd1_line(f, Df, base, c->stride[dir], c->dim[dir], g, nv, bias);
  }
}

/* MPIExchangeBoundariesnD (MPIExchangeBoundariesnD.c:45-176) on one rank per
 * ... of a run decomposed along every periodic dimension, seen from the whole
 * grid: along each dimension d with a periodic zone (InitializeBoundaries.c:
 * 179-181 sets bcperiodic[d] when more than one rank lies along d), the g
 * ghost layers of the interior lines across d receive the opposite g interior
 * layers, all nvars components (:109-167); the corners are not exchanged.
 * Every value read is an interior one, so the dimensions' order is immaterial. */
static int periodic_dim(const orc_case *c, int d)
{
  int k;
  for (k = 0; k < c->nbc; k++)
    if (c->bc[k].dim == d && !strcmp(c->bc[k].type, "periodic")) return 1;
  return 0;
}

static void exchange_periodic_ghosts(const orc_case *c, double *var)
{
  int g = c->ghosts, nv = c->nvars, d, l, k, v;
  for (d = 0; d < c->ndims; d++) {
    int other = (c->ndims == 2) ? 1 - d : -1, N = c->dim[d];
    int nl = (other >= 0) ? c->dim[other] : 1;
    if (!periodic_dim(c, d)) continue;
    for (l = 0; l < nl; l++)
      for (k = 0; k < g; k++) {
        int glo[3] = {0, 0, 0}, ihi[3] = {0, 0, 0}, ghi[3] = {0, 0, 0}, ilo[3] = {0, 0, 0};
        if (other >= 0) glo[other] = ihi[other] = ghi[other] = ilo[other] = l;
        glo[d] = -g + k;  ihi[d] = N - g + k;   /* low ghosts <- the last g interior layers */
        ghi[d] = N + k;   ilo[d] = k;           /* high ghosts <- the first g interior layers */
        for (v = 0; v < nv; v++) {
          var[(size_t)gidx(c, glo) * nv + v] = var[(size_t)gidx(c, ihi) * nv + v];
          var[(size_t)gidx(c, ghi) * nv + v] = var[(size_t)gidx(c, ilo) * nv + v];
        }
      }
  }
}

void orc_set_derivative_exchange(orc_case *c, int on) { c->deriv_exchange = on ? 1 : 0; }

static void parabolic_ns2d(orc_case *c, const double *u, double *par)
{
  int g = c->ghosts, nv = c->nvars, i, j, v, imax = c->dim[0], jmax = c->dim[1];
  size_t n = (size_t)c->npts * nv;
  double two_third = 2.0 / 3.0, inv_gm1 = 1.0 / (c->gamma - 1.0), inv_Re = 1.0 / c->Re, inv_Pr = 1.0 / c->Pr;
  double *Q = c->Q, *QdX = c->QdX, *QdY = c->QdY, *FV = c->FV, *FD = c->FD;
  memset(par, 0, n * sizeof(double));
  if (c->Re <= 0) return;
  memset(Q, 0, n * sizeof(double)); memset(QdX, 0, n * sizeof(double)); memset(QdY, 0, n * sizeof(double));
  memset(FV, 0, n * sizeof(double)); memset(FD, 0, n * sizeof(double));
  for (i = -g; i < imax + g; i++)
    for (j = -g; j < jmax + g; j++) {
      int idx[2] = {i, j};
      size_t p = (size_t)gidx(c, idx) * nv;
      double rho = u[p], vx = (rho == 0) ? 0 : u[p + 1] / rho, vy = (rho == 0) ? 0 : u[p + 2] / rho, e = u[p + 3];
      double vsq = (vx * vx) + (vy * vy), P = (e - 0.5 * rho * vsq) * (c->gamma - 1.0);
      Q[p] = rho; Q[p + 1] = vx; Q[p + 2] = vy;
      Q[p + 3] = c->gamma * P / Q[p];
    }
  first_derivative(c, QdX, Q, 0, 1);
  first_derivative(c, QdY, Q, 1, 1);
  /* NavierStokes2DParabolicFunction.c:91-94: a no-op in a serial run */
  if (c->deriv_exchange) {
    exchange_periodic_ghosts(c, QdX);
    exchange_periodic_ghosts(c, QdY);
  }
  for (i = -g; i < imax + g; i++)
    for (j = -g; j < jmax + g; j++) {
      int idx[2] = {i, j};
      size_t p = (size_t)gidx(c, idx) * nv;
      double dxi = c->dxinv[g + i], dyi = c->dxinv[(imax + 2 * g) + g + j];
      for (v = 0; v < nv; v++) { QdX[p + v] *= dxi; QdY[p + v] *= dyi; }
    }
  for (int dir = 0; dir < 2; dir++) {
    for (i = -g; i < imax + g; i++)
      for (j = -g; j < jmax + g; j++) {
        int idx[2] = {i, j};
        size_t p = (size_t)gidx(c, idx) * nv;
        double uvel = Q[p + 1], vvel = Q[p + 2], T = Q[p + 3];
        double ux = QdX[p + 1], vx = QdX[p + 2], uy = QdY[p + 1], vy = QdY[p + 2];
        double Td = T * c->Tref;
        double mu = exp(1.5 * log(Td / c->T0)) * (c->T0 + c->TS) / (Td + c->TS);
        double kap = exp(1.5 * log(Td / c->T0)) * (c->T0 + c->TA * exp(-c->TB / c->T0))
                   / (Td + c->TA * exp(-c->TB / Td));
        FV[p] = 0.0;
        if (dir == 0) {
          double Tx = QdX[p + 3];
          double txx = two_third * (mu * inv_Re) * (2 * ux - vy);
          double txy = (mu * inv_Re) * (uy + vx);
          double qx = (kap * inv_Re * inv_gm1 * inv_Pr) * Tx;
          FV[p + 1] = txx; FV[p + 2] = txy; FV[p + 3] = uvel * txx + vvel * txy + qx;
        } else {
          double Ty = QdY[p + 3];
          double tyx = (mu * inv_Re) * (uy + vx);
          double tyy = two_third * (mu * inv_Re) * (-ux + 2 * vy);
          double qy = (kap * inv_Re * inv_gm1 * inv_Pr) * Ty;
          FV[p + 1] = tyx; FV[p + 2] = tyy; FV[p + 3] = uvel * tyx + vvel * tyy + qy;
        }
      }
    first_derivative(c, FD, FV, dir, -1);
    for (i = 0; i < imax; i++)
      for (j = 0; j < jmax; j++) {
        int idx[2] = {i, j};
        size_t p = (size_t)gidx(c, idx) * nv;
        double di = (dir == 0) ? c->dxinv[g + i] : c->dxinv[(imax + 2 * g) + g + j];
        for (v = 0; v < nv; v++) par[p + v] += (di * FD[p + v]);
      }
  }
}

/* ------------------------------------------------------------------ */
/* source term: SourceFunction.c:21-60 -> ChemistrySource (ChemistrySolve.c:18-115) */
/* ------------------------------------------------------------------ */

static void source(orc_case *c, const double *u, double t, double *src)
{
  int nv = c->nvars, nz = c->nz_layers, q, iz, d;
  size_t n = (size_t)c->npts * nv;
  memset(src, 0, n * sizeof(double));
  if (!c->include_chem) { sponge_source(c, u, src); return; }
  for (q = 0; q < c->npts_int; q++) {
    int idx[3] = {0, 0, 0}, r = q, p;
    double sigma = c->t_pulse_norm / 2.35482, I0 = 0.0;
    for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
    p = gidx(c, idx);
    if (t > c->t_start_norm) {
      double tp = t - (c->t_start_norm + c->t_pulse_norm);
      I0 = c->I0 * exp(-(tp * tp) / (2 * sigma * sigma)) * c->imap[p];
    }
    c->nv_hnu[nz * p + 0] = I0 / (c->c_light * c->h_planck * c->nu * c->n_O2);
    for (iz = 1; iz < nz; iz++) {
      double sg = c->sO3 * c->n_O2;
      double damp = 1.0 - c->dz * sg * u[(size_t)c->grid_stride * p + c->nflow + c->z_stride * (iz - 1) + 1];
      c->nv_hnu[nz * p + iz] = c->nv_hnu[nz * p + (iz - 1)] * damp;
    }
  }
  for (iz = 0; iz < nz; iz++)
    for (q = 0; q < c->npts_int; q++) {
      int idx[3] = {0, 0, 0}, r = q, p, nfv = c->nflow + c->z_stride * iz;
      double *f, nh;
      const double *uu;
      double nO2, nO3, n1D, n1Dg, n3Su, n1Sg, nCO2;
      for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
      p = gidx(c, idx);
      f = src + (size_t)c->grid_stride * p;
      uu = u + (size_t)c->grid_stride * p;
      nh = c->nv_hnu[nz * p + iz];
      nO2 = uu[nfv + 0]; nO3 = uu[nfv + 1]; n1D = uu[nfv + 2]; n1Dg = uu[nfv + 3];
      n3Su = uu[nfv + 4]; n1Sg = uu[nfv + 5]; nCO2 = uu[nfv + 6];
      /* _ChemistrySetRHS_ chemistry.h:30-78 */
      f[nfv + 0] = 0.0;
      f[nfv + 1] = -(c->k0a + c->k0b) * nh * nO3 - (c->k2a + c->k2b) * nO3 * n1D
                 - (c->k3a + c->k3b) * nO3 * n1Sg - c->k5 * nO3 * n3Su - c->k6 * n1Dg * nO3;
      f[nfv + 2] = c->k0a * nh * nO3 - (c->k1a + c->k1b) * n1D * nO2 - (c->k2a + c->k2b) * n1D * nO3
                 - c->k4 * n1D * nCO2;
      f[nfv + 3] = c->k0a * nh * nO3 + c->k5 * nO3 * n3Su - c->k6 * n1Dg * nO3;
      f[nfv + 4] = c->k2a * nO3 * n1D - c->k5 * nO3 * n3Su;
      f[nfv + 5] = c->k1a * n1D * nO2 - (c->k3a + c->k3b) * nO3 * n1Sg;
      f[nfv + 6] = 0.0;
      /* _ChemistrySetQ_ chemistry.h:84-127 into the energy slot */
      f[c->nflow - 1] = ((c->q0a * c->k0a + c->q0b * c->k0b) * nh * nO3
                         + (c->q1a * c->k1a + c->q1b * c->k1b) * nO2 * n1D
                         + (c->q2a * c->k2a + c->q2b * c->k2b) * nO3 * n1D
                         + (c->q3a * c->k3a + c->q3b * c->k3b) * nO3 * n1Sg
                         + c->q4 * c->k4 * n1D * nCO2
                         + c->q5 * c->k5 * nO3 * n3Su
                         + c->q6 * c->k6 * n1Dg * nO3) / (c->ch_gamma - 1.0);
    }
  sponge_source(c, u, src);
}

/* ------------------------------------------------------------------ */
/* RHS and time stepping                                                */
/* ------------------------------------------------------------------ */

/* TimeRHSFunctionExplicit.c:29-77 (u is modified: BCs applied in place) */
static void rhs_eval(orc_case *c, double *rhs, double *u, double t)
{
  size_t n = (size_t)c->npts * c->nvars, i;
  apply_bcs(c, u);
  memset(rhs, 0, n * sizeof(double));
  hyperbolic(c, u, c->hyp_a);
  for (i = 0; i < n; i++) rhs[i] += -1.0 * c->hyp_a[i];
  if (c->model_id == 2) {
    parabolic_ns2d(c, u, c->par_a);
    for (i = 0; i < n; i++) rhs[i] += 1.0 * c->par_a[i];
  }
  source(c, u, t, c->src_a);
  for (i = 0; i < n; i++) rhs[i] += 1.0 * c->src_a[i];
}

/* CFL: Euler1DComputeCFL.c:19-56, NavierStokes2DComputeCFL.c:19-56 */
static double compute_cfl(const orc_case *c)
{
  double mx = 0;
  int q, d, nv = c->nvars, g = c->ghosts;
  for (q = 0; q < c->npts_int; q++) {
    int idx[3] = {0, 0, 0}, r = q;
    const double *uu;
    for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
    uu = c->u + (size_t)gidx(c, idx) * nv;
    if (c->model_id == 1) {
      double rho = uu[0], v = uu[1] / rho, P = (uu[2] - 0.5 * rho * v * v) * (c->gamma - 1.0);
      double cs = sqrt(c->gamma * P / rho), l = (fabs(v) + cs) * c->dt * c->dxinv[g + idx[0]];
      if (l > mx) mx = l;
    } else {
      double rho = uu[0], vx = (rho == 0) ? 0 : uu[1] / rho, vy = (rho == 0) ? 0 : uu[2] / rho;
      double vsq = (vx * vx) + (vy * vy), P = (uu[3] - 0.5 * rho * vsq) * (c->gamma - 1.0);
      double cs = sqrt(c->gamma * P / rho);
      double lx = (fabs(vx) + cs) * c->dt * c->dxinv[g + idx[0]];
      double ly = (fabs(vy) + cs) * c->dt * c->dxinv[(c->dim[0] + 2 * g) + g + idx[1]];
      if (lx > mx) mx = lx;
      if (ly > mx) mx = ly;
    }
  }
  return mx;
}

int orc_step(orc_case *c, int nsteps, double *cfl, double *norm)
{
  size_t n = (size_t)c->npts * c->nvars, i;
  int it, s, j;
  double *u0 = (norm) ? (double *)malloc(n * sizeof(double)) : NULL;
  for (it = 0; it < nsteps; it++) {
    /* TimePreStep.c:21-81 */
    apply_bcs(c, c->u);
    if (u0) memcpy(u0, c->u, n * sizeof(double));
    if (cfl) cfl[it] = compute_cfl(c);
    /* TimeRK.c:34-102 */
    for (s = 0; s < c->nstages; s++) {
      double st = c->waqt + c->c[s] * c->dt;
      memcpy(c->U[s], c->u, n * sizeof(double));
      for (j = 0; j < s; j++) {
        double a = c->dt * c->A[s * c->nstages + j];
        for (i = 0; i < n; i++) c->U[s][i] += a * c->Udot[j][i];
      }
      rhs_eval(c, c->Udot[s], c->U[s], st);
    }
    for (s = 0; s < c->nstages; s++) {
      double bb = c->dt * c->b[s];
      for (i = 0; i < n; i++) c->u[i] += bb * c->Udot[s][i];
    }
    /* TimePostStep.c:27-64 */
    c->waqt += c->dt;
    if (norm) {
      double sum = 0;
      int q, d;
      for (i = 0; i < n; i++) u0[i] += -1.0 * c->u[i];
      for (q = 0; q < c->npts_int; q++) {
        int idx[3] = {0, 0, 0}, r = q, v;
        for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
        for (v = 0; v < c->nvars; v++) {
          double z = u0[(size_t)gidx(c, idx) * c->nvars + v];
          sum += z * z;
        }
      }
      norm[it] = sqrt(sum / (double)c->npts_int);
    }
  }
  free(u0);
  return 0;
}

static void interior_copy(const orc_case *c, double *dst, const double *src, int to_interior)
{
  int q, d, v, nv = c->nvars;
  for (q = 0; q < c->npts_int; q++) {
    int idx[3] = {0, 0, 0}, r = q;
    size_t p;
    for (d = 0; d < c->ndims; d++) { idx[d] = r % c->dim[d]; r /= c->dim[d]; }
    p = (size_t)gidx(c, idx);
    for (v = 0; v < nv; v++) {
      if (to_interior) dst[(size_t)q * nv + v] = src[p * nv + v];
      else ((double *)dst)[p * nv + v] = src[(size_t)q * nv + v];
    }
  }
}

void orc_get_solution(const orc_case *c, double *u_out) { interior_copy(c, u_out, c->u, 1); }
void orc_set_solution(orc_case *c, const double *u_in) { interior_copy(c, c->u, u_in, 0); }

/* The pieces of TimeRHSFunctionExplicit.c:29-77 at time t (see the header) */
int orc_rhs_parts(orc_case *c, const double *u_in, double t, double *u_ghosted, double *flux0, double *flux1,
                  double *fi0, double *fi1, double *hyp, double *par, double *src)
{
  size_t n = (size_t)c->npts * c->nvars, p;
  int d;
  double *w = (double *)calloc(n, sizeof(double));
  double *fl[2] = {flux0, flux1};
  interior_copy(c, w, u_in, 0);
  apply_bcs(c, w);
  if (u_ghosted) memcpy(u_ghosted, w, n * sizeof(double));
  for (d = 0; d < c->ndims; d++) {
    if (!fl[d]) continue;
    for (p = 0; p < (size_t)c->npts; p++) {
      if (c->model_id == 1) flux_euler1d(c, w + p * c->nvars, fl[d] + p * c->nvars);
      else flux_ns2d(c, w + p * c->nvars, fl[d] + p * c->nvars, d);
    }
  }
  c->fi_cap[0] = fi0;
  c->fi_cap[1] = c->ndims > 1 ? fi1 : NULL;
  hyperbolic(c, w, c->hyp_a);
  c->fi_cap[0] = c->fi_cap[1] = NULL;
  if (hyp) interior_copy(c, hyp, c->hyp_a, 1);
  memset(c->par_a, 0, n * sizeof(double));
  if (c->model_id == 2) parabolic_ns2d(c, w, c->par_a);
  if (par) interior_copy(c, par, c->par_a, 1);
  source(c, w, t, c->src_a);
  if (src) interior_copy(c, src, c->src_a, 1);
  free(w);
  return 0;
}

int orc_rhs_interfaces(orc_case *c, const double *u_in, double t, int dir, double *uL, double *uR, double *fL,
                       double *fR, double *w1, double *w2, double *w3)
{
  size_t n = (size_t)c->npts * c->nvars;
  int N, nl, r;
  double *w;
  if (dir < 0 || dir >= c->ndims) return 1;
  N = c->dim[dir];
  nl = (c->ndims == 2) ? c->dim[1 - dir] : 1;
  w = (double *)calloc(n, sizeof(double));
  interior_copy(c, w, u_in, 0);
  apply_bcs(c, w);
  c->if_cap[0] = uL; c->if_cap[1] = uR; c->if_cap[2] = fL; c->if_cap[3] = fR;
  c->w_cap[0] = w1; c->w_cap[1] = w2; c->w_cap[2] = w3;
  c->w_size = (long)(N + 1) * nl * c->nvars;
  c->cap_dir = dir;
  if (c->is_upw5) c->w_cap[0] = c->w_cap[1] = c->w_cap[2] = NULL;  /* upw5 has no weights */
  hyperbolic(c, w, c->hyp_a);
  c->cap_dir = -1;
  for (r = 0; r < 4; r++) c->if_cap[r] = NULL;
  c->w_cap[0] = c->w_cap[1] = c->w_cap[2] = NULL;
  free(w);
  return 0;
}

void orc_first_derivative(const orc_case *c, const double *f, int dir, int bias, double *Df)
{
  int g = c->ghosts, nv = c->nvars, l, nl = (c->ndims == 2) ? c->dim[1 - dir] : 1;
  for (l = 0; l < nl; l++) {
    int idx[3] = {0, 0, 0}, base;
    if (c->ndims == 2) idx[1 - dir] = l;
    base = gidx(c, idx);
    if (c->par4) fd4_line(f, Df, base, c->stride[dir], c->dim[dir], g, nv);
    else fd1_line(f, Df, base, c->stride[dir], c->dim[dir], g, nv, bias);
  }
}

int orc_rhs(orc_case *c, const double *u_in, double t, double *rhs_out)
{
  size_t n = (size_t)c->npts * c->nvars;
  double *w = (double *)calloc(n, sizeof(double)), *r = (double *)calloc(n, sizeof(double));
  interior_copy(c, w, u_in, 0);
  rhs_eval(c, r, w, t);
  interior_copy(c, rhs_out, r, 1);
  free(w); free(r);
  return 0;
}
