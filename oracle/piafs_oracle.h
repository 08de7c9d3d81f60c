/*
 * piafs_oracle.h -- CPU restatement of the PIAFS right-hand-side and
 * explicit Runge-Kutta time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (piafs_amd) never links or calls it.
 *
 * Layout follows the reference exactly: ghosted arrays, x fastest, the nvars
 * components of a grid point contiguous (include/arrayfunctions.h:43-50 in the
 * reference).  Evaluation order follows the reference sources line by line so
 * that, compiled with gcc -O2 (no FMA contraction on x86-64), results agree
 * with the reference bit for bit where the reference itself is deterministic.
 *
 * Pinning (see tests/test_oracle_*.py): the full-step path is compared bit for
 * bit against the reference binary built from source (oracle/_ref) and against
 * the fixtures shipped in the reference tree; individual functions are checked
 * against closed forms (Thomas vs dense solve, linear WENO5/compact-5 on
 * polynomials, Roe consistency, free-stream preservation, RK tableaux).
 */
#ifndef PIAFS_ORACLE_H
#define PIAFS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case orc_case;

/* Parse the input decks in `dir` (solver.inp, boundary.inp, physics.inp,
 * weno.inp, chemistry.inp, initial.inp, imap.inp), allocate and initialise.
 * Returns 0 on success.  Mirrors main.cpp:133-212 init order. */
int  orc_load(const char *dir, orc_case **out);
void orc_free(orc_case *c);

/* Basic queries. */
int    orc_ndims(const orc_case *c);
int    orc_nvars(const orc_case *c);
int    orc_dim(const orc_case *c, int d);
int    orc_ghosts(const orc_case *c);
double orc_time(const orc_case *c);
double orc_dt(const orc_case *c);
int    orc_n_iter(const orc_case *c);
/* derived physics constants (gamma, Re, Pr after the chemistry overrides) */
double orc_param(const orc_case *c, const char *name);

/* Advance `nsteps` time steps (TimePreStep + TimeRK + TimePostStep).
 * If cfl/norm are non-NULL they receive per-step max CFL and step norm
 * (TimePreStep.c:54-69, TimePostStep.c:36-64). */
int  orc_step(orc_case *c, int nsteps, double *cfl, double *norm);

/* Interior solution in WriteBinary order (point-major, x fastest, vars
 * innermost): prod(dim)*nvars doubles. */
void orc_get_solution(const orc_case *c, double *u_out);
void orc_set_solution(orc_case *c, const double *u_in);

/* One right-hand-side evaluation of the interior state u_in at time t
 * (TimeRHSFunctionExplicit.c:29-77).  Output interior only, same order. */
int  orc_rhs(orc_case *c, const double *u_in, double t, double *rhs_out);

/* The pieces of the same evaluation, for unit parity of the per-operator
 * entry points (any output may be NULL): the state after
 * ApplyBoundaryConditions (ghosted, npoints-with-ghosts x nvars, AoS);
 * the flux along dimensions 0 and 1 at every ghosted point (FFunction:
 * NavierStokes2DFlux.c:24-52, Euler1DFlux.c:19); the upwind interface flux
 * of every line along dimensions 0 and 1 ([line][interface][var], the fI of
 * HyperbolicFunction.c:68-81); and the hyperbolic, parabolic and source terms
 * at the interior points (WriteBinary order). */
int  orc_rhs_parts(orc_case *c, const double *u_in, double t, double *u_ghosted, double *flux0, double *flux1,
                   double *fi0, double *fi1, double *hyp, double *par, double *src);

/* The interface values of dimension dir in the same evaluation, as
 * InterpolateInterfacesHyp writes them for ReconstructHyperbolic
 * (HyperbolicFunction.c:204-207): uL, uR, fL, fR of every interior line,
 * [line][interface][var]; and the nonlinear weights they were made with
 * (SetInterpLimiterVar, WENOFifthOrderCalculateWeights.c:112-237 and the
 * M / Z / YC / characteristic forms), four sets LF, LU, RF, RU at w + k*size
 * with size = (dim[dir]+1)*lines*nvars, each in the reference's interface
 * order (interface points x fastest, nvars contiguous).  Any output may be
 * NULL; upw5 writes no weights. */
int  orc_rhs_interfaces(orc_case *c, const double *u_in, double t, int dir, double *uL, double *uR, double *fL,
                        double *fR, double *w1, double *w2, double *w3);

/* FirstDerivativePar on a ghosted array: FirstDerivativeFourthOrderCentral
 * (FirstDerivativeFourthOrder.c:39-132) for par_space_scheme 4, else
 * FirstDerivativeFirstOrder (FirstDerivativeFirstOrder.c:20-95) with `bias`,
 * at every position (ghosts included) of the interior lines across dir;
 * other points of Df are left as they are. */
void orc_first_derivative(const orc_case *c, const double *f, int dir, int bias, double *Df);

/* Derivative ghosts of the viscous term as a run decomposed over more than
 * one rank along each periodic dimension computes them: the MPI exchange of
 * QDerivX / QDerivY (NavierStokes2DParabolicFunction.c:91-94,
 * MPIExchangeBoundariesnD.c:45-176), a no-op in a serial run, copies the
 * periodic images into the ghost layers of the interior lines.  Off by
 * default (serial semantics). */
void orc_set_derivative_exchange(orc_case *c, int on);

/* ---- function-level entry points used by unit tests ---- */

/* tridiagLU serial path (tridiagLU.c:86-277): n rows, ns systems,
 * a/b/c/x each [n][ns].  Returns -1 on a zero pivot. */
int  orc_tridiag(double *a, double *b, double *c, double *x, int n, int ns);

/* WENO5 weights for one stencil: kind 0=JS,1=mapped,2=Z(borges),3=YC.
 * (interpolation.h:255-490) */
void orc_weno_weights(int kind, double c1, double c2, double c3,
                      double m3, double m2, double m1, double p1, double p2,
                      double eps, double *w);

/* 4th-order first derivative along one line of N interior points with g
 * ghosts (FirstDerivativeFourthOrder.c:39-132); f and df have N+2g entries. */
void orc_fd4_line(const double *f, double *df, int N, int g);

#ifdef __cplusplus
}
#endif
#endif
