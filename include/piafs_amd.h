/*
 * piafs_amd.h -- C ABI of the MI355X-native PIAFS right-hand-side and
 * explicit Runge-Kutta time-step path.
 *
 * Conventions: extern "C", opaque handles, plain pointers and sizes, int
 * status (0 = ok, > 0 = input/configuration error, < 0 = device or numerical
 * error).  No exceptions cross the ABI.  A handle is used by one host thread
 * at a time; different handles are independent.  Host arrays use the
 * reference's point-major layout without ghost points (x fastest, the nvars
 * components of a point contiguous) -- the WriteBinary order of
 * src/IOFunctions/WriteBinary.c:37-93.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the PIAFS source tree).
 */
#ifndef PIAFS_AMD_H
#define PIAFS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct piafs_case piafs_case;

typedef struct {
  int    ndims, nvars, ghosts;
  int    dim[3];
  int    n_iter, restart_iter, nstages;
  double dt, t;          /* fixed step size and current simulation time */
  double gamma, Re, Pr;  /* after the photochemistry overrides */
  int    model;          /* 1 = euler1d, 2 = navierstokes2d */
  int    scheme;         /* 0 = weno5, 1 = crweno5, 2 = upw5 */
  int    characteristic; /* 1 = characteristic interpolation */
  int    include_chem;
  size_t device_bytes;   /* HBM held by this case */
} piafs_info;

/* ---------------------------------------------------------------------------
 * Case lifecycle.
 * Replaces the initialisation sequence of src/main.cpp:133-212
 * (Simulation::define -> ReadInputs.c:185, Initialize.c:25,
 * InitialSolution.c:24, InitializeBoundaries.c:35, InitializeSolvers.c:38,
 * InitializePhysics.c:27 -> NavierStokes2DInitialize.c:68 /
 * Euler1DInitialize.c:58 -> ChemistryInitialize.c:69).  Reads solver.inp,
 * boundary.inp, physics.inp, weno.inp, chemistry.inp, initial.inp and
 * imap.inp from `dir` (NULL = current directory) with the reference's
 * keywords, defaults and selection strings, and uploads the state to
 * device `device`. */
int  piafs_create(const char *dir, int device, piafs_case **out);
/* Replaces the cleanup of src/Simulation/Cleanup.c. */
int  piafs_destroy(piafs_case *c);
int  piafs_get_info(const piafs_case *c, piafs_info *info);
/* Last error message for this case (or the last failed create). */
const char *piafs_last_error(const piafs_case *c);

/* State transfer in WriteBinary order (interior points only).
 * Replaces ReadArray (src/IOFunctions/ReadArray.c:28) into HyPar::u and the
 * gather of WriteArraySerial (src/IOFunctions/WriteArray.c:77). */
int  piafs_upload_state(piafs_case *c, const double *u_host);
int  piafs_download_state(piafs_case *c, double *u_host);
int  piafs_set_time(piafs_case *c, double t);

/* Advance nsteps time steps.  Replaces the body of the loop in
 * src/Simulation/Solve.cpp:54-94: TimePreStep (TimePreStep.c:21) ->
 * TimeStep -> TimeIntegration::TimeIntegrate = TimeRK (TimeRK.c:34) ->
 * TimePostStep (TimePostStep.c:27).  Asynchronous on the case's stream. */
int  piafs_step(piafs_case *c, int nsteps);
/* Same, recording per-step max CFL (TimePreStep.c:54-69) and step norm
 * (TimePostStep.c:36-64) into host arrays of length nsteps (either may be
 * NULL).  Synchronous. */
int  piafs_step_diag(piafs_case *c, int nsteps, double *cfl, double *norm);
/* Wait for all queued work of this case. */
int  piafs_sync(piafs_case *c);

/* One right-hand-side evaluation F(u) at time t of an interior state given in
 * WriteBinary order.  Replaces TimeIntegration::RHSFunction =
 * TimeRHSFunctionExplicit (src/TimeIntegration/TimeRHSFunctionExplicit.c:29:
 * ApplyBoundaryConditions, HyperbolicFunction, ParabolicFunction,
 * SourceFunction). */
int  piafs_rhs(piafs_case *c, const double *u_host, double t, double *rhs_host);

/* Per-operator entry points: the pieces of one right-hand-side evaluation
 * of the interior state u_host at time t (the case's own solution and time are
 * left as they were).  "Ghosted" arrays hold every point of the grid with its
 * ghost layers, x fastest, the nvars components of a point contiguous: the
 * layout of HyPar::u (include/hypar.h:106-111).
 *
 * The state after the boundary fill.  Replaces ApplyBoundaryConditions
 * (src/HyParFunctions/ApplyBoundaryConditions.c:31-54) as
 * TimeRHSFunctionExplicit.c:45 calls it; corner points stay zero. */
int  piafs_boundary_fill(piafs_case *c, const double *u_host, double t, double *u_ghosted);
/* The physical flux along dimension dir at every position of every interior
 * line along dir (ghosts along dir included; the other points are zero), after
 * the boundary fill.  Replaces the FluxFunction argument, HyPar::FFunction =
 * NavierStokes2DFlux (src/PhysicalModels/NavierStokes2D/NavierStokes2DFlux.c:24-52)
 * / Euler1DFlux (src/PhysicalModels/Euler1D/Euler1DFlux.c:19), as
 * src/HyParFunctions/HyperbolicFunction.c:84 calls it.  Ghosted layout. */
int  piafs_flux(piafs_case *c, const double *u_host, double t, int dir, double *f_ghosted);
/* The upwind interface flux fI of every interior line along dir, as
 * [line][interface 0..N][var].  Replaces ReconstructHyperbolic
 * (src/HyParFunctions/HyperbolicFunction.c:168-214: InterpolateInterfacesHyp
 * for uL, uR, fL, fR, then the UpwindFunction argument, e.g.
 * NavierStokes2DUpwindRoe NavierStokes2DUpwind.c:26-92). */
int  piafs_interface_flux(piafs_case *c, const double *u_host, double t, int dir, double *fi);
/* The three terms of the right-hand side at the interior points (WriteBinary
 * order; any output may be NULL): hyp, the flux divergence that
 * TimeRHSFunctionExplicit subtracts (HyperbolicFunction,
 * src/HyParFunctions/HyperbolicFunction.c:34-127); par, the viscous term
 * (ParabolicFunction = NavierStokes2DParabolicFunction.c:41-204; zero for
 * euler1d); src, the photochemistry and sponge source (SourceFunction,
 * src/HyParFunctions/SourceFunction.c:21-60).  rhs = -hyp + par + src. */
int  piafs_rhs_terms(piafs_case *c, const double *u_host, double t, double *hyp, double *par, double *src);

/* ---------------------------------------------------------------------------
 * Device-buffer per-operator entry points: the same pieces on arrays the
 * caller keeps on the device in the reference's own layouts, enqueued on
 * `stream` (the caller's HIP stream handle; NULL = the case's own stream) without host
 * synchronisation.  Ghosted arrays as above; the interface array of
 * dimension dir holds the points of the grid with dim[dir]+1 interfaces along
 * dir and no ghosts, x fastest, nvars contiguous (the fI, uL, uR, fL, fR of
 * src/HyParFunctions/HyperbolicFunction.c:61-70).  These calls use the case's
 * own device buffers as scratch: calls on one case must not overlap each
 * other or its steps (the reference's function table is not reentrant
 * either).  The case's solution and time are left as they are.  A zero pivot
 * in a line solve shows as error -3 of the next piafs_sync.
 *
 * ApplyBoundaryConditions (src/HyParFunctions/ApplyBoundaryConditions.c:
 * 31-54): fills the ghost points of u_ghosted in place from its interior
 * (corners are not written). */
int  piafs_dev_boundary_fill(piafs_case *c, double *u_ghosted, double t, void *stream);
/* FFunction (NavierStokes2DFlux.c:24-52, Euler1DFlux.c:19): the flux along
 * dir at every point of the ghosted grid. */
int  piafs_dev_flux(piafs_case *c, const double *u_ghosted, int dir, double *f_ghosted, void *stream);
/* InterpolateInterfacesHyp (HyperbolicFunction.c:204-207 calls it four
 * times per dimension; Interp1PrimFifthOrderWENO.c:77-171,
 * Interp1PrimFifthOrderCRWENO.c:81-240 and their characteristic forms): the
 * interface values fI of fC along dir, left-biased for upw = 1, right-biased
 * for upw = -1, on every interior line, with the deck's scheme and weights.
 * As ReconstructHyperbolic computes the weights (SetInterpLimiterVar on the
 * flux and the solution), they are those of fC itself: uflag = 1 with fC =
 * u_ghosted interpolates the solution, uflag = 0 with fC the flux of
 * u_ghosted along dir the flux; characteristic interpolation takes its
 * eigenvectors from u_ghosted. */
int  piafs_dev_interp_interfaces(piafs_case *c, double *fI, const double *fC, const double *u_ghosted, int upw,
                                 int dir, int uflag, void *stream);
/* Upwind (HyPar::Upwind, e.g. NavierStokes2DUpwindRoe
 * NavierStokes2DUpwind.c:26-92, Euler1DUpwindRoe Euler1DUpwind.c:27-93, and the
 * rf-char / llf-char / rusanov forms): the upwind interface flux fI along dir
 * from the four interface arrays and the cell-centred ghosted state. */
int  piafs_dev_upwind(piafs_case *c, double *fI, const double *fL, const double *fR, const double *uL,
                      const double *uR, const double *u_ghosted, int dir, double t, void *stream);
/* FirstDerivativePar (FirstDerivativeFourthOrder.c:39-132 for
 * par_space_scheme 4; FirstDerivativeFirstOrder.c:20-95 for "2", as
 * InitializeSolvers.c:69-72 selects it): Df = d f / d x_dir (unscaled) at every
 * position, ghosts included, of every interior line across dir; other points
 * of Df are not written. */
int  piafs_dev_first_derivative_par(piafs_case *c, double *Df, const double *f_ghosted, int dir, int bias,
                                    void *stream);
/* SetInterpLimiterVar = WENOFifthOrderCalculateWeights (src/
 * InterpolationFunctions/WENOFifthOrderCalculateWeights.c:41-84 with the JS,
 * mapped, Z and YC forms and their characteristic variants): the nonlinear
 * weights of every interface of dir on every interior line, four sets
 * (left-biased flux, left-biased solution, right-biased flux, right-biased
 * solution) at w + k*size (k = 0..3), each set an interface array of dir.
 * Pass the reference's weno->w1 + weno->offset[dir] (and w2, w3) with
 * size = weno->size to fill its arrays.  Not for upw5 (no weights). */
int  piafs_dev_weno_weights(piafs_case *c, const double *fC, const double *u_ghosted, int dir, double *w1,
                            double *w2, double *w3, long size, void *stream);

/* Write the current solution in the reference's binary layout.  Replaces
 * WriteBinary (src/IOFunctions/WriteBinary.c:37-93) as called from
 * OutputSolution (src/Simulation/OutputSolution.cpp:25). */
int  piafs_write_binary(piafs_case *c, const char *path);

/* Whole solve as the reference's Solve (src/Simulation/Solve.cpp:29-123)
 * followed by SimWriteErrors (src/Simulation/SimulationWriteErrors.c:21-100):
 * n_iter steps from restart_iter, screen output every screen_op_iter, the
 * solution files op[_NNNNN].{bin,dat} in the deck's op_file_format
 * (OutputSolution.cpp:22-70), op_species_z* with chemistry
 * (ChemistryWriteSpecies.c:18-80), then errors.dat (against exact.inp when
 * present, CalculateError.c:28-121), conservation.dat and function_counts.dat
 * in dir.  Returns the wall time of the time loop in *seconds (may be NULL). */
int  piafs_solve(piafs_case *c, const char *dir, int verbose, double *seconds);

/* Derivative ghosts of the viscous term with the semantics of a run
 * decomposed over more than one rank along each periodic dimension (on = 1)
 * or of a serial run (on = 0, the default).  Replaces the call
 * MPIExchangeBoundariesnD(QDerivX / QDerivY) of
 * src/PhysicalModels/NavierStokes2D/NavierStokes2DParabolicFunction.c:91-94
 * (src/MPIFunctions/MPIExchangeBoundariesnD.c:45-176), a no-op in a serial
 * run: along a dimension with a periodic zone, the ghost points of the
 * interior lines across it receive the derivatives of their periodic images;
 * the viscous term is then invariant under a periodic translation of the
 * state.  Affects navierstokes2d with Re > 0 only. */
int  piafs_set_derivative_exchange(piafs_case *c, int on);

/* Conservation error of every variable after the steps taken so far (nvars
 * values; -1 before the first step): |volume integral + time integral of the
 * boundary flux - initial volume integral| / max(1, |initial volume
 * integral|).  Replaces CalculateConservationError
 * (src/HyParFunctions/CalculateConservationError.c:19-43) with the boundary
 * integral of TimeRK.c:86-97 and BoundaryIntegral.c:37-58.  Needs
 * conservation_check yes in solver.inp (returns 1 otherwise). */
int  piafs_conservation_errors(piafs_case *c, double *errors);

/* Device timing of the dominant kernel: average duration (ms) of each
 * kernel class over the last piafs_step call when profiling is enabled.
 * names/ms arrays of length *n (in: capacity, out: count). */
int  piafs_kernel_times(piafs_case *c, int enable, const char **names, double *ms, int *n);

/* Tile-shape autotuner (no counterpart in the reference): times the
 * reconstruction launch for line-solve segments of 4, 8, 12 and 16 rows per
 * lane on this case and keeps the fastest.  Results do not depend on the
 * choice (the line solves are exact).  Optional outputs: the chosen rows per
 * lane and its reconstruction time in ms. */
int  piafs_autotune(piafs_case *c, int *rows_per_lane, double *recon_ms);

/* ---------------------------------------------------------------------------
 * Per-operator entry points for unit parity (host arrays in, host arrays out;
 * the device work runs on an internal stream and is synchronised).
 */

/* Batched tridiagonal solve: a team of lanes per system of up to 4097 rows,
 * one thread per longer system.  Replaces tridiagLU
 * (src/TridiagLU/tridiagLU.c:86-277, serial stages 1 and 4): a, b, c, x each
 * [n][ns] (row-major over rows, systems contiguous, as tridiagLU.h:60-62);
 * x is overwritten with the solution.  Returns -1 on a zero pivot. */
int  piafs_tridiag_host(const double *a, const double *b, const double *c, double *x, int n, int ns);

/* The same solve without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * device copies of a, b, c, x and the zero-pivot flag live in a fixed device
 * array of 4,194,304 doubles (32 MiB) that each device holds until the
 * process ends: n*ns must not exceed 1,048,575 (returns 2 otherwise, having
 * done nothing).  Every call on a device shares that array, so calls that
 * overlap on one device race.  Returns as piafs_tridiag_host. */
int  piafs_tridiag_host_noalloc(const double *a, const double *b, const double *c, double *x, int n, int ns);

/* Nonlinear WENO5 weights for `count` stencils (five values each, ordered
 * m3 m2 m1 p1 p2).  kind: 0 JS, 1 mapped, 2 borges (Z), 3 YC.  Replaces the
 * weight macros of include/interpolation.h:255-490 as used by
 * WENOFifthOrderCalculateWeights (src/InterpolationFunctions/
 * WENOFifthOrderCalculateWeights.c:41-84).  w_out: count x 3. */
int  piafs_weno_weights_host(int kind, const double *copt, const double *stencils, int count,
                             double eps, double *w_out);

/* ---------------------------------------------------------------------------
 * Batched ensembles: many independent cases advanced together, one launch per
 * kernel class per RK stage covering every case, the step captured in a
 * HIP graph.  Replaces the reference's serial ensemble loop
 * (src/Simulation/EnsembleSimulationsDefine.cpp:31-119, TimeRK.c:46-81).
 * Every launch block belongs to one member case and reads that case's own
 * descriptor, so members may differ in model, grid size, scheme, weights, RK
 * tableau, dt, Re and beam map.
 */
typedef struct piafs_batch piafs_batch;
/* Cases are addressed by slot numbers (0..ncases-1 after create).  On a
 * failed create *out still receives a handle whose last_error names the
 * failing case; destroy it. */
int  piafs_batch_create(const char *const *dirs, int ncases, int device, piafs_batch **out);
int  piafs_batch_destroy(piafs_batch *b);
/* Advance every member case by nsteps (asynchronous; graph replay by default). */
int  piafs_batch_step(piafs_batch *b, int nsteps);
int  piafs_batch_sync(piafs_batch *b);
/* Replay steps from a captured HIP graph (default) or launch them eagerly. */
int  piafs_batch_use_graph(piafs_batch *b, int on);
/* Continuous batching: admit a case between steps (its state is uploaded and
 * stays resident) and release a finished one.  The launches of the next step
 * cover the new member set. */
int  piafs_batch_add(piafs_batch *b, const char *dir, int *slot);
int  piafs_batch_remove(piafs_batch *b, int slot);
int  piafs_batch_steps_done(piafs_batch *b, int slot, int *steps);
/* piafs_autotune over all members of the batch together. */
int  piafs_batch_autotune(piafs_batch *b, int *rows_per_lane, double *recon_ms);
int  piafs_batch_download(piafs_batch *b, int slot, double *u_host);
int  piafs_batch_info(piafs_batch *b, int slot, piafs_info *info);
const char *piafs_batch_last_error(const piafs_batch *b);

/* Ensemble directory: replaces EnsembleSimulation::define and the ensemble
   branch of main (EnsembleSimulationsDefine.cpp:31-119, main.cpp:92-111).
   `dir` holds simulation.inp ("begin nsims N end"), one solver.inp whose
   `size` (and `size_exact`) line lists one tuple per member, and per-member
   files initial_<i>.inp, boundary_<i>.inp (else the shared boundary.inp) and
   exact_<i>.inp, <i> zero-padded to floor(log10 N) + 1 digits
   (InitialSolution.c:32-41, InitializeBoundaries.c:51-75).  Creates one batch
   whose slot i is member i.  Returns 0; 1 when simulation.inp is absent or
   malformed or a member fails to load -- the batch handle is still returned
   so that piafs_batch_last_error can say why; destroy it. */
int  piafs_ensemble_create(const char *dir, int device, piafs_batch **out);
/* The reference's run of the ensemble (Solve.cpp:28-110, TimePrintStep.c:
   19-75, SimulationWriteErrors.c:21-100): n_iter steps of every member,
   writing op_<i>[_NNNNN], op_species_z<k>_<i>_[_NNNNN], errors_<i>.dat,
   conservation_<i>.dat and function_counts_<i>.dat into dir.  The batch must
   hold the ensemble as created (no member added or removed).  Returns 0, or
   nonzero with piafs_batch_last_error set. */
int  piafs_ensemble_solve(piafs_batch *b, const char *dir, int verbose, double *seconds);

/* Device memory a case built from `dir` would hold (admission control for
 * serving); for an ensemble directory (simulation.inp) the sum over its
 * members.  Parses the decks on the host only.  Returns 0 on success, 1 with
 * piafs_last_error(NULL) saying why. */
int  piafs_estimate_bytes(const char *dir, size_t *bytes);

/* ---------------------------------------------------------------------------
 * Serving: a request loop with continuous batching.  No counterpart in the
 * reference, which runs one simulation per process (src/main.cpp:133-235) or
 * a serial loop over an ensemble (EnsembleSimulationsDefine.cpp:31-119).
 * A request is the solve of one deck directory to its n_iter steps.  Requests
 * are parsed on the submitting thread, queued, and admitted in arrival order
 * while the device memory they hold fits the budget (hbm_budget = 0: 80% of
 * the free device memory at create); admitted requests stay resident and are
 * advanced together with the other members of their class (1D or 2D) by the
 * ensemble launches, each class on its own stream and host thread; members
 * join and leave between rounds, and every result equals that of the case
 * run alone.  Thread-safe: any thread may submit and poll.
 */
typedef struct piafs_server piafs_server;
int  piafs_server_create(int device, size_t hbm_budget, piafs_server **out);
int  piafs_server_destroy(piafs_server *s);
/* Returns 1 for an invalid deck, 2 for a request larger than the budget. */
int  piafs_server_submit(piafs_server *s, const char *dir, int *id);
/* state: 0 queued, 1 running, 2 done, < 0 failed; latency (seconds from
 * submit to completion) once done or failed, else -1. */
int  piafs_server_poll(piafs_server *s, int id, int *state, double *latency_seconds);
/* Final state of a finished request (WriteBinary order) and its info. */
int  piafs_server_result(piafs_server *s, int id, double *u_host, piafs_info *info);
/* Forget a finished request (frees its host copy of the result). */
int  piafs_server_release(piafs_server *s, int id);
/* Block until no request is queued or running. */
int  piafs_server_wait_all(piafs_server *s);
int  piafs_server_stats(piafs_server *s, size_t *bytes_in_use, int *running, int *queued);
const char *piafs_server_last_error(const piafs_server *s);

#ifdef __cplusplus
}
#endif
#endif
