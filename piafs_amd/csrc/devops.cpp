// Device-buffer per-operator entry points of the C ABI: the pieces of the
// right-hand side on arrays the caller keeps on the device in the reference's
// own layouts, enqueued on the caller's stream.
//
// Layouts (include/hypar.h:106-111, HyperbolicFunction.c:61-70): a ghosted
// array holds every grid point with its ghost layers, x fastest, the nvars
// components of a point contiguous; the interface array of dimension dir holds
// the points of the grid with dim[dir]+1 interfaces along dir and no ghosts,
// in the same order.
//
// The reconstruction and upwinding run the step's own kernels on staged
// copies: the case's stage-state buffers (natural and transposed), its flux
// buffers and its interface buffers in split mode, all of which a step
// rewrites from the solution before reading them.  The case's solution is
// left alone.
#include <cstring>

#include "case.hpp"

using namespace piafs;

namespace {

hipStream_t stream_of(piafs_case *pc, void *stream) {
  return stream ? (hipStream_t)stream : pc->eng.stream;
}

AosGrid aos_grid(const Case *c) {
  const Grid &G = c->h.G;
  return AosGrid{G.ndims, G.nv, G.g, G.nx, G.ny};
}

int fail(piafs_case *pc, int rc, const std::string &msg) {
  pc->err = msg;
  return rc;
}

int launched(piafs_case *pc) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(pc, -1, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// The split-mode descriptor of the case and the one-member engine that runs
// it, built at the first call and rebuilt when the case's descriptor changes
// (autotuned segment lengths, the derivative exchange, a rebuilt plan).
int unit_prepare(piafs_case *pc, hipStream_t st) {
  Case *c = pc->c;
  if (!pc->unit) {
    auto un = std::make_unique<UnitOps>();
    un->eng.owns_cases = false;
    if (un->eng.init(c->device)) return fail(pc, -1, un->eng.err);
    un->eng.add(c);
    const Grid &G = c->h.G;
    un->u = c->dalloc((size_t)G.nv * G.plane, st);
    if (!un->u) return fail(pc, -2, "device allocation failed");
    if (c->h.paired) {  // a paired case keeps two interface arrays per sweep; split mode needs four
      const SweepGeom &Sx = c->h.Sx, &Sy = c->h.Sy;
      for (int k = 0; k < 2; k++) {
        un->io[0][k] = c->dalloc((size_t)G.nv * Sx.nlines * (Sx.N + 1), st);
        if (G.ndims == 2) un->io[1][k] = c->dalloc((size_t)G.nv * Sy.nlines * (Sy.N + 1), st);
        if (!un->io[0][k] || (G.ndims == 2 && !un->io[1][k])) return fail(pc, -2, "device allocation failed");
      }
    }
    if (hipStreamSynchronize(st) != hipSuccess) return fail(pc, -1, "stream synchronisation failed");
    pc->unit = std::move(un);
  }
  UnitOps &un = *pc->unit;
  CaseDev d = c->h;
  d.paired = 0;
  d.terms = 7;
  d.u = un.u;
  d.status = pc->eng.d_status;  // zero pivots reach the case's own status word (piafs_sync)
  if (c->h.paired) {
    d.ix[2] = un.io[0][0];
    d.ix[3] = un.io[0][1];
    d.iy[2] = un.io[1][0];
    d.iy[3] = un.io[1][1];
  }
  if (!un.eng.built || memcmp(&d, &un.desc, sizeof d) != 0) {
    if (hipStreamSynchronize(st) != hipSuccess) return fail(pc, -1, "stream synchronisation failed");
    un.desc = d;
    un.eng.override_desc = &un.desc;
    if (un.eng.build()) return fail(pc, -1, un.eng.err);
  }
  return 0;
}

// the ghosted reference-layout state into the case's stage-state buffers
void stage_state(piafs_case *pc, const double *u_ghosted, hipStream_t st) {
  Case *c = pc->c;
  launch_dev_planar(c->h.G, u_ghosted, c->h.U, c->h.G.ndims == 2 ? c->h.UT : nullptr, 1, 0, nullptr, st);
}

}  // namespace

extern "C" {

int piafs_dev_boundary_fill(piafs_case *pc, double *u_ghosted, double t, void *stream) {
  (void)t;  // the boundary types of this path do not depend on time
  if (!pc || !pc->c) return 1;
  if (!u_ghosted) return fail(pc, 1, "state pointer is NULL");
  const hipStream_t st = stream_of(pc, stream);
  int rc = unit_prepare(pc, st);
  if (rc) return rc;
  Case *c = pc->c;
  UnitOps &un = *pc->unit;
  // the interior points are the state the ghost rules read (stage 0: U = u)
  launch_dev_planar(c->h.G, u_ghosted, un.u, nullptr, 1, 0, nullptr, st);
  launch_class(un.eng.d_cases, un.eng.plan, K_GHOST, 0, st);
  launch_dev_planar(c->h.G, nullptr, c->h.U, nullptr, 0, 1, u_ghosted, st);
  return launched(pc);
}

int piafs_dev_flux(piafs_case *pc, const double *u_ghosted, int dir, double *f_ghosted, void *stream) {
  if (!pc || !pc->c) return 1;
  if (!u_ghosted || !f_ghosted) return fail(pc, 1, "state or output pointer is NULL");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) return fail(pc, 1, "dir must be 0 .. ndims-1");
  launch_dev_flux(aos_grid(c), c->h.model, c->h.P.gamma, dir, u_ghosted, f_ghosted, stream_of(pc, stream));
  return launched(pc);
}

int piafs_dev_interp_interfaces(piafs_case *pc, double *fI, const double *fC, const double *u_ghosted, int upw,
                                int dir, int uflag, void *stream) {
  if (!pc || !pc->c) return 1;
  if (!fI || !fC || !u_ghosted) return fail(pc, 1, "an array pointer is NULL");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) return fail(pc, 1, "dir must be 0 .. ndims-1");
  if (upw != 1 && upw != -1) return fail(pc, 1, "upw must be 1 or -1");
  const hipStream_t st = stream_of(pc, stream);
  int rc = unit_prepare(pc, st);
  if (rc) return rc;
  UnitOps &un = *pc->unit;
  const CaseDev &d = un.desc;
  const Grid &G = c->h.G;
  // the solution systems read the stage state, the flux systems the flux
  // buffers; the systems of the other biases and of the other sweep are
  // solved too, on the state's own flux, and discarded
  stage_state(pc, uflag ? fC : u_ghosted, st);
  launch_class(un.eng.d_cases, un.eng.plan, K_FLUX, 0, st);
  if (!uflag)
    launch_dev_planar(G, fC, dir ? nullptr : d.flux_x, dir ? d.flux_y : nullptr, 1, 0, nullptr, st);
  for (int cls : {K_ENDS, K_RECON, K_RECONF, K_RCHAR, K_CWROWS, K_CWSOLVE})
    launch_class(un.eng.d_cases, un.eng.plan, cls, 0, st);
  const int r = (uflag ? 0 : 2) + (upw > 0 ? 0 : 1);
  const SweepGeom &S = dir ? d.Sy : d.Sx;
  launch_dev_iface(G.ndims, dir, G.nx, S.N, S.nlines, G.nv, dir ? d.iy[r] : d.ix[r], fI, 0, st);
  return launched(pc);
}

int piafs_dev_upwind(piafs_case *pc, double *fI, const double *fL, const double *fR, const double *uL,
                     const double *uR, const double *u_ghosted, int dir, double t, void *stream) {
  (void)t;  // the upwinders of this path do not depend on time
  if (!pc || !pc->c) return 1;
  if (!fI || !fL || !fR || !uL || !uR || !u_ghosted) return fail(pc, 1, "an array pointer is NULL");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) return fail(pc, 1, "dir must be 0 .. ndims-1");
  const hipStream_t st = stream_of(pc, stream);
  int rc = unit_prepare(pc, st);
  if (rc) return rc;
  UnitOps &un = *pc->unit;
  const CaseDev &d = un.desc;
  const Grid &G = c->h.G;
  const SweepGeom &S = dir ? d.Sy : d.Sx;
  double *const *io = dir ? d.iy : d.ix;
  stage_state(pc, u_ghosted, st);
  const double *in[4] = {uL, uR, fL, fR};
  for (int r = 0; r < 4; r++) launch_dev_iface(G.ndims, dir, G.nx, S.N, S.nlines, G.nv, in[r], io[r], 1, st);
  launch_class(un.eng.d_cases, un.eng.plan, K_UPWIND, 0, st);
  launch_dev_iface(G.ndims, dir, G.nx, S.N, S.nlines, G.nv, dir ? d.fIy : d.fIx, fI, 0, st);
  return launched(pc);
}

int piafs_dev_first_derivative_par(piafs_case *pc, double *Df, const double *f, int dir, int bias, void *stream) {
  if (!pc || !pc->c) return 1;
  if (!Df || !f) return fail(pc, 1, "an array pointer is NULL");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) return fail(pc, 1, "dir must be 0 .. ndims-1");
  launch_dev_fd(aos_grid(c), dir, c->h.P.par4, bias, f, Df, stream_of(pc, stream));
  return launched(pc);
}

int piafs_dev_weno_weights(piafs_case *pc, const double *fC, const double *u_ghosted, int dir, double *w1,
                           double *w2, double *w3, long size, void *stream) {
  if (!pc || !pc->c) return 1;
  if (!fC || !u_ghosted || !w1 || !w2 || !w3) return fail(pc, 1, "an array pointer is NULL");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) return fail(pc, 1, "dir must be 0 .. ndims-1");
  const ReconParams &R = c->h.R;
  if (R.kind == 4) return fail(pc, 1, "upw5 interpolation has no nonlinear weights");
  const SweepGeom &S = dir ? c->h.Sy : c->h.Sx;
  if (size < (long)(S.N + 1) * S.nlines * c->cfg.nvars) return fail(pc, 1, "size is smaller than one set of weights");
  launch_dev_weights(aos_grid(c), c->h.model, c->h.P.gamma, dir, R.kind, R.crweno, R.characteristic, R.no_limiting,
                     R.eps, fC, u_ghosted, w1, w2, w3, size, stream_of(pc, stream));
  return launched(pc);
}

}  // extern "C"
