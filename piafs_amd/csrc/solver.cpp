// Case runtime and the C ABI of include/piafs_amd.h.
//
// One case = one PIAFS simulation resident in HBM.  A time step is a fixed
// sequence of stream-ordered kernel launches (no host synchronisation):
// per RK stage  stage+BC -> x-sweep (flux, reconstruction, upwind,
// divergence) -> y-sweep -> viscous derivatives -> assembly of Udot; after
// the last stage the solution update.  See TimeRK.c:34-102 for the reference
// control flow that this replaces.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/piafs_amd.h"
#include "case.hpp"

using namespace piafs;

namespace piafs {

static const char *kClassNames[NCLASS] = {"stage_bc", "flux", "recon", "upwind", "divergence",
                                          "transpose", "qderiv", "viscflux", "assemble", "update"};

const char *class_name(int c) { return kClassNames[c]; }

int Case::fail(int code, const std::string &msg) {
  err = msg;
  return code;
}

#define HIPCHK(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(-1, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

double *Case::dalloc(size_t n) {
  void *p = nullptr;
  if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
  bytes += n * sizeof(double);
  allocs.push_back(p);
  hipMemsetAsync(p, 0, n * sizeof(double), stream);
  return (double *)p;
}

size_t Case::plan_bytes(const Config &c) {
  Grid g = make_grid(c);
  size_t plane = g.plane, nv = c.nvars;
  size_t n = 0;
  n += nv * plane * (3 + c.nstages + 1);  // u, U, UT, Udot[s], flux
  size_t iface = 0;
  for (int d = 0; d < c.ndims; d++) {
    size_t lines = (c.ndims == 2) ? (size_t)c.dim[1 - d] : 1;
    iface = std::max(iface, nv * (size_t)(c.dim[d] + 1) * lines);
  }
  n += 5 * iface;
  if (c.crweno) n += 8 * iface;
  if (c.ndims == 2) {  // y-sweep buffers of the second stream
    n += nv * plane + 4 * iface;
    if (c.crweno) n += 8 * iface;
  }
  size_t nint = (size_t)c.npts_int;
  n += 2 * nv * nint;                      // hx, hxT
  if (c.model_id == MODEL_NS2D) n += 12 * (size_t)g.nxg * g.nyg;
  n += plane + c.x.size() + 4096;
  return n * sizeof(double);
}

Grid make_grid(const Config &c) {
  Grid G{};
  G.ndims = c.ndims;
  G.nv = c.nvars;
  G.g = c.ghosts;
  G.nx = c.dim[0];
  G.ny = (c.ndims == 2) ? c.dim[1] : 1;
  G.gy = (c.ndims == 2) ? c.ghosts : 0;
  G.nxg = G.nx + 2 * G.g;
  G.nyg = G.ny + 2 * G.gy;
  long pl = (long)G.nxg * G.nyg;
  G.plane = (pl + 63) / 64 * 64;
  return G;
}

int Case::init(const std::string &dir, int dev) {
  if (cfg.parse(dir, err)) return 1;
  device = dev;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  G = make_grid(cfg);
  const int nv = cfg.nvars;
  // sweeps
  Sx.N = G.nx; Sx.g = G.g; Sx.plane = G.plane;
  if (cfg.ndims == 1) {
    Sx.nlines = 1; Sx.pos_stride = 1; Sx.line_off = 0;
  } else {
    Sx.nlines = G.ny; Sx.pos_stride = G.nyg; Sx.line_off = G.gy;   // transposed layout
    Sy.N = G.ny; Sy.g = G.g; Sy.plane = G.plane;
    Sy.nlines = G.nx; Sy.pos_stride = G.nxg; Sy.line_off = G.g;    // natural layout
  }
  // boundary map (one zone per face; later zones on the same face replace earlier ones)
  for (int d = 0; d < 2; d++)
    for (int f = 0; f < 2; f++) bc.f[d][f] = FaceBC{-1, 0, 0};
  for (const Boundary &b : cfg.bcs) {
    if (!b.on) continue;
    FaceBC fb;
    fb.type = (b.bctype == BC_PERIODIC) ? 0 : 1;
    if (cfg.ndims == 2) {
      int o = 1 - b.dim;
      fb.lo = b.is[o];
      fb.hi = b.ie[o];
    } else {
      fb.lo = 0;
      fb.hi = 1;
    }
    bc.f[b.dim][b.face == 1 ? 0 : 1] = fb;
  }
  // physics parameters
  memset(&P, 0, sizeof(P));
  P.gamma = cfg.gamma;
  P.viscous = (cfg.model_id == MODEL_NS2D && cfg.Re > 0) ? 1 : 0;
  P.par4 = cfg.par4 ? 1 : 0;
  if (cfg.model_id == MODEL_NS2D) {
    P.inv_Re = 1.0 / cfg.Re;
    P.inv_Pr = 1.0 / cfg.Pr;
    P.inv_gm1 = 1.0 / (cfg.gamma - 1.0);
    P.Tref = cfg.Tref; P.T0 = cfg.T0; P.TS = cfg.TS; P.TA = cfg.TA; P.TB = cfg.TB;
    P.two_third = 2.0 / 3.0;
    P.kappa_num = cfg.T0 + cfg.TA * exp(-cfg.TB / cfg.T0);
  }
  P.nflow = cfg.nflow;
  if (cfg.include_chem) {
    const Chem &ch = cfg.chem;
    P.chem = 1;
    P.nlayers = ch.nlayers;
    P.z_stride = ch.z_stride;
    P.k0a = ch.k0a; P.k0b = ch.k0b; P.k1a = ch.k1a; P.k1b = ch.k1b; P.k2a = ch.k2a; P.k2b = ch.k2b;
    P.k3a = ch.k3a; P.k3b = ch.k3b; P.k4 = ch.k4; P.k5 = ch.k5; P.k6 = ch.k6;
    P.q0a = ch.q0a; P.q0b = ch.q0b; P.q1a = ch.q1a; P.q1b = ch.q1b; P.q2a = ch.q2a; P.q2b = ch.q2b;
    P.q3a = ch.q3a; P.q3b = ch.q3b; P.q4 = ch.q4; P.q5 = ch.q5; P.q6 = ch.q6;
    P.ch_gm1 = ch.gamma - 1.0;
    P.hnu_den = ch.c_light * ch.h_planck * ch.nu * ch.n_O2;
    P.damp_coef = ch.dz * (ch.sO3 * ch.n_O2);
  }
  R.crweno = cfg.crweno ? 1 : 0;
  R.kind = cfg.weight_kind;
  R.no_limiting = cfg.no_limiting;
  R.eps = cfg.eps;
  // device buffers
  const size_t st = (size_t)nv * G.plane;
  u = dalloc(st);
  U = dalloc(st);
  if (cfg.ndims == 2) UT = dalloc(st);
  for (int s = 0; s < cfg.nstages; s++) Udot[s] = dalloc(st);
  flux = dalloc(st);
  size_t iface = 0;
  iface = std::max(iface, (size_t)nv * (Sx.N + 1) * Sx.nlines);
  if (cfg.ndims == 2) iface = std::max(iface, (size_t)nv * (Sy.N + 1) * Sy.nlines);
  for (int k = 0; k < 4; k++) ifc[k] = dalloc(iface);
  fIx = dalloc(iface);
  if (cfg.ndims == 2) fIy = dalloc((size_t)nv * (Sy.N + 1) * Sy.nlines);
  if (cfg.crweno) {
    scr_b = dalloc(4 * iface);
    scr_c = dalloc(4 * iface);
  }
  if (cfg.ndims == 2) {
    dual = true;
    HIPCHK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    const size_t ify = (size_t)nv * (Sy.N + 1) * Sy.nlines;
    flux_y = dalloc(st);
    for (int k = 0; k < 4; k++) ifcy[k] = dalloc(ify);
    if (cfg.crweno) {
      scr_by = dalloc(4 * ify);
      scr_cy = dalloc(4 * ify);
    }
  }
  const size_t nint = (size_t)G.nx * G.ny;
  hx = dalloc((size_t)nv * nint);
  if (cfg.ndims == 2) hxT = dalloc((size_t)nv * nint);
  if (P.viscous) {
    Qd = dalloc(6 * (size_t)G.nxg * G.nyg);
    FV = dalloc(6 * (size_t)G.nxg * G.nyg);
  }
  dxinv = dalloc(cfg.dxinv.size());
  imap = dalloc(G.plane);
  nparts = cfl_norm_blocks(G);
  partials = dalloc(2 * (size_t)nparts);
  uprev = dalloc(st);
  if (!u || !U || !flux || !dxinv || !imap || !hx || !partials || !uprev)
    return fail(-2, "device allocation failed");
  HIPCHK(hipMemcpyAsync(dxinv, cfg.dxinv.data(), cfg.dxinv.size() * sizeof(double), hipMemcpyHostToDevice,
                        stream));
  if (cfg.include_chem) {
    std::vector<double> im(G.plane, 0.0);
    for (long p = 0; p < (long)cfg.imap.size(); p++) im[p] = cfg.imap[p];
    HIPCHK(hipMemcpyAsync(imap, im.data(), G.plane * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  t = (double)cfg.restart_iter * cfg.dt;
  int rc = upload(cfg.u0.data());
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(stream));
  return 0;
}

Case::~Case() {
  if (device >= 0) hipSetDevice(device);
  for (auto &e : ev) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  for (void *p : allocs) hipFree(p);
  if (ev_fork) hipEventDestroy(ev_fork);
  if (ev_join) hipEventDestroy(ev_join);
  if (stream2) hipStreamDestroy(stream2);
  if (stream) hipStreamDestroy(stream);
}

// point-major interior (WriteBinary order) <-> natural planar ghosted
int Case::upload(const double *uh) {
  const int nv = cfg.nvars;
  std::vector<double> buf((size_t)nv * G.plane, 0.0);
  for (long q = 0; q < cfg.npts_int; q++) {
    int i = (int)(q % G.nx), j = (int)(q / G.nx);
    long p = (long)(j + G.gy) * G.nxg + (i + G.g);
    for (int v = 0; v < nv; v++) buf[(size_t)v * G.plane + p] = uh[(size_t)q * nv + v];
  }
  HIPCHK(hipMemcpyAsync(u, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return 0;
}

int Case::download(double *uh, const double *src) {
  const int nv = cfg.nvars;
  std::vector<double> buf((size_t)nv * G.plane);
  HIPCHK(hipMemcpyAsync(buf.data(), src ? src : u, buf.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  for (long q = 0; q < cfg.npts_int; q++) {
    int i = (int)(q % G.nx), j = (int)(q / G.nx);
    long p = (long)(j + G.gy) * G.nxg + (i + G.g);
    for (int v = 0; v < nv; v++) uh[(size_t)q * nv + v] = buf[(size_t)v * G.plane + p];
  }
  return 0;
}

void Case::mark(int cls, bool begin, hipStream_t st) {
  if (!profiling) return;
  if (!st) st = stream;
  if (begin) {
    if (ev_used >= ev.size()) {
      hipEvent_t a, b;
      hipEventCreate(&a);
      hipEventCreate(&b);
      ev.push_back({a, b});
      ev_cls.push_back(cls);
    }
    ev_cls[ev_used] = cls;
    hipEventRecord(ev[ev_used].first, st);
  } else {
    hipEventRecord(ev[ev_used].second, st);
    ev_used++;
  }
}

// Right-hand side of stage state U (already holding the BC-filled state) into
// out: TimeRHSFunctionExplicit.c:29-77.
void Case::rhs_from_stage(double time, double *out) {
  const int model = cfg.model_id;
  hipStream_t sy = dual ? stream2 : stream;
  double *fy = dual ? flux_y : flux;
  double *const *iy = dual ? ifcy : ifc;
  double *sby = dual ? scr_by : scr_b, *scy = dual ? scr_cy : scr_c;
  if (dual) {
    hipEventRecord(ev_fork, stream);
    hipStreamWaitEvent(stream2, ev_fork, 0);
  }
  if (cfg.ndims == 2) {
    // ---- y direction (natural layout) and viscous derivatives ----
    mark(1, true, sy);
    launch_flux(G, Sy, model, 1, cfg.gamma, U, fy, sy);
    mark(1, false, sy);
    mark(2, true, sy);
    launch_recon_components(G, Sy, R, U, fy, iy[0], iy[1], iy[2], iy[3], sby, scy, sy);
    mark(2, false, sy);
    mark(3, true, sy);
    launch_upwind(G, Sy, model, cfg.upwind, 1, cfg.gamma, U, iy[0], iy[1], iy[2], iy[3], fIy, sy);
    mark(3, false, sy);
    if (P.viscous) {
      mark(6, true, sy);
      launch_qderiv(G, P, U, dxinv, Qd, sy);
      mark(6, false, sy);
      mark(7, true, sy);
      launch_viscflux(G, P, U, Qd, FV, sy);
      mark(7, false, sy);
    }
    if (dual) hipEventRecord(ev_join, stream2);
  }
  // ---- x direction (transposed copy in 2D) ----
  const double *qx = (cfg.ndims == 2) ? UT : U;
  mark(1, true);
  launch_flux(G, Sx, model, 0, cfg.gamma, qx, flux, stream);
  mark(1, false);
  mark(2, true);
  if (cfg.characteristic)
    launch_recon_char_euler1d(G, Sx, R, cfg.gamma, qx, flux, ifc[0], ifc[1], ifc[2], ifc[3], stream);
  else
    launch_recon_components(G, Sx, R, qx, flux, ifc[0], ifc[1], ifc[2], ifc[3], scr_b, scr_c, stream);
  mark(2, false);
  mark(3, true);
  launch_upwind(G, Sx, model, cfg.upwind, 0, cfg.gamma, qx, ifc[0], ifc[1], ifc[2], ifc[3], fIx, stream);
  mark(3, false);
  mark(4, true);
  launch_divergence(G, Sx, dxinv, fIx, (cfg.ndims == 2) ? hxT : hx, stream);
  mark(4, false);
  if (cfg.ndims == 2) {
    mark(5, true);
    launch_transpose(cfg.nvars, G.nx, G.ny, hxT, hx, stream);
    mark(5, false);
  }
  if (dual) hipStreamWaitEvent(stream, ev_join, 0);
  // ---- source time factor (ChemistrySolve.c:36-43), evaluated on the host ----
  double hs = 0.0;
  if (cfg.include_chem) {
    const Chem &ch = cfg.chem;
    if (time > ch.t_start_norm) {
      double sigma = ch.t_pulse_norm / 2.35482;
      double tp = time - (ch.t_start_norm + ch.t_pulse_norm);
      hs = ch.I0 * exp(-(tp * tp) / (2 * sigma * sigma));
    }
  }
  mark(8, true);
  launch_assemble(G, P, U, hx, fIy, FV, dxinv, imap, hs, out, stream);
  mark(8, false);
}

void Case::enqueue_step() {
  const int ns = cfg.nstages;
  for (int s = 0; s < ns; s++) {
    const double stage_time = t + cfg.c[s] * cfg.dt;
    const double *dd[3] = {nullptr, nullptr, nullptr};
    double aa[3] = {0, 0, 0};
    int np = 0;
    for (int j = 0; j < s; j++) {
      const double a = cfg.dt * cfg.A[s * ns + j];
      if (a == 0.0) continue;  // U += 0*Udot leaves U unchanged
      dd[np] = Udot[j];
      aa[np] = a;
      np++;
    }
    mark(0, true);
    launch_stage_bc(G, bc, u, dd, aa, np, U, UT, stream);
    mark(0, false);
    rhs_from_stage(stage_time, Udot[s]);
  }
  double bb[4];
  const double *dd[4];
  for (int s = 0; s < ns; s++) {
    bb[s] = cfg.dt * cfg.b[s];
    dd[s] = Udot[s];
  }
  mark(9, true);
  launch_update(G, u, dd, bb, ns, stream);
  mark(9, false);
  t += cfg.dt;  // TimePostStep.c:33
}

int Case::step(int n, double *cfl, double *norm) {
  HIPCHK(hipSetDevice(device));
  for (int it = 0; it < n; it++) {
    const bool diag = (cfl || norm);
    if (diag) HIPCHK(hipMemcpyAsync(uprev, u, (size_t)cfg.nvars * G.plane * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream));
    enqueue_step();
    if (diag) {
      launch_cfl_norm(G, cfg.model_id, cfg.gamma, cfg.dt, u, uprev, dxinv, partials, nparts, stream);
      std::vector<double> h(2 * nparts);
      HIPCHK(hipMemcpyAsync(h.data(), partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      double mx = 0, sm = 0;
      for (int b = 0; b < nparts; b++) {
        mx = std::max(mx, h[2 * b]);
        sm += h[2 * b + 1];
      }
      if (cfl) cfl[it] = mx;
      if (norm) norm[it] = sqrt(sm / (double)cfg.npts_int);
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int Case::rhs_host(const double *uh, double time, double *rh) {
  // state into u (restored afterwards), stage = u with BCs, RHS into Udot[0]
  const size_t st = (size_t)cfg.nvars * G.plane;
  HIPCHK(hipMemcpyAsync(uprev, u, st * sizeof(double), hipMemcpyDeviceToDevice, stream));
  int rc = upload(uh);
  if (rc) return rc;
  launch_stage_bc(G, bc, u, nullptr, nullptr, 0, U, UT, stream);
  rhs_from_stage(time, Udot[0]);
  rc = download(rh, Udot[0]);
  HIPCHK(hipMemcpyAsync(u, uprev, st * sizeof(double), hipMemcpyDeviceToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return rc;
}

int Case::write_binary(const std::string &path) {
  std::vector<double> uh((size_t)cfg.npts_int * cfg.nvars);
  int rc = download(uh.data());
  if (rc) return rc;
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return fail(2, "could not open " + path + " for writing");
  int nd = cfg.ndims, nv = cfg.nvars;
  fwrite(&nd, sizeof(int), 1, f);
  fwrite(&nv, sizeof(int), 1, f);
  fwrite(cfg.dim, sizeof(int), nd, f);
  long off = 0;
  for (int d = 0; d < nd; d++) {
    fwrite(&cfg.x[off + cfg.ghosts], sizeof(double), cfg.dim[d], f);
    off += cfg.dim[d] + 2 * cfg.ghosts;
  }
  fwrite(uh.data(), sizeof(double), uh.size(), f);
  fclose(f);
  return 0;
}

}  // namespace piafs

// ===========================================================================
// C ABI
// ===========================================================================
struct piafs_case : public piafs::Case {};

static std::string g_create_error;

extern "C" {

int piafs_create(const char *dir, int device, piafs_case **out) {
  *out = nullptr;
  piafs_case *c = new piafs_case();
  int rc = c->init(dir ? dir : "", device);
  if (rc) {
    g_create_error = c->err;
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int piafs_destroy(piafs_case *c) {
  delete c;
  return 0;
}

const char *piafs_last_error(const piafs_case *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int piafs_get_info(const piafs_case *c, piafs_info *info) {
  if (!c || !info) return 1;
  memset(info, 0, sizeof(*info));
  info->ndims = c->cfg.ndims;
  info->nvars = c->cfg.nvars;
  info->ghosts = c->cfg.ghosts;
  for (int d = 0; d < 3; d++) info->dim[d] = c->cfg.dim[d];
  info->n_iter = c->cfg.n_iter;
  info->restart_iter = c->cfg.restart_iter;
  info->nstages = c->cfg.nstages;
  info->dt = c->cfg.dt;
  info->t = c->t;
  info->gamma = c->cfg.gamma;
  info->Re = c->cfg.Re;
  info->Pr = c->cfg.Pr;
  info->model = c->cfg.model_id;
  info->scheme = c->cfg.crweno ? 1 : 0;
  info->characteristic = c->cfg.characteristic ? 1 : 0;
  info->include_chem = c->cfg.include_chem ? 1 : 0;
  info->device_bytes = c->bytes;
  return 0;
}

int piafs_upload_state(piafs_case *c, const double *u) { return c->upload(u); }
int piafs_download_state(piafs_case *c, double *u) { return c->download(u); }
int piafs_set_time(piafs_case *c, double t) {
  c->t = t;
  return 0;
}
int piafs_step(piafs_case *c, int n) { return c->step(n, nullptr, nullptr); }
int piafs_step_diag(piafs_case *c, int n, double *cfl, double *norm) {
  int rc = c->step(n, cfl, norm);
  if (rc) return rc;
  return piafs_sync(c);
}
int piafs_sync(piafs_case *c) {
  hipSetDevice(c->device);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->fail(-1, std::string("stream: ") + hipGetErrorString(e));
  return 0;
}
int piafs_rhs(piafs_case *c, const double *u, double t, double *rhs) { return c->rhs_host(u, t, rhs); }
int piafs_write_binary(piafs_case *c, const char *path) { return c->write_binary(path); }

int piafs_kernel_times(piafs_case *c, int enable, const char **names, double *ms, int *n) {
  if (enable >= 0) {
    c->profiling = (enable != 0);
    c->ev_used = 0;
  }
  if (!n) return 0;
  int cap = *n;
  piafs_sync(c);
  double tot[NCLASS] = {0};
  int cnt[NCLASS] = {0};
  for (size_t k = 0; k < c->ev_used; k++) {
    float m = 0;
    hipEventElapsedTime(&m, c->ev[k].first, c->ev[k].second);
    tot[c->ev_cls[k]] += m;
    cnt[c->ev_cls[k]]++;
  }
  int w = 0;
  for (int k = 0; k < NCLASS && w < cap; k++) {
    if (!cnt[k]) continue;
    if (names) names[w] = piafs::class_name(k);
    if (ms) ms[w] = tot[k] / cnt[k];
    w++;
  }
  // second half of ms (if room) holds the launch counts
  *n = w;
  c->ev_used = 0;
  return 0;
}

int piafs_estimate_bytes(const char *dir, size_t *bytes) {
  piafs::Config cfg;
  std::string err;
  if (cfg.parse(dir ? dir : "", err)) {
    g_create_error = err;
    return 1;
  }
  *bytes = piafs::Case::plan_bytes(cfg);
  return 0;
}

int piafs_solve(piafs_case *c, const char *dir, int verbose, double *seconds) {
  const Config &cf = c->cfg;
  std::string d = dir ? dir : "";
  auto outname = [&](int index) {
    std::string fn = d.empty() ? "op" : d + "/op";
    if (cf.op_overwrite == "no") {
      char b[16];
      snprintf(b, sizeof b, "_%05d", index);
      fn += b;
    }
    return fn + ".bin";
  };
  bool binary = (cf.op_file_format == "binary" || cf.op_file_format == "bin");
  int index = 0;
  if (cf.restart_iter) index = cf.restart_iter / cf.file_op_iter;  // InitializeSolvers.c:240-245
  auto t0 = std::chrono::steady_clock::now();
  std::vector<double> cfl(1), norm(1);
  for (int iter = cf.restart_iter; iter < cf.n_iter; iter++) {
    if (iter == 0 && binary) {
      if (c->write_binary(outname(index))) return c->err.empty() ? -1 : 2;
      index++;
    }
    bool screen = ((iter + 1) % cf.screen_op_iter == 0);
    int rc = screen && verbose ? c->step(1, cfl.data(), norm.data()) : c->step(1, nullptr, nullptr);
    if (rc) return rc;
    if (screen && verbose)
      printf("iter=%7d  t=%1.3E  CFL=%1.3E  norm=%1.4E\n", iter + 1, c->t, cfl[0], norm[0]);
    if (((iter + 1) % cf.file_op_iter == 0) && ((iter + 1) < cf.n_iter) && binary) {
      if (c->write_binary(outname(index))) return 2;
      index++;
    }
  }
  int rc = piafs_sync(c);
  if (rc) return rc;
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (seconds) *seconds = secs;
  if (binary) {
    if (c->write_binary(outname(index))) return 2;
  }
  return 0;
}

}  // extern "C"
