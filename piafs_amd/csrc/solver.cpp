// Case runtime, ensemble engine and the single-case C ABI of
// include/piafs_amd.h.
//
// A time step is a fixed sequence of stream-ordered launches with no host
// synchronisation: per RK stage  stage+BC -> fluxes -> reconstruction (x and
// y sweeps in one launch) -> upwinding -> x divergence -> transpose ->
// viscous derivatives -> assembly of Udot; after the last stage the update.
// TimeRK.c:34-102 is the reference control flow this replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/piafs_amd.h"
#include "case.hpp"

namespace piafs {

#define HIPCHK(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(-1, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

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

// ---------------------------------------------------------------------------
// Case
// ---------------------------------------------------------------------------
int Case::fail(int code, const std::string &msg) {
  err = msg;
  return code;
}

double *Case::dalloc(size_t n, hipStream_t st) {
  void *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
  bytes += n * sizeof(double);
  allocs.push_back(p);
  if (hipMemsetAsync(p, 0, n * sizeof(double), st) != hipSuccess) return nullptr;
  return (double *)p;
}

// Paired reconstruction (two interface arrays per sweep instead of four):
// component-wise interpolation with Roe or Rusanov upwinding, which read only
// 0.5*(uR - uL) and 0.5*(fL + fR).  Characteristic interpolation and the
// rf-char / llf-char upwinders need uL, uR, fL, fR separately.
static bool paired_mode(const Config &c) {
  return !c.characteristic && (c.upwind == UPW_ROE || c.upwind == UPW_RUSANOV);
}

// Cases whose species Roe fluxes the reconstruction can write (NS2D with
// species, paired Roe, component-wise): they keep |v_n| per interface.
static bool species_dsp(const Config &c) {
  return PIAFS_RECON_SPECIES && c.model_id == MODEL_NS2D && c.ndims == 2 && c.nvars > 4 && paired_mode(c) &&
         c.upwind == UPW_ROE;
}

size_t Case::plan_bytes(const Config &c) {
  Grid g = make_grid(c);
  size_t plane = g.plane, nv = c.nvars;
  size_t n = nv * plane * (3 + c.nstages + 3);  // u, U, UT, Udot[s], flux_x, flux_y, uprev
  size_t ix = nv * (size_t)(c.dim[0] + 1) * (c.ndims == 2 ? c.dim[1] : 1);
  size_t iy = c.ndims == 2 ? nv * (size_t)(c.dim[1] + 1) * c.dim[0] : 0;
  n += (paired_mode(c) ? 3 : 5) * (ix + iy);
  if (species_dsp(c)) n += (ix + iy) / nv;
  if (c.characteristic && c.crweno) {
    const int nf = c.model_id == MODEL_EULER1D ? 3 : 4;
    n += (size_t)(cw_row_fields(nf, (int)nv) + 4 * cw_back_fields(nf, (int)nv)) * (ix + iy) / nv;
  }
  if (c.model_id == MODEL_NS2D) n += 12 * (size_t)g.nxg * g.nyg;
  if (!c.characteristic) {  // long-line scratch of sweeps longer than one team
    int T, m;
    for (int d = 0; d < c.ndims; d++)
      if (!team_shape(c.dim[d], kMaxRows, T, m))
        n += (size_t)4 * (c.dim[d] + 1) * (c.ndims == 2 ? c.dim[1 - d] : 1) * nv * 4;
  }
  n += plane + c.x.size() + 4096;
  return n * sizeof(double);
}

int Case::init(const std::string &dir, int dev, hipStream_t st) {
  if (cfg.parse(dir, err)) return 1;
  return setup(dev, st);
}

int Case::setup(int dev, hipStream_t st) {
  device = dev;
  HIPCHK(hipSetDevice(device));
  CaseDev &D = h;
  memset(&D, 0, sizeof(D));
  D.G = make_grid(cfg);
  const Grid &G = D.G;
  const int nv = cfg.nvars;
  // sweep geometry: x lines are rows of the natural layout, y lines rows of
  // the transposed copy
  D.Sx.N = G.nx; D.Sx.g = G.g; D.Sx.plane = G.plane;
  if (cfg.ndims == 1) {
    D.Sx.nlines = 1; D.Sx.base = G.g; D.Sx.line_stride = G.nxg;
  } else {
    D.Sx.nlines = G.ny; D.Sx.base = (long)G.gy * G.nxg + G.g; D.Sx.line_stride = G.nxg;
    D.Sy.N = G.ny; D.Sy.g = G.g; D.Sy.plane = G.plane;
    D.Sy.nlines = G.nx; D.Sy.base = (long)G.g * G.nyg + G.gy; D.Sy.line_stride = G.nyg;
  }
  if (set_team_rows(kMaxRows)) return 1;
  // boundary zones, in list order (ApplyBoundaryConditions.c:20-40)
  std::vector<double> zvals;
  D.nzones = 0;
  D.sponge = 0;
  for (const Boundary &b : cfg.bcs) {
    if (!b.on) continue;
    if (D.nzones == kMaxZones) return fail(1, "more than " + std::to_string(kMaxZones) + " boundary zones");
    Zone &z = D.zones[D.nzones++];
    z.kind = b.bctype;
    z.dim = b.dim;
    z.face = b.face;
    for (int d = 0; d < 2; d++) {
      z.is[d] = d < cfg.ndims ? b.is[d] : 0;
      z.ie[d] = d < cfg.ndims ? b.ie[d] : 1;
    }
    z.rho = b.rho;
    z.vel[0] = b.vel[0];
    z.vel[1] = b.vel[1];
    z.p = b.p;
    z.gamma = b.gamma;
    z.xmin = b.xmin[b.dim];
    z.xmax = b.xmax[b.dim];
    z.vals = -1;
    if (!b.vals.empty()) {
      z.vals = (int)zvals.size();
      zvals.insert(zvals.end(), b.vals.begin(), b.vals.end());
    }
    if (b.bctype == BC_SPONGE) D.sponge = 1;
  }
  PhysParams &P = D.P;
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
    P.I0 = ch.I0;
    P.t_start_norm = ch.t_start_norm;
    P.t_pulse_norm = ch.t_pulse_norm;
  }
  D.terms = 7;
  D.R.crweno = cfg.crweno ? 1 : 0;
  D.R.kind = cfg.weight_kind;
  D.R.no_limiting = cfg.no_limiting;
  D.R.characteristic = cfg.characteristic ? 1 : 0;
  D.R.eps = cfg.eps;
  {
    const char *fr = getenv("PIAFS_AMD_LINE_SOLVE_PLAIN_START");  // run the relay from trivial estimates
    D.R.plain_start = (fr && fr[0] == '1') ? 1 : 0;
  }
  D.model = cfg.model_id;
  D.upwind = cfg.upwind;
  D.paired = paired_mode(cfg) ? 1 : 0;
  D.nstages = cfg.nstages;
  D.dt = cfg.dt;
  for (int k = 0; k < 16; k++) D.A[k] = cfg.A[k];
  for (int k = 0; k < 4; k++) { D.b[k] = cfg.b[k]; D.c[k] = cfg.c[k]; }
  // device buffers
  const size_t st_sz = (size_t)nv * G.plane;
  const bool two = (cfg.ndims == 2);
  D.u = dalloc(st_sz, st);
  D.U = dalloc(st_sz, st);
  if (two) D.UT = dalloc(st_sz, st);
  for (int s = 0; s < cfg.nstages; s++) D.Udot[s] = dalloc(st_sz, st);
  D.flux_x = dalloc(st_sz, st);
  const size_t ix = (size_t)nv * (D.Sx.N + 1) * D.Sx.nlines;
  const int narr = D.paired ? 2 : 4;
  for (int k = 0; k < narr; k++) D.ix[k] = dalloc(ix, st);
  D.fIx = dalloc(ix, st);
  if (D.Sx.longl) D.lrow[0] = dalloc(lrow_doubles(D.Sx, nv), st);
  if (two && D.Sy.longl) D.lrow[1] = dalloc(lrow_doubles(D.Sy, nv), st);
  if ((D.Sx.longl && !D.lrow[0]) || (two && D.Sy.longl && !D.lrow[1])) return fail(-2, "device allocation failed");
  if (two) {
    const size_t iy = (size_t)nv * (D.Sy.N + 1) * D.Sy.nlines;
    D.flux_y = dalloc(st_sz, st);
    for (int k = 0; k < narr; k++) D.iy[k] = dalloc(iy, st);
    D.fIy = dalloc(iy, st);
  }
  D.dsp[0] = D.dsp[1] = nullptr;
  if (species_dsp(cfg)) {  // |v_n| per interface, for the species fluxes written by the reconstruction
    D.dsp[0] = dalloc((size_t)(D.Sx.N + 1) * D.Sx.nlines, st);
    D.dsp[1] = dalloc((size_t)(D.Sy.N + 1) * D.Sy.nlines, st);
    if (!D.dsp[0] || !D.dsp[1]) return fail(-2, "device allocation failed");
  }
  if (P.viscous) {
    D.Qd = dalloc(6 * (size_t)G.nxg * G.nyg, st);
    D.FV = dalloc(6 * (size_t)G.nxg * G.nyg, st);
  }
  D.dxinv = dalloc(cfg.dxinv.size(), st);
  D.xc = dalloc(cfg.x.size(), st);
  D.bcvals = dalloc(zvals.size() + 1, st);
  D.imap = dalloc(G.plane, st);
  D.time = dalloc(1, st);
  // characteristic CRWENO5 scratch (block rows and back-substitution data per sweep)
  if (cfg.characteristic && cfg.crweno) {
    const int nf = cfg.model_id == MODEL_EULER1D ? 3 : 4;
    const SweepGeom *sw[2] = {&D.Sx, &D.Sy};
    for (int d = 0; d < cfg.ndims; d++) {
      const size_t ni = (size_t)(sw[d]->N + 1) * sw[d]->nlines;
      D.cwr[d] = dalloc((size_t)cw_row_fields(nf, nv) * ni, st);
      D.cwb[d] = dalloc((size_t)cw_back_fields(nf, nv) * 4 * ni, st);
      if (!D.cwr[d] || !D.cwb[d]) return fail(-2, "device allocation failed");
    }
  }
  // conservation check (InitialSolution.c:115-135, TimePostStep.c:68-80)
  D.cons = (cfg.conservation_check == "yes") ? 1 : 0;
  if (D.cons) {
    long off = 0;
    double dxm[2] = {1.0, 1.0};  // dxinv at the middle of each dimension (BoundaryIntegral.c:38-41)
    for (int d = 0; d < cfg.ndims; d++) {
      dxm[d] = cfg.dxinv[off + cfg.ghosts + cfg.dim[d] / 2];
      off += cfg.dim[d] + 2 * cfg.ghosts;
    }
    for (int d = 0; d < 2; d++) {
      double dS = 1.0;
      for (int k = 0; k < cfg.ndims; k++)
        if (k != d) dS *= (1.0 / dxm[k]);
      D.dS[d] = dS;
    }
    D.bflux = dalloc(4 * 2 * (size_t)cfg.ndims * nv, st);
    D.tbi = dalloc(nv, st);
    if (!D.bflux || !D.tbi) return fail(-2, "device allocation failed");
  }
  if (D.paired) {
    D.ends = dalloc(4 * 2 * (size_t)nv * (D.Sx.nlines + (two ? D.Sy.nlines : 0)), st);
    if (!D.ends) return fail(-2, "device allocation failed");
  }
  nparts = cfl_norm_blocks(G);
  partials = dalloc(2 * (size_t)nparts, st);
  uprev = dalloc(st_sz, st);
  for (void *p : allocs)
    if (!p) return fail(-2, "device allocation failed");
  if (!D.u || !D.U || !D.flux_x || !D.dxinv || !D.imap || !D.time || !partials || !uprev ||
      (two && (!D.UT || !D.flux_y || !D.fIy)))
    return fail(-2, "device allocation failed");
  if (!D.xc || !D.bcvals) return fail(-2, "device allocation failed");
  HIPCHK(hipMemcpyAsync(D.dxinv, cfg.dxinv.data(), cfg.dxinv.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(D.xc, cfg.x.data(), cfg.x.size() * sizeof(double), hipMemcpyHostToDevice, st));
  if (!zvals.empty())
    HIPCHK(hipMemcpyAsync(D.bcvals, zvals.data(), zvals.size() * sizeof(double), hipMemcpyHostToDevice, st));
  if (cfg.include_chem) {
    std::vector<double> im(G.plane, 0.0);
    for (long p = 0; p < (long)cfg.imap.size(); p++) im[p] = cfg.imap[p];
    HIPCHK(hipMemcpyAsync(D.imap, im.data(), G.plane * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  t = (double)cfg.restart_iter * cfg.dt;
  HIPCHK(hipMemcpyAsync(D.time, &t, sizeof(double), hipMemcpyHostToDevice, st));
  int rc = upload(cfg.u0.data(), st);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int Case::set_team_rows(int max_rows) {
  CaseDev &D = h;
  if (cfg.characteristic) {  // characteristic WENO5 runs one thread per interface: no line teams
    D.Sx.team = D.Sy.team = 1;
    D.Sx.rows = D.Sy.rows = 1;
    return 0;
  }
  for (int d = 0; d < cfg.ndims; d++) {
    SweepGeom &S = d ? D.Sy : D.Sx;
    int T, m;
    if (!team_shape(S.N, kMaxRows, T, m)) {  // longer than a 256-lane team covers: the sequential solver
      S.longl = 1;
      S.team = S.rows = 1;
      continue;
    }
    S.longl = 0;
    if (!team_shape(S.N, max_rows, S.team, S.rows))
      return fail(1, "lines of " + std::to_string(S.N) + " points need more than 256 lanes at " +
                         std::to_string(max_rows) + " rows per lane");
  }
  return 0;
}

Case::~Case() {
  if (device >= 0) hipSetDevice(device);
  for (void *p : allocs) hipFree(p);
}

// point-major interior (WriteBinary order) <-> natural planar ghosted
int Case::upload(const double *uh, hipStream_t st) {
  const Grid &G = h.G;
  const int nv = cfg.nvars;
  std::vector<double> buf((size_t)nv * G.plane, 0.0);
  for (long q = 0; q < cfg.npts_int; q++) {
    int i = (int)(q % G.nx), j = (int)(q / G.nx);
    long p = (long)(j + G.gy) * G.nxg + (i + G.g);
    for (int v = 0; v < nv; v++) buf[(size_t)v * G.plane + p] = uh[(size_t)q * nv + v];
  }
  HIPCHK(hipMemcpyAsync(h.u, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int Case::download(double *uh, hipStream_t st, const double *src) {
  const Grid &G = h.G;
  const int nv = cfg.nvars;
  std::vector<double> buf((size_t)nv * G.plane);
  HIPCHK(hipMemcpyAsync(buf.data(), src ? src : h.u, buf.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (long q = 0; q < cfg.npts_int; q++) {
    int i = (int)(q % G.nx), j = (int)(q / G.nx);
    long p = (long)(j + G.gy) * G.nxg + (i + G.g);
    for (int v = 0; v < nv; v++) uh[(size_t)q * nv + v] = buf[(size_t)v * G.plane + p];
  }
  return 0;
}

// WriteBinary layout (WriteBinary.c:37-93)
int Case::write_binary(const std::string &path, hipStream_t st) {
  std::vector<double> uh((size_t)cfg.npts_int * cfg.nvars);
  int rc = download(uh.data(), st);
  if (rc) return rc;
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return fail(2, "could not open " + path + " for writing");
  int nd = cfg.ndims, nv = cfg.nvars;
  bool ok = fwrite(&nd, sizeof(int), 1, f) == 1 && fwrite(&nv, sizeof(int), 1, f) == 1 &&
            fwrite(cfg.dim, sizeof(int), nd, f) == (size_t)nd;
  long off = 0;
  for (int d = 0; d < nd && ok; d++) {
    ok = fwrite(&cfg.x[off + cfg.ghosts], sizeof(double), cfg.dim[d], f) == (size_t)cfg.dim[d];
    off += cfg.dim[d] + 2 * cfg.ghosts;
  }
  ok = ok && fwrite(uh.data(), sizeof(double), uh.size(), f) == uh.size();
  fclose(f);
  if (!ok) return fail(2, "short write to " + path);
  return 0;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
int Engine::fail(int code, const std::string &msg) {
  err = msg;
  return code;
}

int Engine::init(int dev) {
  device = dev;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  // PIAFS_AMD_ONE_STREAM=1 keeps every launch of a step on one stream
  const char *one = getenv("PIAFS_AMD_ONE_STREAM");
  if (!(one && one[0] == '1')) {
    HIPCHK(hipStreamCreateWithFlags(&fork.side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork.fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&fork.join, hipEventDisableTiming));
  }
  return 0;
}

int Engine::add(Case *c) {
  cases.push_back(c);
  return 0;
}

void Engine::release_plan() {
  if (gexec) hipGraphExecDestroy(gexec);
  if (graph) hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
  for (void *p : plan_allocs) hipFree(p);
  plan_allocs.clear();
  d_cases = nullptr;
  d_status = nullptr;
  plan = Plan();
  built = false;
}

int Engine::build() {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamSynchronize(stream));
  release_plan();
  const int nc = (int)cases.size();
  plan.ncases = nc;
  if (nc == 0) {
    built = true;
    return 0;
  }
  void *p = nullptr;
  HIPCHK(hipMalloc(&p, nc * sizeof(int)));
  plan_allocs.push_back(p);
  d_status = (int *)p;
  HIPCHK(hipMemsetAsync(d_status, 0, nc * sizeof(int), stream));
  std::vector<CaseDev> hd(nc);
  for (int k = 0; k < nc; k++) {
    if (override_desc && nc == 1) {
      hd[k] = *override_desc;  // the descriptor brings its own status word
    } else {
      hd[k] = cases[k]->h;
      hd[k].status = d_status + k;
    }
    hd[k].fuse_sp = species_in_recon(hd[k]) ? 1 : 0;
  }
  plan.stage64 = true;
  for (int k = 0; k < nc; k++) plan.stage64 = plan.stage64 && stage_tile_64(hd[k]);
  for (int k = 0; k < nc; k++) hd[k].stage_tile = plan.stage64 ? 64 : 32;
  HIPCHK(hipMalloc(&p, nc * sizeof(CaseDev)));
  plan_allocs.push_back(p);
  d_cases = (CaseDev *)p;
  HIPCHK(hipMemcpyAsync(d_cases, hd.data(), nc * sizeof(CaseDev), hipMemcpyHostToDevice, stream));
  for (int cls = 0; cls < K_NCLASS; cls++) {
    std::vector<int> bc, first(nc);
    for (int k = 0; k < nc; k++) {
      int b[K_NCLASS];
      case_blocks(hd[k], b);
      first[k] = (int)bc.size();
      for (int j = 0; j < b[cls]; j++) bc.push_back(k);
    }
    plan.nblocks[cls] = (int)bc.size();
    if (bc.empty()) continue;
    HIPCHK(hipMalloc(&p, bc.size() * sizeof(int)));
    plan_allocs.push_back(p);
    plan.blockcase[cls] = (int *)p;
    HIPCHK(hipMemcpyAsync(p, bc.data(), bc.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));  // bc and first are host temporaries
    HIPCHK(hipMalloc(&p, nc * sizeof(int)));
    plan_allocs.push_back(p);
    plan.firstblock[cls] = (int *)p;
    HIPCHK(hipMemcpyAsync(p, first.data(), nc * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  HIPCHK(hipStreamSynchronize(stream));  // hd is a host temporary
  built = true;
  return 0;
}

static void mark_cb(void *ctx, int cls, bool begin, hipStream_t st) { ((Engine *)ctx)->mark(cls, begin, st); }

// Events bracketing one launch of class cls, recorded on the stream the launch
// goes to (the viscous branch runs on the side stream beside the main one).
void Engine::mark(int cls, bool begin, hipStream_t st) {
  if (!profiling) return;
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

void Engine::enqueue_step() {
  int ns = 0;
  for (Case *c : cases) ns = std::max(ns, c->cfg.nstages);
  for (int s = 0; s < ns; s++)
    launch_stage(d_cases, plan, s, stream, profiling ? mark_cb : nullptr, this, fork.side ? &fork : nullptr);
  if (profiling) mark(K_UPDATE, true, stream);
  launch_update(d_cases, plan, stream);
  if (profiling) mark(K_UPDATE, false, stream);
}

// Capture one step of the current member set into a graph (once per member set).
int Engine::prepare() {
  HIPCHK(hipSetDevice(device));
  if (!built) return fail(-1, "the ensemble plan is not built");
  if (!use_graph || profiling || gexec || cases.empty()) return 0;
  HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  enqueue_step();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(stream, &g);
  if (e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return fail(-1, std::string("step graph capture: ") + hipGetErrorString(e));
  }
  graph = g;
  HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  return 0;
}

int Engine::step(int n) {
  HIPCHK(hipSetDevice(device));
  if (!built) return fail(-1, "the ensemble plan is not built");
  if (cases.empty()) return 0;
  if (use_graph && !profiling) {
    if (!gexec) {
      int rc = prepare();
      if (rc) return rc;
    }
    for (int i = 0; i < n; i++) HIPCHK(hipGraphLaunch(gexec, stream));
  } else {
    for (int i = 0; i < n; i++) enqueue_step();
  }
  HIPCHK(hipGetLastError());
  for (Case *c : cases)
    for (int i = 0; i < n; i++) c->t += c->cfg.dt;  // TimePostStep.c:33, same sum as on the device
  return 0;
}

// Device time (ms) per stage of the kernel classes `cls` and `cls2` (-1:
// none) -- the sum of their average launch times -- over `reps` evaluations
// of stage 0 of every member.  A stage leaves the solution and the time alone
// (it writes the stage state, the interface arrays and Udot[0] only).
double Engine::time_class(int cls, int reps, int cls2) {
  if (!built || cases.empty()) return 0.0;
  const bool was = profiling;
  profiling = true;
  ev_used = 0;
  for (int r = 0; r < reps; r++) launch_stage(d_cases, plan, 0, stream, mark_cb, this);
  hipStreamSynchronize(stream);
  double tot[2] = {0.0, 0.0};
  int cnt[2] = {0, 0};
  for (size_t k = 0; k < ev_used; k++) {
    const int which = ev_cls[k] == cls ? 0 : (ev_cls[k] == cls2 ? 1 : -1);
    if (which < 0) continue;
    float m = 0.f;
    hipEventElapsedTime(&m, ev[k].first, ev[k].second);
    tot[which] += m;
    cnt[which]++;
  }
  ev_used = 0;
  profiling = was;
  return (cnt[0] ? tot[0] / cnt[0] : 0.0) + (cnt[1] ? tot[1] / cnt[1] : 0.0);
}

// Segment length of the line solves (rows per lane, which sets each case's
// team size): every candidate is timed on the members themselves and the
// fastest kept.
int Engine::autotune(const int *cands, int ncand, int reps, int *best_rows, double *best_ms) {
  HIPCHK(hipSetDevice(device));
  if (cases.empty()) return 0;
  int best = -1;
  double bms = 0.0;
  for (int k = 0; k < ncand; k++) {
    bool ok = true;
    for (Case *c : cases) ok = ok && c->set_team_rows(cands[k]) == 0;
    if (!ok) continue;
    int rc = build();
    if (rc) return rc;
    time_class(K_RECON, 1, K_RECONF);  // warm
    const double ms = time_class(K_RECON, reps, K_RECONF);
    if (best < 0 || ms < bms) {
      best = cands[k];
      bms = ms;
    }
  }
  if (best < 0) return fail(1, "no candidate segment length fits the members' lines");
  for (Case *c : cases) c->set_team_rows(best);
  int rc = build();
  if (rc) return rc;
  if (best_rows) *best_rows = best;
  if (best_ms) *best_ms = bms;
  return 0;
}

int Engine::stage_only(int s, int last) {
  HIPCHK(hipSetDevice(device));
  if (!built) return fail(-1, "the ensemble plan is not built");
  launch_stage(d_cases, plan, s, stream, nullptr, nullptr, nullptr, last);
  HIPCHK(hipGetLastError());
  return 0;
}

int Engine::sync() {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamSynchronize(stream));
  if (!cases.empty() && d_status) {
    std::vector<int> st(cases.size());
    HIPCHK(hipMemcpyAsync(st.data(), d_status, st.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (size_t k = 0; k < st.size(); k++)
      if (st[k]) {
        const std::string msg = "zero pivot in a CRWENO5 line solve (tridiagLU returns -1)";
        cases[k]->err = msg;
        return fail(-3, cases.size() > 1 ? "case " + std::to_string(k) + ": " + msg : msg);
      }
  }
  return 0;
}

Engine::~Engine() {
  if (device >= 0 && stream) hipSetDevice(device);
  if (stream) hipStreamSynchronize(stream);
  release_plan();
  for (auto &e : ev) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  if (fork.side) hipStreamSynchronize(fork.side);
  if (owns_cases)
    for (Case *c : cases) delete c;
  if (fork.side) hipStreamDestroy(fork.side);
  if (fork.fork) hipEventDestroy(fork.fork);
  if (fork.join) hipEventDestroy(fork.join);
  if (stream) hipStreamDestroy(stream);
}

}  // namespace piafs

namespace piafs {

// One step of every member with the screen diagnostics of TimePreStep /
// TimePostStep (TimePostStep.c:37-60): the largest CFL number of any member
// and the norm of the change sqrt(sum over members of |u_new - u_old|^2 /
// their total points).
int step_with_diagnostics(Engine &eng, const std::vector<Case *> &cs, double *cfl, double *norm,
                          std::string &err) {
  hipStream_t st = eng.stream;
  for (Case *c : cs)
    if (hipMemcpyAsync(c->uprev, c->h.u, (size_t)c->cfg.nvars * c->h.G.plane * sizeof(double),
                       hipMemcpyDeviceToDevice, st) != hipSuccess) {
      err = "copy failed";
      return -1;
    }
  int rc = eng.step(1);
  if (rc) {
    err = eng.err;
    return rc;
  }
  double mx = 0, sm = 0, npts = 0;
  for (Case *c : cs) {
    std::vector<double> hbuf(2 * (size_t)c->nparts);
    launch_cfl_norm(c->h.G, c->cfg.model_id, c->cfg.gamma, c->cfg.dt, c->h.u, c->uprev, c->h.dxinv, c->partials,
                    c->nparts, st);
    if (hipMemcpyAsync(hbuf.data(), c->partials, hbuf.size() * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      err = "diagnostics failed";
      return -1;
    }
    for (int b = 0; b < c->nparts; b++) {
      mx = std::max(mx, hbuf[2 * b]);
      sm += hbuf[2 * b + 1];
    }
    npts += (double)c->cfg.npts_int;
  }
  if (cfl) *cfl = mx;
  if (norm) *norm = sqrt(sm / npts);
  return 0;
}

// The reference's run (Solve.cpp:28-110 with TimePrintStep.c:19-75, then
// SimWriteErrors) over the members of one engine: a single simulation, or the
// nsims members of an ensemble directory, which share solver.inp's steps, dt
// and output schedule (ReadInputs.c:259-445) and write their files under
// their member tags.
int run_solve(Engine &eng, const std::vector<Case *> &cs, const std::string &d, int verbose, double *seconds,
              std::string &err) {
  const auto t_main = std::chrono::steady_clock::now();
  const Config &cf = cs[0]->cfg;
  const int nsims = (int)cs.size();
  const std::string ext = output_extension(cf);
  const bool write = !ext.empty();
  const std::string pre = d.empty() ? std::string() : d + "/";
  auto fail_with = [&](Case *c, int rc) {
    err = (c && !c->err.empty()) ? c->err : eng.err;
    return rc;
  };
  // op[_<member>][_NNNNN]: the index advances after every write when
  // op_overwrite is "no" (OutputSolution.cpp:42-62); a restart fast-forwards
  // it (InitializeSolvers.c:240-245)
  int index = 0;
  if (cf.op_overwrite == "no" && cf.restart_iter)
    for (int it = 0; it < cf.restart_iter; it++)
      if ((it + 1) % cf.file_op_iter == 0) index++;
  auto suffix = [&]() {
    if (cf.op_overwrite != "no") return std::string();
    char b[16];
    snprintf(b, sizeof b, "_%05d", index);
    return std::string(b);
  };
  // time of each member's latest right-hand-side evaluation (none yet)
  std::vector<double> stage_time(nsims, -1.0);
  auto output = [&]() -> int {  // PhysicsOutput of every member, then OutputSolution (Solve.cpp:95-107)
    if (!write) return 0;
    for (int k = 0; k < nsims; k++)
      if (cs[k]->cfg.species_output && cs[k]->write_species(d, suffix(), stage_time[k], eng.stream))
        return fail_with(cs[k], 2);
    for (Case *c : cs) {
      std::string fn = pre + "op" + c->cfg.tag() + suffix() + ext;
      if (verbose) printf("Writing solution file %s.\n", fn.c_str());
      if (c->write_solution(fn, eng.stream)) return fail_with(c, 2);
    }
    if (cf.op_overwrite == "no") index++;
    return 0;
  };
  auto t0 = std::chrono::steady_clock::now();
  long steps = 0;
  for (int iter = cf.restart_iter; iter < cf.n_iter; iter++) {
    if (iter == 0 && output()) return 2;
    std::vector<double> t_start(nsims);
    for (int k = 0; k < nsims; k++) t_start[k] = cs[k]->t;
    const bool screen = ((iter + 1) % cf.screen_op_iter == 0);
    const auto t_it = std::chrono::steady_clock::now();
    double cfl = 0, norm = 0;
    int rc = 0;
    if (screen && verbose) rc = step_with_diagnostics(eng, cs, &cfl, &norm, err);
    else if ((rc = eng.step(1))) err = eng.err;
    if (rc) return rc;
    steps++;
    for (int k = 0; k < nsims; k++) stage_time[k] = t_start[k] + cs[k]->cfg.c[cs[k]->cfg.nstages - 1] * cs[k]->cfg.dt;
    if (screen && verbose) {  // TimePrintStep; the wall time is the host's view of an asynchronous step
      const double wc = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_it).count();
      const double t = cs[0]->t;
      if (nsims > 1) {
        printf("--\n");
        printf("iter=%7d,  t=%1.3e\n", iter + 1, t);
        printf("  CFL=%1.3E\n", cfl);
        printf("  norm=%1.4E\n", norm);
        printf("  wctime=%1.1E (s)\n", wc);
      } else {
        printf("iter=%7d  t=%1.3E  CFL=%1.3E  norm=%1.4E  wctime: %1.1E (s)  ", iter + 1, t, cfl, norm, wc);
      }
      if (cs[0]->h.cons) {  // the conservation error over every member's variables
        double e2 = 0;
        for (Case *c : cs) {
          std::vector<double> ce;
          if (c->conservation_errors(ce, eng.stream)) return fail_with(c, -1);
          for (double e : ce) e2 += e * e;
        }
        printf("  cons_err=%1.4E\n", sqrt(e2));
      } else if (nsims == 1) {
        printf("\n");
      }
      if (nsims > 1) printf("--\n\n");
    }
    if (((iter + 1) % cf.file_op_iter == 0) && ((iter + 1) < cf.n_iter) && output()) return 2;
  }
  if (eng.sync()) {
    err = eng.err;
    return -1;
  }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (seconds) *seconds = secs;
  // CalculateError of every member, then the final output (Solve.cpp:86-107),
  // then SimWriteErrors (main.cpp)
  std::vector<std::array<double, 3>> errs(nsims);
  for (int k = 0; k < nsims; k++)
    if (cs[k]->exact_errors(d, errs[k].data(), eng.stream)) return fail_with(cs[k], 1);
  if (output()) return 2;
  double main_secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_main).count();
  for (int k = 0; k < nsims; k++)
    if (cs[k]->write_run_files(d, errs[k].data(), steps, secs, main_secs, eng.stream)) return fail_with(cs[k], 2);
  for (int k = 0; k < nsims; k++) {
    Case *c = cs[k];
    if (verbose && errs[k][0] >= 0) {
      printf("Computed errors for domain %d:\n", k);
      printf("  L1         Error           : %1.16E\n", errs[k][0]);
      printf("  L2         Error           : %1.16E\n", errs[k][1]);
      printf("  Linfinity  Error           : %1.16E\n", errs[k][2]);
    }
    if (verbose && c->h.cons) {
      std::vector<double> ce;
      if (c->conservation_errors(ce, eng.stream)) return fail_with(c, -1);
      printf("Conservation Errors:\n");
      for (double e : ce) printf("\t%1.16E\n", e);
    }
  }
  return 0;
}

}  // namespace piafs

// ===========================================================================
// C ABI: single case = an engine with one member
// ===========================================================================
using namespace piafs;

static thread_local std::string g_create_error;  // the calling thread's last failed create

static int null_arg(piafs_case *pc, const char *what) {
  pc->err = std::string(what) + " pointer is NULL";
  return 1;
}

static int case_fail(piafs_case *pc, int rc) {
  pc->err = !pc->c->err.empty() ? pc->c->err : pc->eng.err;
  return rc;
}

extern "C" {

int piafs_create(const char *dir, int device, piafs_case **out) {
  if (!out) {
    g_create_error = "output handle pointer is NULL";
    return 1;
  }
  *out = nullptr;
  piafs_case *pc = new piafs_case();
  pc->eng.use_graph = true;  // steps replay a captured graph (eager launches while profiling)
  int rc = pc->eng.init(device);
  if (rc) {
    g_create_error = pc->eng.err;
    delete pc;
    return rc;
  }
  Case *c = new Case();
  rc = c->init(dir ? dir : "", device, pc->eng.stream);
  if (rc) {
    g_create_error = c->err;
    delete c;
    delete pc;
    return rc;
  }
  pc->c = c;
  pc->eng.add(c);
  rc = pc->eng.build();
  if (rc) {
    g_create_error = pc->eng.err;
    delete pc;
    return rc;
  }
  *out = pc;
  return 0;
}

int piafs_destroy(piafs_case *pc) {
  delete pc;
  return 0;
}

const char *piafs_last_error(const piafs_case *pc) { return pc ? pc->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

namespace piafs {
void fill_info(const Case *c, piafs_info *info) {
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
  info->scheme = c->cfg.crweno ? 1 : c->cfg.upw5 ? 2 : 0;
  info->characteristic = c->cfg.characteristic ? 1 : 0;
  info->include_chem = c->cfg.include_chem ? 1 : 0;
  info->device_bytes = c->bytes;
}
}  // namespace piafs

extern "C" {

int piafs_get_info(const piafs_case *pc, piafs_info *info) {
  if (!pc || !pc->c || !info) return 1;  // no case or no output
  fill_info(pc->c, info);
  return 0;
}

int piafs_upload_state(piafs_case *pc, const double *u) {
  if (!pc || !pc->c) return 1;  // no case
  if (!u) return null_arg(pc, "state");
  int rc = pc->c->upload(u, pc->eng.stream);
  return rc ? case_fail(pc, rc) : 0;
}

int piafs_download_state(piafs_case *pc, double *u) {
  if (!pc || !pc->c) return 1;  // no case
  if (!u) return null_arg(pc, "state");
  int rc = pc->c->download(u, pc->eng.stream);
  return rc ? case_fail(pc, rc) : 0;
}

int piafs_set_time(piafs_case *pc, double t) {
  if (!pc || !pc->c) return 1;  // no case
  pc->c->t = t;
  if (hipMemcpyAsync(pc->c->h.time, &t, sizeof(double), hipMemcpyHostToDevice, pc->eng.stream) != hipSuccess ||
      hipStreamSynchronize(pc->eng.stream) != hipSuccess) {
    pc->err = "set_time failed";
    return -1;
  }
  return 0;
}

int piafs_step(piafs_case *pc, int n) {
  if (!pc || !pc->c) return 1;  // no case
  int rc = pc->eng.step(n);
  return rc ? case_fail(pc, rc) : 0;
}

int piafs_sync(piafs_case *pc) {
  if (!pc || !pc->c) return 1;  // no case
  int rc = pc->eng.sync();
  return rc ? case_fail(pc, rc) : 0;
}

int piafs_step_diag(piafs_case *pc, int n, double *cfl, double *norm) {
  if (!pc || !pc->c) return 1;  // no case
  std::vector<Case *> one{pc->c};
  for (int it = 0; it < n; it++) {
    int rc = step_with_diagnostics(pc->eng, one, cfl ? cfl + it : nullptr, norm ? norm + it : nullptr, pc->err);
    if (rc) return rc;
  }
  return 0;
}

int piafs_rhs(piafs_case *pc, const double *u, double t, double *rhs) {
  if (!pc || !pc->c) return 1;  // no case
  if (!u || !rhs) return null_arg(pc, "state or output");
  Case *c = pc->c;
  hipStream_t st = pc->eng.stream;
  const size_t n = (size_t)c->cfg.nvars * c->h.G.plane;
  const double t_saved = c->t;
  int rc = 0;
  if (hipMemcpyAsync(c->uprev, c->h.u, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
  if ((rc = c->upload(u, st))) return case_fail(pc, rc);
  if ((rc = piafs_set_time(pc, t))) return rc;
  if ((rc = pc->eng.stage_only(0))) return case_fail(pc, rc);  // stage 0: U = u, time = t
  if ((rc = c->download(rhs, st, c->h.Udot[0]))) return case_fail(pc, rc);
  if (hipMemcpyAsync(c->h.u, c->uprev, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
  return piafs_set_time(pc, t_saved);
}

}  // extern "C"

// One evaluation of stage 0 at time t on the state u (interior, WriteBinary
// order), up to kernel class `last`, then `body` reads what it needs; the
// case's own solution and time are put back afterwards, as piafs_rhs does.
template <class F>
static int eval_pieces(piafs_case *pc, const double *u, double t, int last, int terms, F body) {
  Case *c = pc->c;
  hipStream_t st = pc->eng.stream;
  const size_t n = (size_t)c->cfg.nvars * c->h.G.plane;
  const double t_saved = c->t;
  int rc = 0;
  if (!u) {
    pc->err = "state pointer is NULL";
    return 1;
  }
  if (hipMemcpyAsync(c->uprev, c->h.u, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
  if ((rc = c->upload(u, st))) return case_fail(pc, rc);
  if ((rc = piafs_set_time(pc, t))) return rc;
  int *dterms = &pc->eng.d_cases[0].terms;
  if (terms != 7 && hipMemcpy(dterms, &terms, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return -1;
  // the stream is idle here (upload and set_time synchronise), so the
  // blocking copies of the mask cannot race the kernels that read it
  rc = pc->eng.stage_only(0, last);
  if (!rc) rc = body();
  const int all = 7;
  hipStreamSynchronize(st);
  if (terms != 7 && hipMemcpy(dterms, &all, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return -1;
  if (rc) return case_fail(pc, rc);
  if (hipMemcpyAsync(c->h.u, c->uprev, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
  return piafs_set_time(pc, t_saved);
}

static int copy_to_host(Case *c, double *dst, const double *src, size_t n, hipStream_t st) {
  const hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, st);
  const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(st) : e;
  if (e2 != hipSuccess) {
    c->err = std::string("device to host copy: ") + hipGetErrorString(e2);
    return -1;
  }
  return 0;
}

// planar device array (nvars planes of the ghosted grid) -> host
static int download_planes(Case *c, const double *src, std::vector<double> &buf, hipStream_t st) {
  buf.resize((size_t)c->cfg.nvars * c->h.G.plane);
  return copy_to_host(c, buf.data(), src, buf.size(), st);
}

extern "C" {

int piafs_boundary_fill(piafs_case *pc, const double *u, double t, double *ug) {
  if (!pc || !pc->c) return 1;  // no case
  if (!ug) return null_arg(pc, "output");
  Case *c = pc->c;
  return eval_pieces(pc, u, t, K_GHOST, 7, [&]() {
    std::vector<double> buf;
    int rc = download_planes(c, c->h.U, buf, pc->eng.stream);
    if (rc) return rc;
    // the device planes are padded (plane >= nxg*nyg); the caller's array holds the nxg*nyg points
    const Grid &G = c->h.G;
    const long pl = G.plane, np = (long)G.nxg * G.nyg;
    const int nv = c->cfg.nvars;
    for (long p = 0; p < np; p++)
      for (int v = 0; v < nv; v++) ug[(size_t)p * nv + v] = buf[(size_t)v * pl + p];
    return 0;
  });
}

int piafs_flux(piafs_case *pc, const double *u, double t, int dir, double *f) {
  if (!pc || !pc->c) return 1;  // no case
  if (!f) return null_arg(pc, "output");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) {
    pc->err = "dir must be 0 .. ndims-1";
    return 1;
  }
  return eval_pieces(pc, u, t, K_FLUX, 7, [&]() {
    std::vector<double> buf;
    int rc = download_planes(c, dir ? c->h.flux_y : c->h.flux_x, buf, pc->eng.stream);
    if (rc) return rc;
    const Grid &G = c->h.G;
    const SweepGeom &S = dir ? c->h.Sy : c->h.Sx;
    const int nv = c->cfg.nvars;
    std::fill(f, f + (size_t)G.nxg * G.nyg * nv, 0.0);  // the caller's nxg*nyg points (the planes are padded)
    for (int l = 0; l < S.nlines; l++)
      for (int pos = -S.g; pos < S.N + S.g; pos++) {
        const long b = S.base + (long)l * S.line_stride + pos;
        const int i = dir ? l : pos, j = dir ? pos : l;  // point (i, j) of the natural ghosted grid
        const long p = (long)(j + G.gy) * G.nxg + (i + G.g);
        for (int v = 0; v < nv; v++) f[(size_t)p * nv + v] = buf[(size_t)v * S.plane + b];
      }
    return 0;
  });
}

int piafs_interface_flux(piafs_case *pc, const double *u, double t, int dir, double *fi) {
  if (!pc || !pc->c) return 1;  // no case
  if (!fi) return null_arg(pc, "output");
  Case *c = pc->c;
  if (dir < 0 || dir >= c->cfg.ndims) {
    pc->err = "dir must be 0 .. ndims-1";
    return 1;
  }
  return eval_pieces(pc, u, t, K_UPWIND, 7, [&]() {
    const SweepGeom &S = dir ? c->h.Sy : c->h.Sx;
    const int nv = c->cfg.nvars, N = S.N, nl = S.nlines;
    std::vector<double> buf((size_t)nv * nl * (N + 1));
    const int rc = copy_to_host(c, buf.data(), dir ? c->h.fIy : c->h.fIx, buf.size(), pc->eng.stream);
    if (rc) return rc;
    for (int v = 0; v < nv; v++)
      for (int l = 0; l < nl; l++)
        for (int k = 0; k <= N; k++)
          fi[((size_t)l * (N + 1) + k) * nv + v] = buf[((size_t)v * nl + l) * (N + 1) + k];
    return 0;
  });
}

int piafs_rhs_terms(piafs_case *pc, const double *u, double t, double *hyp, double *par, double *src) {
  if (!pc || !pc->c) return 1;  // no case
  Case *c = pc->c;
  double *out[3] = {hyp, par, src};
  const size_t n = (size_t)c->cfg.npts_int * c->cfg.nvars;
  for (int k = 0; k < 3; k++) {
    if (!out[k]) continue;
    const int rc = eval_pieces(pc, u, t, K_NCLASS, 1 << k, [&]() {
      return c->download(out[k], pc->eng.stream, c->h.Udot[0]);
    });
    if (rc) return rc;
    if (k == 0)  // k_assemble wrote -hyp
      for (size_t q = 0; q < n; q++) hyp[q] = -hyp[q];
  }
  return 0;
}

int piafs_write_binary(piafs_case *pc, const char *path) {
  if (!pc || !pc->c) return 1;  // no case
  if (!path) return null_arg(pc, "path");
  int rc = pc->c->write_binary(path, pc->eng.stream);
  return rc ? case_fail(pc, rc) : 0;
}

int piafs_set_derivative_exchange(piafs_case *pc, int on) {
  if (!pc || !pc->c) return 1;  // no case
  Case *c = pc->c;
  for (int d = 0; d < 2; d++) {
    bool periodic = false;
    for (const Boundary &b : c->cfg.bcs)
      periodic = periodic || (b.bctype == BC_PERIODIC && b.dim == d && d < c->cfg.ndims);
    // only the viscous term of navierstokes2d exchanges derivative ghosts
    c->h.P.xchg[d] = (on && periodic && c->cfg.model_id == MODEL_NS2D) ? 1 : 0;
  }
  int rc = pc->eng.build();  // the device descriptor table and a fresh step graph
  return rc ? case_fail(pc, rc) : 0;
}

int piafs_conservation_errors(piafs_case *pc, double *errors) {
  if (!pc || !pc->c) return 1;  // no case
  if (!errors) return null_arg(pc, "output");
  if (!pc->c->h.cons) {
    pc->err = "the deck does not set conservation_check yes";
    return 1;
  }
  std::vector<double> ce;
  int rc = pc->eng.sync();
  if (!rc) rc = pc->c->conservation_errors(ce, pc->eng.stream);
  if (rc) return case_fail(pc, rc);
  std::copy(ce.begin(), ce.end(), errors);
  return 0;
}

int piafs_kernel_times(piafs_case *pc, int enable, const char **names, double *ms, int *n) {
  if (!pc || !pc->c) return 1;  // no case
  Engine &e = pc->eng;
  if (enable >= 0) {
    e.profiling = (enable != 0);
    e.ev_used = 0;
  }
  if (!n) return 0;
  int cap = *n;
  e.sync();
  double tot[K_NCLASS] = {0};
  int cnt[K_NCLASS] = {0};
  for (size_t k = 0; k < e.ev_used; k++) {
    float m = 0;
    hipEventElapsedTime(&m, e.ev[k].first, e.ev[k].second);
    tot[e.ev_cls[k]] += m;
    cnt[e.ev_cls[k]]++;
  }
  int w = 0;
  for (int k = 0; k < K_NCLASS && w < cap; k++) {
    if (!cnt[k]) continue;
    if (names) names[w] = kclass_name(k);
    if (ms) ms[w] = tot[k] / cnt[k];
    w++;
  }
  *n = w;
  e.ev_used = 0;
  return 0;
}

int piafs_autotune(piafs_case *pc, int *rows_per_lane, double *recon_ms) {
  if (!pc || !pc->c) return 1;  // no case
  static const int cands[] = {4, 8, 12, 16};
  int rc = pc->eng.autotune(cands, 4, 3, rows_per_lane, recon_ms);
  if (rc) pc->err = pc->eng.err;
  return rc;
}

int piafs_estimate_bytes(const char *dir, size_t *bytes) {
  if (!bytes) {
    g_create_error = "output pointer is NULL";
    return 1;
  }
  const std::string d = dir ? dir : "";
  std::string err;
  int nsims = 1;
  const int ens = read_ensemble(d, nsims, err);  // 0: ensemble, 1: single simulation, 2: bad simulation.inp
  if (ens == 2) {
    g_create_error = err;
    return 1;
  }
  if (ens == 1) nsims = 1;
  size_t total = 0;
  for (int k = 0; k < nsims; k++) {  // an ensemble's members, each parsed as piafs_ensemble_create does
    Config cfg;
    if (ens == 0) {
      cfg.member = k;
      cfg.nsims = nsims;
    }
    if (cfg.parse(d, err)) {
      g_create_error = ens == 0 ? "domain " + std::to_string(k) + ": " + err : err;
      return 1;
    }
    total += Case::plan_bytes(cfg);
  }
  *bytes = total;
  return 0;
}

int piafs_solve(piafs_case *pc, const char *dir, int verbose, double *seconds) {
  if (!pc || !pc->c) return 1;  // no case
  std::vector<Case *> one{pc->c};
  int rc = run_solve(pc->eng, one, dir ? dir : "", verbose, seconds, pc->err);
  return rc;
}

}  // extern "C"
