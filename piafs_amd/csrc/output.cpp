// Files a PIAFS run writes, in the reference's names and layouts:
//  solution op[_NNNNN].{bin,dat}: OutputSolution.cpp:22-70, WriteArray.c:131-142
//    (WriteArraySerial), WriteBinary.c:37-93, WriteText.c:21-55, WriteTecplot2D.c:22-80;
//  species op_species_z<k>[_NNNNN]: ChemistryWriteSpecies.c:18-80 with the photon
//    density of the last right-hand-side evaluation (ChemistrySolve.c:18-65);
//  errors.dat / conservation.dat / function_counts.dat: SimulationWriteErrors.c:21-100,
//    with the exact-solution errors of CalculateError.c:28-121.
// The state is computed on the device; these writers format what it downloads.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "case.hpp"

namespace piafs {

static bool fmt_binary(const std::string &f) { return f == "binary" || f == "bin"; }

std::string output_extension(const Config &c) {
  if (c.op_file_format == "text" || c.op_file_format == "tecplot2d") return ".dat";
  if (fmt_binary(c.op_file_format)) return ".bin";
  return "";
}

// One array in the configured format.  x holds the interior coordinates of
// each dimension in turn, u the interior values point-major (x fastest).
static int write_array(const Config &c, const std::string &path, int nvars, const double *x, const double *u,
                       std::string &err) {
  const int nd = c.ndims;
  long npts = 1;
  for (int d = 0; d < nd; d++) npts *= c.dim[d];
  if (fmt_binary(c.op_file_format)) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { err = "could not open " + path + " for writing"; return 2; }
    long xs = 0;
    for (int d = 0; d < nd; d++) xs += c.dim[d];
    bool ok = fwrite(&nd, sizeof(int), 1, f) == 1 && fwrite(&nvars, sizeof(int), 1, f) == 1 &&
              fwrite(c.dim, sizeof(int), nd, f) == (size_t)nd && fwrite(x, sizeof(double), xs, f) == (size_t)xs &&
              fwrite(u, sizeof(double), (size_t)npts * nvars, f) == (size_t)npts * nvars;
    fclose(f);
    if (!ok) { err = "short write to " + path; return 2; }
    return 0;
  }
  const bool tec = (c.op_file_format == "tecplot2d");
  if (tec && nd != 2) {
    err = "WriteTecplot2D() is for 2-dimensional problems only";
    return 2;
  }
  FILE *f = fopen(path.c_str(), "w");
  if (!f) { err = "could not open " + path + " for writing"; return 2; }
  if (tec) {
    fprintf(f, "VARIABLES=\"I\",\"J\",\"X\",\"Y\",");
    char vn[3] = "00";
    for (int v = 0; v < nvars; v++) {
      fprintf(f, "\"%s\",", vn);
      if (vn[1] == '9') { vn[0]++; vn[1] = '0'; }
      else vn[1]++;
    }
    fprintf(f, "\n");
    fprintf(f, "ZONE I=%d,J=%d,F=POINT\n", c.dim[0], c.dim[1]);
  }
  for (long q = 0; q < npts; q++) {
    int idx[3] = {0, 0, 0};
    long r = q;
    for (int d = 0; d < nd; d++) { idx[d] = (int)(r % c.dim[d]); r /= c.dim[d]; }
    for (int d = 0; d < nd; d++) fprintf(f, "%4d ", idx[d]);
    long off = 0;
    for (int d = 0; d < nd; d++) {
      fprintf(f, "%+1.16E ", x[off + idx[d]]);
      off += c.dim[d];
    }
    for (int v = 0; v < nvars; v++) fprintf(f, "%+1.16E ", u[q * nvars + v]);
    fprintf(f, "\n");
  }
  fclose(f);
  return 0;
}

static std::vector<double> interior_x(const Config &c) {
  std::vector<double> xi;
  long off = 0;
  for (int d = 0; d < c.ndims; d++) {
    for (int i = 0; i < c.dim[d]; i++) xi.push_back(c.x[off + c.ghosts + i]);
    off += c.dim[d] + 2 * c.ghosts;
  }
  return xi;
}

int Case::write_solution(const std::string &path, hipStream_t st) {
  std::vector<double> uh((size_t)cfg.npts_int * cfg.nvars);
  int rc = download(uh.data(), st);
  if (rc) return rc;
  std::vector<double> xi = interior_x(cfg);
  std::string e;
  rc = write_array(cfg, path, cfg.nvars, xi.data(), uh.data(), e);
  if (rc) return fail(rc, e);
  return 0;
}

// op_species_z<k>: the seven species of the z-layer z_i from the solution
// and, last, the photon density of layer k as the latest right-hand-side
// evaluation left it (zero before the first one).  stage_time < 0: none yet.
int Case::write_species(const std::string &dir, const std::string &suffix, double stage_time, hipStream_t st) {
  const Chem &ch = cfg.chem;
  const int nv = cfg.nvars, nz = ch.z_i + 1, nsp = 8;
  const long np = cfg.npts_int;
  std::vector<double> uh((size_t)np * nv), Us;
  int rc = download(uh.data(), st);
  if (rc) return rc;
  if (stage_time >= 0.0 && nz > 1) {
    Us.resize((size_t)np * nv);
    if ((rc = download(Us.data(), st, h.U))) return rc;
  }
  // ChemistrySetPhotonDensity (ChemistrySolve.c:33-60) at the interior points
  std::vector<double> hnu((size_t)np * nz, 0.0);
  if (stage_time >= 0.0) {
    const Grid G = make_grid(cfg);
    for (long q = 0; q < np; q++) {
      const int i = (int)(q % G.nx), j = (int)(q / G.nx);
      const long p = (long)(j + G.gy) * G.nxg + (i + G.g);
      const double sigma = ch.t_pulse_norm / 2.35482;
      double I0 = 0.0;
      if (stage_time > ch.t_start_norm) {
        const double tp = stage_time - (ch.t_start_norm + ch.t_pulse_norm);
        I0 = ch.I0 * exp(-(tp * tp) / (2 * sigma * sigma)) * cfg.imap[p];
      }
      hnu[q * nz] = I0 / (ch.c_light * ch.h_planck * ch.nu * ch.n_O2);
      for (int iz = 1; iz < nz; iz++) {
        const double sg = ch.sO3 * ch.n_O2;
        const double damp = 1.0 - ch.dz * sg * Us[(size_t)q * nv + ch.nflow + ch.z_stride * (iz - 1) + 1];
        hnu[q * nz + iz] = hnu[q * nz + iz - 1] * damp;
      }
    }
  }
  const int width = (int)log10((double)nz) + 1;
  std::vector<double> arr((size_t)np * nsp);
  std::vector<double> xi = interior_x(cfg);
  for (int iz = cfg.species_all_zlocs ? 0 : ch.z_i; iz < nz; iz++) {
    for (long q = 0; q < np; q++) {
      for (int m = 0; m < ch.z_stride; m++) arr[q * nsp + m] = uh[(size_t)q * nv + ch.nflow + ch.z_stride * ch.z_i + m];
      arr[q * nsp + nsp - 1] = hnu[q * nz + iz];
    }
    char idx[32];
    snprintf(idx, sizeof idx, "%0*d", width, iz);
    // an ensemble member's root gains "_<member>_" (ChemistryWriteSpecies.c:44-50)
    const std::string mem = cfg.nsims > 1 ? cfg.tag() + "_" : std::string();
    std::string path = (dir.empty() ? std::string() : dir + "/") + "op_species_z" + idx + mem + suffix +
                       output_extension(cfg);
    std::string e;
    if ((rc = write_array(cfg, path, nsp, xi.data(), arr.data(), e))) return fail(rc, e);
  }
  return 0;
}

// CalculateError (CalculateError.c:28-121): L1, L2, Linf of u - u_exact over
// the interior (sums point by point, variables innermost), relative to the
// norms of u_exact when all three exceed 1e-15.  err = -1 without exact.inp.
int Case::exact_errors(const std::string &dir, double err3[3], hipStream_t st) {
  err3[0] = err3[1] = err3[2] = -1.0;
  std::vector<double> uex;
  bool found = false;
  std::string e;
  if (read_exact(cfg, dir, uex, found, e)) return fail(1, e);
  if (!found) return 0;
  for (int d = 0; d < cfg.ndims; d++)
    if (cfg.dim_exact[d] != cfg.dim[d])
      return fail(1, "an exact solution of another size (size_exact) is not supported on this path");
  const long n = (long)cfg.npts_int * cfg.nvars;
  std::vector<double> uh(n);
  int rc = download(uh.data(), st);
  if (rc) return rc;
  const double npg = (double)cfg.npts_int;
  auto norms = [&](const std::vector<double> &a, double out[3]) {
    double s1 = 0, s2 = 0, mx = 0;
    for (long k = 0; k < n; k++) s1 += fabs(a[k]);          // ArraySumAbsnD
    for (long k = 0; k < n; k++) s2 += a[k] * a[k];         // ArraySumSquarenD
    for (long k = 0; k < n; k++) if (fabs(a[k]) > mx) mx = fabs(a[k]);  // ArrayMaxnD
    out[0] = s1 / npg;
    out[1] = sqrt(s2 / npg);
    out[2] = mx;
  };
  double sn[3];
  norms(uex, sn);
  for (long k = 0; k < n; k++) uex[k] += -1.0 * uh[k];  // _ArrayAXPY_(u, -1.0, uex)
  norms(uex, err3);
  const double tol = 1e-15;
  if (sn[0] > tol && sn[1] > tol && sn[2] > tol) {
    err3[0] /= sn[0];
    err3[1] /= sn[1];
    err3[2] /= sn[2];
  }
  return 0;
}

// VolumeIntegral (VolumeIntegral.c:18-50): sum over the interior points (x
// fastest) of u dV, dV the product of the local 1/dxinv.
static void volume_integral(const Config &c, const double *u, std::vector<double> &vi) {
  const int nv = c.nvars, nd = c.ndims;
  vi.assign(nv, 0.0);
  for (long q = 0; q < c.npts_int; q++) {
    int idx[3] = {0, 0, 0};
    long r = q, off = 0;
    for (int d = 0; d < nd; d++) { idx[d] = (int)(r % c.dim[d]); r /= c.dim[d]; }
    double dV = 1.0;
    for (int d = 0; d < nd; d++) {
      dV *= (1.0 / c.dxinv[off + c.ghosts + idx[d]]);
      off += c.dim[d] + 2 * c.ghosts;
    }
    for (int v = 0; v < nv; v++) vi[v] += (u[q * nv + v] * dV);
  }
}

int Case::conservation_errors(std::vector<double> &ce, hipStream_t st) {
  const int nv = cfg.nvars;
  ce.assign(nv, -1.0);  // Initialize.c:195-196
  if (!h.cons || t == (double)cfg.restart_iter * cfg.dt) return 0;  // computed after each step only
  std::vector<double> uh((size_t)cfg.npts_int * nv), tbi(nv), vi0, vi;
  int rc = download(uh.data(), st);
  if (rc) return rc;
  if (hipMemcpyAsync(tbi.data(), h.tbi, nv * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(-1, "copy of the boundary integral failed");
  volume_integral(cfg, cfg.u0.data(), vi0);
  volume_integral(cfg, uh.data(), vi);
  for (int v = 0; v < nv; v++) {
    const double base = fabs(vi0[v]) > 1.0 ? fabs(vi0[v]) : 1.0;
    const double e = (vi[v] + tbi[v] - vi0[v]) * (vi[v] + tbi[v] - vi0[v]);
    ce[v] = sqrt(e) / base;
  }
  return 0;
}

// SimWriteErrors (SimulationWriteErrors.c:21-100) for one process.
int Case::write_run_files(const std::string &dir, const double err3[3], long steps, double solver_rt,
                          double main_rt, hipStream_t st) {
  const std::string pre = dir.empty() ? std::string() : dir + "/";
  const std::string t = cfg.tag();  // errors[_<member>].dat etc. (SimulationWriteErrors.c:39-61)
  FILE *f = fopen((pre + "errors" + t + ".dat").c_str(), "w");
  if (!f) return fail(2, "could not open errors" + t + ".dat for writing");
  for (int d = 0; d < cfg.ndims; d++) fprintf(f, "%4d ", cfg.dim[d]);
  for (int d = 0; d < cfg.ndims; d++) fprintf(f, "%4d ", 1);
  fprintf(f, "%1.16E  ", cfg.dt);
  fprintf(f, "%1.16E %1.16E %1.16E   ", err3[0], err3[1], err3[2]);
  fprintf(f, "%1.16E %1.16E\n", solver_rt, main_rt);
  fclose(f);
  std::vector<double> ce;
  int rc = conservation_errors(ce, st);
  if (rc) return rc;
  f = fopen((pre + "conservation" + t + ".dat").c_str(), "w");
  if (!f) return fail(2, "could not open conservation" + t + ".dat for writing");
  for (int d = 0; d < cfg.ndims; d++) fprintf(f, "%4d ", cfg.dim[d]);
  for (int d = 0; d < cfg.ndims; d++) fprintf(f, "%4d ", 1);
  fprintf(f, "%1.16E  ", cfg.dt);
  for (int v = 0; v < cfg.nvars; v++) fprintf(f, "%1.16E ", ce[v]);
  fprintf(f, "\n");
  fclose(f);
  // every right-hand-side evaluation calls the hyperbolic and source functions;
  // the parabolic one counts only when viscous (NavierStokes2DParabolicFunction.c:64-65)
  const long evals = steps * cfg.nstages;
  const bool visc = (cfg.model_id == MODEL_NS2D && cfg.Re > 0);
  f = fopen((pre + "function_counts" + t + ".dat").c_str(), "w");
  if (!f) return fail(2, "could not open function_counts" + t + ".dat for writing");
  fprintf(f, "%d\n%ld\n%ld\n%ld\n", cfg.n_iter, evals, visc ? evals : 0L, evals);
  fclose(f);
  return 0;
}

}  // namespace piafs
