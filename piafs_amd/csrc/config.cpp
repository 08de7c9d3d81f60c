// Input decks -> Config.  Keywords, defaults and selection strings follow the
// reference readers named at each function; unknown keywords skip one token
// as the reference does.
#include "config.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace piafs {

namespace {

std::string join(const std::string &dir, const char *name) {
  if (dir.empty()) return name;
  return dir + "/" + name;
}

// Tokenised "begin ... end" file.  Returns 0 if absent (defaults apply),
// 1 on a format error, 2 when present and read.
template <class F>
int read_keywords(const std::string &path, F &&on_key, std::string &err) {
  std::ifstream in(path);
  if (!in) return 0;
  std::string w;
  if (!(in >> w) || w != "begin") {
    err = "illegal format in " + path;
    return 1;
  }
  while (in >> w && w != "end") {
    if (!on_key(w, in)) {
      std::string skip;
      if (!(in >> skip)) {
        err = "truncated " + path;
        return 1;
      }
    }
    if (in.fail()) {
      err = "bad value for keyword " + w + " in " + path;
      return 1;
    }
  }
  return 2;
}

// FindInterval.c:22-45
void find_interval(double a, double b, const double *x, int N, int *imin, int *imax) {
  *imax = -1;
  *imin = N;
  double min_dx = x[1] - x[0];
  for (int i = 2; i < N; i++) min_dx = std::min(min_dx, x[i] - x[i - 1]);
  double tol = 1e-10 * min_dx;
  for (int i = 0; i < N; i++)
    if (x[i] <= b + tol) *imax = i + 1;
  for (int i = N - 1; i > -1; i--)
    if (x[i] >= a - tol) *imin = i;
}

}  // namespace

// ReadArraySerial, ReadArray.c:153-300 (+ ghost x extrapolation :71-91).
static int read_array(const Config &c, const std::string &dir, const char *root, int nv,
                      std::vector<double> *x, std::vector<double> &u_pointmajor, bool &found,
                      std::string &err) {
  found = false;
  bool bin = (c.ip_file_type == "bin" || c.ip_file_type == "binary");
  std::string fn = join(dir, (std::string(root) + ".inp").c_str());
  FILE *f = fopen(fn.c_str(), bin ? "rb" : "r");
  if (!f) return 0;
  found = true;
  long size = 1, xs = 0;
  for (int d = 0; d < c.ndims; d++) {
    size *= c.dim[d];
    xs += c.dim[d];
  }
  std::vector<double> xg(xs);
  u_pointmajor.assign(size * nv, 0.0);
  bool ok = true;
  if (bin) {
    ok = fread(xg.data(), sizeof(double), xs, f) == (size_t)xs &&
         fread(u_pointmajor.data(), sizeof(double), size * nv, f) == (size_t)(size * nv);
  } else {
    for (long i = 0; i < xs && ok; i++) ok = fscanf(f, "%lf", &xg[i]) == 1;
    for (int v = 0; v < nv && ok; v++)
      for (long i = 0; i < size && ok; i++) ok = fscanf(f, "%lf", &u_pointmajor[i * nv + v]) == 1;
  }
  fclose(f);
  if (!ok) {
    err = "unable to read " + fn;
    return 1;
  }
  if (x) {
    int g = c.ghosts;
    long off = 0, og = 0;
    for (int d = 0; d < c.ndims; d++) {
      int N = c.dim[d];
      double *X = x->data() + off;
      for (int i = 0; i < N; i++) X[i + g] = xg[og + i];
      for (int i = 0; i < g; i++) X[i] = X[g] + ((double)(g - i)) * (X[g] - X[g + 1]);
      for (int i = N + g; i < N + 2 * g; i++)
        X[i] = X[N + g - 1] + ((double)(i - (N + g - 1))) * (X[N + g - 1] - X[N + g - 2]);
      off += N + 2 * g;
      og += N;
    }
  }
  return 0;
}

int read_exact(const Config &c, const std::string &dir, std::vector<double> &u, bool &found, std::string &err) {
  // exact[_<member>] (CalculateError.c:45-51)
  return read_array(c, dir, ("exact" + c.tag()).c_str(), c.nvars, nullptr, u, found, err);
}

std::string Config::tag() const {
  if (nsims <= 1) return std::string();
  char b[32];
  snprintf(b, sizeof b, "_%0*d", (int)log10((double)nsims) + 1, member);
  return b;
}

int read_ensemble(const std::string &dir, int &nsims, std::string &err) {
  nsims = -1;
  const std::string fn = join(dir, "simulation.inp");
  {
    std::ifstream probe(fn);
    if (!probe) return 1;
  }
  auto on_key = [&](const std::string &w, std::istream &s) -> bool {
    if (w == "nsims") {
      s >> nsims;
      return true;
    }
    return false;
  };
  if (read_keywords(fn, on_key, err) != 2) {
    if (err.empty()) err = "illegal format in simulation.inp";
    return 2;
  }
  if (nsims < 1) {
    err = "invalid value for nsims (" + std::to_string(nsims) + ") in simulation.inp";
    return 2;
  }
  return 0;
}

int Config::parse(const std::string &dir, std::string &err) {
  // ---- solver.inp: ReadInputs.c:200-480 -----------------------------------
  {
    std::string fn = join(dir, "solver.inp");
    std::ifstream in(fn);
    if (!in) {
      err = "file solver.inp not found";
      return 1;
    }
    auto on_key = [&](const std::string &w, std::istream &s) -> bool {
      if (w == "ndims") s >> ndims;
      else if (w == "nvars") s >> nvars;
      else if (w == "size" || w == "size_exact") {  // one tuple per member (ReadInputs.c:272-305)
        int *to = (w == "size") ? dim : dim_exact;
        for (int n = 0; n < nsims; n++)
          for (int d = 0; d < ndims; d++) {
            int v = 0;
            s >> v;
            if (n == member && d < 3) to[d] = v;
          }
      }
      else if (w == "ghost") s >> ghosts;
      else if (w == "n_iter") s >> n_iter;
      else if (w == "restart_iter") s >> restart_iter;
      else if (w == "time_scheme") s >> time_scheme;
      else if (w == "time_scheme_type") s >> time_scheme_type;
      else if (w == "hyp_space_scheme") s >> hyp_scheme;
      else if (w == "hyp_interp_type") s >> interp_type;
      else if (w == "par_space_scheme") s >> par_scheme;
      else if (w == "dt") s >> dt;
      else if (w == "conservation_check") s >> conservation_check;
      else if (w == "screen_op_iter") s >> screen_op_iter;
      else if (w == "file_op_iter") s >> file_op_iter;
      else if (w == "op_file_format") s >> op_file_format;
      else if (w == "ip_file_type") s >> ip_file_type;
      else if (w == "input_mode" || w == "output_mode") {
        std::string &m = (w == "input_mode") ? input_mode : output_mode;
        s >> m;
        if (m != "serial") { int nio; s >> nio; }
      } else if (w == "op_overwrite") s >> op_overwrite;
      else if (w == "model") s >> model;
      else return false;
      return true;
    };
    int r = read_keywords(fn, on_key, err);
    if (r != 2) {
      if (err.empty()) err = "illegal format in solver.inp";
      return 1;
    }
    for (int d = 0; d < 3; d++)
      if (dim_exact[d] <= 0) dim_exact[d] = dim[d];  // ReadInputs.c:283-288: defaults to size
    if (op_file_format != "text" && op_file_format != "tecplot2d" && op_file_format != "binary" &&
        op_file_format != "bin" && op_file_format != "none") {  // InitializeSolvers.c:224-239
      err = op_file_format + " is not a supported file format";
      return 1;
    }
    if (screen_op_iter <= 0) screen_op_iter = 1;
    if (file_op_iter <= 0) file_op_iter = n_iter;
    if (restart_iter != 0 && op_file_format != "binary") {
      err = "restart is supported only for binary output files";
      return 1;
    }
  }
  if (ndims < 1 || ndims > 2) {
    err = "ndims must be 1 or 2 on this path";
    return 1;
  }
  if (ghosts < 3) {
    err = "fifth-order schemes need ghost >= 3";
    return 1;
  }
  npts_int = 1;
  npts_g = 1;
  for (int d = 0; d < ndims; d++) {
    // the reference needs 4 points per dimension for its one rank
    // (ComputeProcessorDistribution, ReadInputs.c:39-49, _MIN_GRID_PTS_PER_PROC_
    // of basic.h:42); a periodic or mirrored ghost layer reads as many
    // interior points as there are ghosts
    if (dim[d] < 4 || dim[d] < ghosts) {
      err = "grid too small: " + std::to_string(dim[d]) + " points along dimension " + std::to_string(d) +
            " (at least 4, and at least the " + std::to_string(ghosts) + " ghost points)";
      return 1;
    }
    npts_int *= dim[d];
    npts_g *= dim[d] + 2 * ghosts;
  }
  long xs = 0;
  for (int d = 0; d < ndims; d++) xs += dim[d] + 2 * ghosts;
  x.assign(xs, 0.0);
  dxinv.assign(xs, 0.0);

  // ---- initial solution: InitialSolution.c:24-114 -------------------------
  {
    bool found = false;
    if (read_array(*this, dir, ("initial" + tag()).c_str(), nvars, &x, u0, found, err)) return 1;
    if (!found) {
      err = "initial solution file not found";
      return 1;
    }
    int g = ghosts;
    long off = 0;
    for (int d = 0; d < ndims; d++) {
      for (int i = 0; i < dim[d]; i++)
        dxinv[i + off + g] = 2.0 / (x[i + 1 + off + g] - x[i - 1 + off + g]);
      for (int i = 0; i < g; i++) dxinv[off + i] = dxinv[off + g];
      for (int i = dim[d] + g; i < dim[d] + 2 * g; i++) dxinv[off + i] = dxinv[off + dim[d] + g - 1];
      off += dim[d] + 2 * g;
    }
  }

  // ---- boundary.inp: InitializeBoundaries.c:35-290 ------------------------
  {
    // boundary[_<member>].inp, else boundary.inp (InitializeBoundaries.c:51-75)
    const std::string own = "boundary" + tag() + ".inp";
    std::ifstream in(join(dir, own.c_str()));
    if (!in && nsims > 1) in.open(join(dir, "boundary.inp"));
    if (!in) {
      err = "boundary condition file " + own + (nsims > 1 ? " or boundary.inp" : "") + " not found";
      return 1;
    }
    int nb = 0;
    if (!(in >> nb) || nb < 0) {
      err = "bad boundary.inp";
      return 1;
    }
    for (int n = 0; n < nb; n++) {
      Boundary b;
      if (!(in >> b.type >> b.dim >> b.face)) {
        err = "bad boundary.inp";
        return 1;
      }
      for (int d = 0; d < ndims; d++)
        if (!(in >> b.xmin[d] >> b.xmax[d])) {
          err = "bad boundary.inp";
          return 1;
        }
      static const char *names[BC_NTYPES] = {"periodic", "extrapolate", "dirichlet", "reflect", "sponge",
                                             "noslip-wall", "slip-wall", "subsonic-inflow", "subsonic-outflow",
                                             "subsonic-ambivalent", "supersonic-inflow", "supersonic-outflow"};
      b.bctype = -1;
      for (int k = 0; k < BC_NTYPES; k++)
        if (b.type == names[k]) b.bctype = k;
      if (b.bctype < 0) {
        err = "boundary type " + b.type + " is not supported on this path";
        return 1;
      }
      if (b.dim < 0 || b.dim >= ndims || (b.face != 1 && b.face != -1)) {
        err = "invalid boundary dim/face";
        return 1;
      }
      // type-specific values, read as the reference reads them (a missing
      // value leaves zero): InitializeBoundaries.c:95-150
      auto rd = [&](double &v) { if (!(in >> v)) { in.clear(); v = 0.0; } };
      const int nsc = nvars - (2 + ndims);
      switch (b.bctype) {
        case BC_DIRICHLET:
        case BC_SPONGE:
          b.vals.assign(nvars, 0.0);
          for (double &v : b.vals) rd(v);
          break;
        case BC_SLIP_WALL:
        case BC_NOSLIP_WALL:
          for (int d = 0; d < ndims; d++) rd(b.vel[d]);
          break;
        case BC_SUBSONIC_INFLOW:
          rd(b.rho);
          for (int d = 0; d < ndims; d++) rd(b.vel[d]);
          b.vals.assign(nsc > 0 ? nsc : 0, 0.0);
          for (double &v : b.vals) rd(v);
          break;
        case BC_SUBSONIC_OUTFLOW: rd(b.p); break;
        case BC_SUBSONIC_AMBIVALENT:
          rd(b.rho);
          for (int d = 0; d < ndims; d++) rd(b.vel[d]);
          rd(b.p);
          break;
        case BC_SUPERSONIC_INFLOW:
          rd(b.rho);
          for (int d = 0; d < ndims; d++) rd(b.vel[d]);
          rd(b.p);
          b.vals.assign(nsc > 0 ? nsc : 0, 0.0);
          for (double &v : b.vals) rd(v);
          break;
        default: break;
      }
      // CalculateLocalExtent, InitializeBoundaries.c:296-377 (one process)
      long off = 0;
      for (int d = 0; d < ndims; d++) {
        if (b.bctype == BC_SPONGE) {  // a sponge zone is a box of interior points
          find_interval(b.xmin[d], b.xmax[d], &x[off + ghosts], dim[d], &b.is[d], &b.ie[d]);
          if (b.ie[d] - b.is[d] <= 0) b.on = false;
        } else if (d == b.dim) {
          if (b.face == 1) { b.is[d] = -ghosts; b.ie[d] = 0; }
          else { b.is[d] = dim[d]; b.ie[d] = dim[d] + ghosts; }
        } else {
          find_interval(b.xmin[d], b.xmax[d], &x[off + ghosts], dim[d], &b.is[d], &b.ie[d]);
          if (b.ie[d] - b.is[d] <= 0) b.on = false;
        }
        off += dim[d] + 2 * ghosts;
      }
      bcs.push_back(b);
    }
  }

  // ---- solvers: InitializeSolvers.c:57-212 --------------------------------
  if (par_scheme == "4") par4 = true;
  else if (par_scheme == "2") par4 = false;
  else {
    err = par_scheme + " is not a supported spatial scheme for the parabolic terms";
    return 1;
  }
  if (interp_type != "characteristic" && interp_type != "components") {
    err = interp_type + " is not a supported interpolation type";
    return 1;
  }
  if (hyp_scheme == "weno5") crweno = false;
  else if (hyp_scheme == "crweno5") crweno = true;
  else if (hyp_scheme == "upw5") upw5 = true;
  else {
    err = hyp_scheme + " is not a supported spatial interpolation scheme on this path";
    return 1;
  }
  characteristic = (nvars > 1 && interp_type == "characteristic");
  {
    double p = 2.0, rc = 0.3, xi = 0.001, tol = 1e-16;
    auto on_key = [&](const std::string &w, std::istream &s) -> bool {
      if (w == "mapped") s >> mapped;
      else if (w == "borges") s >> borges;
      else if (w == "yc") s >> yc;
      else if (w == "no_limiting") s >> no_limiting;
      else if (w == "epsilon") s >> eps;
      else if (w == "p") s >> p;
      else if (w == "rc") s >> rc;
      else if (w == "xi") s >> xi;
      else if (w == "tol") s >> tol;
      else return false;
      return true;
    };
    // upw5 has no weights and the reference reads no weno.inp for it
    // (InitializeSolvers.c:141-148 calls no WENOInitialize)
    if (!upw5 && read_keywords(join(dir, "weno.inp"), on_key, err) == 1) return 1;
    weight_kind = upw5 ? W_UPW5 : yc ? W_YC : borges ? W_Z : mapped ? W_MAPPED : W_JS;
  }
  // RK tableau
  {
    for (double &a : A) a = 0;
    for (int s = 0; s < 4; s++) b[s] = c[s] = 0;
    if (time_scheme == "euler") {
      nstages = 1;
      b[0] = 1.0;
    } else if (time_scheme == "rk") {
      const std::string &t = time_scheme_type;
      if (t == "1fe") { nstages = 1; b[0] = 1.0; }
      else if (t == "22") { nstages = 2; A[2] = 1.0; c[1] = 1.0; b[0] = b[1] = 0.5; }
      else if (t == "33") {
        nstages = 3;
        A[3] = 2.0 / 3.0; A[6] = 2.0 / 3.0 - 1.0 / 4.0; A[7] = 1.0 / 4.0;
        c[1] = 2.0 / 3.0; c[2] = 2.0 / 3.0;
        b[0] = 1.0 / 4.0; b[1] = -1.0 / 4.0; b[2] = 1.0;
      } else if (t == "44") {
        nstages = 4;
        A[4] = 0.5; A[9] = 0.5; A[14] = 1.0;
        c[1] = c[2] = 0.5; c[3] = 1.0;
        b[0] = 1.0 / 6.0; b[1] = 1.0 / 3.0; b[2] = 1.0 / 3.0; b[3] = 1.0 / 6.0;
      } else if (t == "ssprk3" || t == "tvdrk3") {
        nstages = 3;
        A[3] = 1.0; A[6] = 0.25; A[7] = 0.25;
        c[1] = 1.0; c[2] = 0.5;
        b[0] = b[1] = 1.0 / 6.0; b[2] = 2.0 / 3.0;
      } else {
        err = t + " is not a supported multi-stage time integration scheme of class rk";
        return 1;
      }
    } else {
      err = time_scheme + " is not a supported time-integration scheme";
      return 1;
    }
  }

  // ---- physics.inp: Euler1DInitialize.c:58-170 / NavierStokes2DInitialize.c:68-235
  std::string include_chem_s = "no", write_op = "yes";
  if (model == "euler1d") {
    model_id = MODEL_EULER1D;
    nflow = 3;
    if (ndims != 1) {
      err = "euler1d: ndims has to be 1";
      return 1;
    }
  } else if (model == "navierstokes2d") {
    model_id = MODEL_NS2D;
    nflow = 4;
    if (ndims != 2) {
      err = "navierstokes2d: ndims has to be 2";
      return 1;
    }
  } else {
    err = "model " + model + " is not supported on this path";
    return 1;
  }
  {
    auto on_key = [&](const std::string &w, std::istream &s) -> bool {
      if (w == "gamma") s >> gamma;
      else if (w == "upwinding") s >> upw_choice;
      else if (w == "include_chemistry") s >> include_chem_s;
      else if (model_id == MODEL_NS2D && w == "Pr") s >> Pr;
      else if (model_id == MODEL_NS2D && w == "Re") s >> Re;
      else if (model_id == MODEL_NS2D && w == "T_ref") s >> Tref;
      else if (model_id == MODEL_NS2D && w == "T_0") s >> T0;
      else if (model_id == MODEL_NS2D && w == "T_S") s >> TS;
      else if (model_id == MODEL_NS2D && w == "T_A") s >> TA;
      else if (model_id == MODEL_NS2D && w == "T_B") s >> TB;
      else if (model_id == MODEL_NS2D && w == "write_output") s >> write_op;
      else return false;
      return true;
    };
    if (read_keywords(join(dir, "physics.inp"), on_key, err) == 1) return 1;
  }
  if (upw_choice == "roe") upwind = UPW_ROE;
  else if (upw_choice == "rf-char") upwind = UPW_RF;
  else if (upw_choice == "llf-char") upwind = UPW_LLF;
  else if (upw_choice == "rusanov") upwind = UPW_RUSANOV;
  else {
    err = upw_choice + " is not a valid upwinding scheme on this path";
    return 1;
  }
  include_chem = (include_chem_s == "yes");
  // PhysicsOutput: Euler1DInitialize.c:145 always, NavierStokes2DInitialize.c:189
  // when write_output is "yes"
  species_output = include_chem && (model_id == MODEL_EULER1D || write_op == "yes");
  // the zones' gamma: NavierStokes2DInitialize.c:184 copies the physics.inp
  // value before chemistry replaces it; Euler1DInitialize never sets it
  for (Boundary &b : bcs) b.gamma = (model_id == MODEL_NS2D) ? gamma : 0.0;

  // ---- chemistry.inp: ChemistryInitialize.c:69-481 ------------------------
  if (include_chem) {
    double pi = 4.0 * atan(1.0), NA = 6.02214076e23, kB = 1.380649e-23, e = 1.60217663e-19;
    double M_O2 = 0.032, M_CO2 = 0.044, Cp_O2 = 918.45, Cv_O2 = 650.0, Cp_CO2 = 842.86, Cv_CO2 = 654.5;
    double mu0_O2 = 1.99e-5, kappa0_O2 = 2.70e-2, mu0_CO2 = 1.45e-5, kappa0_CO2 = 1.58e-2;
    double lambda_UV = 2.48e-7, theta = 0.17 * pi / 180, f_CO2 = 0.0, f_O3 = 0.005, Ptot = 101325.0;
    double Ti = 288, t_start = 0.0, t_pulse = 10 * 1e-9, Lz = 0.03, z_mm = 0.0;
    int nz = 22;
    double k0a = 0.9 * 3.3e-13, k0b = 0.1 * 3.3e-13, k1a = 0.8 * 3.95e-17, k1b = 0.2 * 3.95e-17;
    double k2a = 1.2e-16, k2b = 1.2e-16, k3a = 1.2e-17, k3b = 1.0e-17, k4 = 1.1e-16, k5 = 2.0e-16;
    double k6 = 0.2 * 3.0e-17, F0 = 2000, sO3 = 1.1e-21, IA = 1.0, IB = 1.0, IC = 0.0;
    std::string wz = "yes";
    auto on_key = [&](const std::string &w, std::istream &s) -> bool {
      struct { const char *k; double *v; } dk[] = {
          {"lambda_UV", &lambda_UV}, {"t_start", &t_start}, {"t_pulse", &t_pulse}, {"theta", &theta},
          {"f_CO2", &f_CO2}, {"f_O3", &f_O3}, {"Ptot", &Ptot}, {"Ti", &Ti}, {"Lz", &Lz},
          {"z_mm", &z_mm}, {"k0a", &k0a}, {"k1a", &k1a}, {"k2a", &k2a}, {"k3a", &k3a},
          {"k0b", &k0b}, {"k1b", &k1b}, {"k2b", &k2b}, {"k3b", &k3b}, {"k4", &k4}, {"k5", &k5},
          {"k6", &k6}, {"F0", &F0}, {"sO3", &sO3}, {"IA", &IA}, {"IB", &IB}, {"IC", &IC}};
      for (auto &p : dk)
        if (w == p.k) { s >> *p.v; return true; }
      if (w == "nz") { s >> nz; return true; }
      if (w == "write_all_zlocs") { s >> wz; return true; }
      return false;
    };
    if (read_keywords(join(dir, "chemistry.inp"), on_key, err) == 1) return 1;
    species_all_zlocs = (wz != "no");  // ChemistryWriteSpecies.c:34: only "no" restricts it
    if (f_CO2 > 1.0 || f_CO2 < 0.0 || f_O3 > 1.0 || f_O3 < 0.0) {
      err = "f_CO2 / f_O3 must be between 0 and 1";
      return 1;
    }
    Chem &ch = chem;
    double f_O2 = 1.0 - f_CO2;
    ch.R = NA * kB / (f_O2 * M_O2 + f_CO2 * M_CO2);
    double Cp = f_O2 * Cp_O2 + f_CO2 * Cp_CO2, Cv = f_O2 * Cv_O2 + f_CO2 * Cv_CO2;
    ch.gamma = Cp / Cv;
    double mu0 = f_O2 * mu0_O2 + f_CO2 * mu0_CO2, kappa0 = f_O2 * kappa0_O2 + f_CO2 * kappa0_CO2;
    double kUV = 2 * pi / lambda_UV, kg = 2 * kUV * sin(theta);
    ch.n_O2 = f_O2 * Ptot / (kB * Ti);
    double cs = sqrt(ch.gamma * ch.R * Ti);
    ch.L_ref = 1.0 / kg;
    ch.v_ref = cs;
    ch.t_ref = ch.L_ref / ch.v_ref;
    ch.rho_ref = Ptot / (ch.R * Ti);
    double P_ref = ch.rho_ref * ch.v_ref * ch.v_ref;
    ch.t_start_norm = t_start / ch.t_ref;
    ch.t_pulse_norm = t_pulse / ch.t_ref;
    double rate = kg * cs / ch.n_O2;
    ch.k0a = k0a / rate; ch.k1a = k1a / rate; ch.k2a = k2a / rate; ch.k3a = k3a / rate;
    ch.k0b = k0b / rate; ch.k1b = k1b / rate; ch.k2b = k2b / rate; ch.k3b = k3b / rate;
    ch.k4 = k4 / rate; ch.k5 = k5 / rate; ch.k6 = k6 / rate;
    ch.I0 = F0 / t_pulse;
    ch.nu = ch.c_light / lambda_UV;
    double Ei = P_ref / ((ch.gamma - 1.0) * ch.n_O2);
    ch.q0a = 0.73 * e / Ei; ch.q0b = 2.8 * e / Ei; ch.q1a = 0.29 * e / Ei; ch.q1b = 0.84 * e / Ei;
    ch.q2a = 4.3 * e / Ei; ch.q2b = 0.81 * e / Ei; ch.q3a = 0.42 * e / Ei; ch.q3b = 1.16 * e / Ei;
    ch.q4 = 1.23 * e / Ei; ch.q5 = 1.81 * e / Ei; ch.q6 = 0.42 * e / Ei;
    ch.dz = Lz / nz;
    ch.z_i = (int)(ceil(z_mm * 0.001 * nz / Lz));
    ch.nlayers = ch.z_i + 1;
    ch.n_react = 7 * ch.nlayers;
    ch.z_stride = 7;
    ch.nflow = nflow;
    ch.grid_stride = nflow + ch.n_react;
    ch.Ti = Ti;
    ch.sO3 = sO3;
    // imap (ChemistryInitialize.c:444-477): ghosted array, interior filled
    imap.assign(npts_g, 0.0);
    std::vector<double> im;
    bool found = false;
    if (read_array(*this, dir, "imap", 1, nullptr, im, found, err)) return 1;
    const int g = ghosts;
    const int nxg = dim[0] + 2 * g;
    for (long q = 0; q < npts_int; q++) {
      int i = (int)(q % dim[0]), j = (ndims == 2) ? (int)(q / dim[0]) : 0;
      long p = (ndims == 2) ? (long)(j + g) * nxg + (i + g) : (long)(i + g);
      if (found) imap[p] = im[q];
      else {
        double x0 = x[g], xx = x[g + i];
        imap[p] = IA + IB * cos(kg * ch.L_ref * (xx - x0) * (1.0 - IC * (xx - x0)));
      }
    }
    // model overrides: NavierStokes2DInitialize.c:188-216, Euler1DInitialize.c:150-158
    gamma = ch.gamma;
    if (model_id == MODEL_NS2D) {
      Tref = Ti;
      if (Re > 0) {
        ch.mu_ref = mu0 * exp(1.5 * log(Ti / T0)) * ((T0 + TS) / (Ti + TS));
        ch.kappa_ref = kappa0 * exp(1.5 * log(Ti / T0)) *
                       ((T0 + TA * exp(-TB / T0)) / (Ti + TA * exp(-TB / Ti)));
        Re = ch.rho_ref * ch.v_ref * ch.L_ref / ch.mu_ref;
        Pr = ch.gamma * ch.R / (ch.gamma - 1) * ch.mu_ref / ch.kappa_ref;
      }
    }
  }
  int expect = nflow + (include_chem ? chem.n_react : 0);
  if (nvars != expect) {
    err = "nvars has to be " + std::to_string(expect) + " in solver.inp";
    return 1;
  }
  if (nvars > 64) {
    err = "nvars > 64 not supported";
    return 1;
  }
  return 0;
}

}  // namespace piafs
