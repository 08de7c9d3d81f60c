// Host-side case description: the parsed PIAFS input decks plus every
// init-time constant the device path consumes.
#pragma once

#include <string>
#include <vector>

namespace piafs {

enum Model { MODEL_EULER1D = 1, MODEL_NS2D = 2 };
// boundary.inp zone types (include/boundaryconditions.h:15-42), in the order
// of the device zone kinds
enum BCType {
  BC_PERIODIC = 0, BC_EXTRAPOLATE, BC_DIRICHLET, BC_REFLECT, BC_SPONGE, BC_NOSLIP_WALL, BC_SLIP_WALL,
  BC_SUBSONIC_INFLOW, BC_SUBSONIC_OUTFLOW, BC_SUBSONIC_AMBIVALENT, BC_SUPERSONIC_INFLOW, BC_SUPERSONIC_OUTFLOW,
  BC_NTYPES
};
// W_UPW5: not a weight; the linear fifth-order upwind interpolation (upw5)
enum WeightKind { W_JS = 0, W_MAPPED = 1, W_Z = 2, W_YC = 3, W_UPW5 = 4 };
enum Upwind { UPW_ROE = 0, UPW_RF = 1, UPW_LLF = 2, UPW_RUSANOV = 3 };

struct Boundary {
  std::string type;
  int bctype = BC_PERIODIC;
  int dim = 0, face = 1;
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};
  int is[3] = {0, 0, 0}, ie[3] = {0, 0, 0};
  bool on = true;
  // type-specific data (InitializeBoundaries.c:95-150); gamma is the
  // physics.inp value for navierstokes2d and 0 for euler1d, which never sets it
  double rho = 0, vel[3] = {0, 0, 0}, p = 0, gamma = 0;
  std::vector<double> vals;  // dirichlet / sponge values (nvars), inflow passive scalars
};

// Photochemistry constants after normalisation (ChemistryInitialize.c:289-353).
struct Chem {
  int    nflow = 0, z_i = 0, nlayers = 1, z_stride = 7, grid_stride = 0, n_react = 0;
  double gamma = 1.4, R = 0, Ti = 288, n_O2 = 0, sO3 = 0, dz = 0, I0 = 0;
  double c_light = 3.0e8, h_planck = 6.62607015e-34, nu = 0;
  double t_start_norm = 0, t_pulse_norm = 0;
  double k0a, k0b, k1a, k1b, k2a, k2b, k3a, k3b, k4, k5, k6;
  double q0a, q0b, q1a, q1b, q2a, q2b, q3a, q3b, q4, q5, q6;
  double mu_ref = 0, kappa_ref = 0, rho_ref = 0, L_ref = 0, v_ref = 0, t_ref = 0;
};

struct Config {
  // solver.inp (ReadInputs.c:200-480)
  int ndims = 1, nvars = 1, ghosts = 1, n_iter = 0, restart_iter = 0;
  int dim[3] = {0, 0, 0}, dim_exact[3] = {0, 0, 0};
  int screen_op_iter = 1, file_op_iter = 1000;
  double dt = 0.0;
  std::string time_scheme = "euler", time_scheme_type = " ", hyp_scheme = "1", par_scheme = "2";
  std::string interp_type = "characteristic", ip_file_type = "ascii", input_mode = "serial";
  std::string output_mode = "serial", op_file_format = "text", op_overwrite = "no", model = "none";
  std::string conservation_check = "no";
  // boundary.inp
  std::vector<Boundary> bcs;
  // physics.inp
  int model_id = 0;
  double gamma = 1.4, Re = -1, Pr = 0.72, Tref = 273.15, T0 = 275.0, TS = 110.4, TA = 245.4, TB = 27.6;
  std::string upw_choice = "roe";
  int upwind = UPW_ROE;
  bool include_chem = false;
  bool species_output = false, species_all_zlocs = true;  // op_species_z* files
  int nflow = 0;
  // weno.inp
  int mapped = 0, borges = 0, yc = 0, no_limiting = 0;
  double eps = 1e-6;
  int weight_kind = W_JS;
  // selections
  bool crweno = false, upw5 = false, characteristic = false, par4 = false;
  // RK tableau (TimeExplicitRKInitialize.c:20-95)
  int nstages = 1;
  double A[16] = {0}, b[4] = {0}, c[4] = {0};
  // chemistry
  Chem chem;
  // grid and data (ghosted x / dxinv, interior u point-major, ghosted imap)
  std::vector<double> x, dxinv, u0, imap;
  long npts_int = 0, npts_g = 0;
  // member `member` of an ensemble of `nsims` simulations (simulation.inp,
  // EnsembleSimulationsDefine.cpp:31-119); set before parse
  int member = 0, nsims = 1;

  int parse(const std::string &dir, std::string &err);
  // the member's file-name tag: "" for a single simulation, else "_" and the
  // index zero-padded to floor(log10(nsims)) + 1 digits (GetStringFromInteger,
  // InitialSolution.c:36-41)
  std::string tag() const;
};

// simulation.inp of an ensemble directory: nsims (EnsembleSimulationsDefine.cpp:
// 47-96).  Returns 0 and nsims >= 1, 1 when the file is absent (a single
// simulation), 2 on a malformed file or nsims < 1 (err says why).
int read_ensemble(const std::string &dir, int &nsims, std::string &err);

// exact.inp (ExactSolution.c:17-60 with ReadArray): the interior solution,
// point-major; found is false when the file is absent
int read_exact(const Config &c, const std::string &dir, std::vector<double> &u, bool &found, std::string &err);

}  // namespace piafs
