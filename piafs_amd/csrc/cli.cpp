// piafs_amd: drop-in driver for the PIAFS executable (src/main.cpp:68-247).
// Reads the input decks from the working directory (or the directory given as
// the first argument), runs n_iter steps on GPU 0 and writes op[_NNNNN].bin
// files as the reference does for op_file_format binary; an ensemble
// directory (simulation.inp) runs all its members as one batch.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/piafs_amd.h"

// An ensemble directory (simulation.inp present): every member advanced
// together, files under the member tags (main.cpp:92-111, Solve.cpp:28-110).
static int run_ensemble(const char *dir, int device) {
  printf("-- Ensemble Simulation --\n");
  piafs_batch *b = nullptr;
  int rc = piafs_ensemble_create(dir, device, &b);
  if (rc) {
    fprintf(stderr, "Error: %s (code %d)\n", piafs_batch_last_error(b), rc);
    piafs_batch_destroy(b);
    return 1;
  }
  for (int k = 0;; k++) {
    piafs_info info;
    if (piafs_batch_info(b, k, &info)) break;
    printf("piafs_amd: domain %d: model %d, ndims %d, nvars %d, size %d x %d, %d steps of dt %g\n", k, info.model,
           info.ndims, info.nvars, info.dim[0], info.ndims > 1 ? info.dim[1] : 1, info.n_iter, info.dt);
  }
  double secs = 0;
  rc = piafs_ensemble_solve(b, dir, 1, &secs);
  if (rc) {
    fprintf(stderr, "Error: %s (code %d)\n", piafs_batch_last_error(b), rc);
    piafs_batch_destroy(b);
    return 1;
  }
  printf("Solver runtime (in seconds): %1.16E\n", secs);
  piafs_batch_destroy(b);
  return 0;
}

int main(int argc, char **argv) {
  const char *dir = argc > 1 ? argv[1] : nullptr;
  int device = 0;
  if (const char *d = getenv("PIAFS_DEVICE")) device = atoi(d);
  {
    std::string sim = dir ? std::string(dir) + "/simulation.inp" : std::string("simulation.inp");
    if (FILE *f = fopen(sim.c_str(), "r")) {
      fclose(f);
      return run_ensemble(dir, device);
    }
  }
  piafs_case *c = nullptr;
  int rc = piafs_create(dir, device, &c);
  if (rc) {
    fprintf(stderr, "Error: %s (code %d)\n", piafs_last_error(nullptr), rc);
    return 1;
  }
  piafs_info info;
  piafs_get_info(c, &info);
  printf("piafs_amd: model %d, ndims %d, nvars %d, size %d x %d, %d steps of dt %g\n", info.model, info.ndims,
         info.nvars, info.dim[0], info.ndims > 1 ? info.dim[1] : 1, info.n_iter, info.dt);
  if (info.include_chem) printf("    gamma %1.4e  Re %1.4e  Pr %1.4e\n", info.gamma, info.Re, info.Pr);
  double secs = 0;
  rc = piafs_solve(c, dir, 1, &secs);
  if (rc) {
    fprintf(stderr, "Error: %s (code %d)\n", piafs_last_error(c), rc);
    piafs_destroy(c);
    return 1;
  }
  printf("Solver runtime (in seconds): %1.16E\n", secs);
  piafs_destroy(c);
  return 0;
}
