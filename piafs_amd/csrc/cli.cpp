// piafs_amd: drop-in driver for the PIAFS executable (src/main.cpp:68-247).
// Reads the input decks from the working directory (or the directory given as
// the first argument), runs n_iter steps on GPU 0 and writes op[_NNNNN].bin
// files as the reference does for op_file_format binary.
#include <cstdio>
#include <cstdlib>

#include "../../include/piafs_amd.h"

int main(int argc, char **argv) {
  const char *dir = argc > 1 ? argv[1] : nullptr;
  int device = 0;
  if (const char *d = getenv("PIAFS_DEVICE")) device = atoi(d);
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
