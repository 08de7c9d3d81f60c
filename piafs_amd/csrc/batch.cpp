// Batched ensembles (include/piafs_amd.h, piafs_batch_*).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/piafs_amd.h"
#include "case.hpp"

struct piafs_batch {
  std::vector<piafs_case *> cases;
  std::string err;
};

extern "C" {

int piafs_batch_create(const char *const *dirs, int ncases, int device, piafs_batch **out) {
  *out = nullptr;
  piafs_batch *b = new piafs_batch();
  for (int k = 0; k < ncases; k++) {
    piafs_case *c = nullptr;
    int rc = piafs_create(dirs[k], device, &c);
    if (rc) {
      b->err = std::string("case ") + std::to_string(k) + ": " + piafs_last_error(nullptr);
      for (auto *x : b->cases) piafs_destroy(x);
      delete b;
      return rc;
    }
    b->cases.push_back(c);
  }
  *out = b;
  return 0;
}

int piafs_batch_destroy(piafs_batch *b) {
  if (!b) return 0;
  for (auto *c : b->cases) piafs_destroy(c);
  delete b;
  return 0;
}

int piafs_batch_step(piafs_batch *b, int nsteps) {
  for (int s = 0; s < nsteps; s++)
    for (auto *c : b->cases) {
      int rc = piafs_step(c, 1);
      if (rc) return rc;
    }
  return 0;
}

int piafs_batch_sync(piafs_batch *b) {
  for (auto *c : b->cases) {
    int rc = piafs_sync(c);
    if (rc) return rc;
  }
  return 0;
}

int piafs_batch_download(piafs_batch *b, int k, double *u) {
  if (k < 0 || k >= (int)b->cases.size()) return 1;
  return piafs_download_state(b->cases[k], u);
}

int piafs_batch_info(piafs_batch *b, int k, piafs_info *info) {
  if (k < 0 || k >= (int)b->cases.size()) return 1;
  return piafs_get_info(b->cases[k], info);
}

const char *piafs_batch_last_error(const piafs_batch *b) { return b ? b->err.c_str() : ""; }

}  // extern "C"
