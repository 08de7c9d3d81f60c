// Ensembles and continuous batching (include/piafs_amd.h, piafs_batch_*).
//
// A batch owns one engine: every member case is advanced by the same
// launches (one per kernel class per RK stage), replayed from a captured HIP
// graph.  Members are addressed by stable slot numbers; admitting or removing
// a case between steps rebuilds the block tables and re-captures the graph.
// Replaces the serial per-simulation loop of the reference's ensemble mode
// (TimeRK.c:46-81, EnsembleSimulationsDefine.cpp:31-119).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/piafs_amd.h"
#include "case.hpp"

using namespace piafs;

struct piafs_batch {
  Engine eng;
  std::vector<Case *> slots;  // nullptr = free slot
  std::vector<int> steps;     // steps taken by each slot since admission
  std::string err;

  int rebuild() {
    eng.cases.clear();
    for (Case *c : slots)
      if (c) eng.cases.push_back(c);
    int rc = eng.build();
    if (rc) err = eng.err;
    return rc;
  }
};

static Case *slot_case(piafs_batch *b, int k) {
  if (!b || k < 0 || k >= (int)b->slots.size()) return nullptr;
  return b->slots[k];
}

extern "C" {

int piafs_batch_create(const char *const *dirs, int ncases, int device, piafs_batch **out) {
  if (!out || (ncases > 0 && !dirs) || ncases < 0) return 1;
  *out = nullptr;
  piafs_batch *b = new piafs_batch();
  b->eng.owns_cases = false;
  b->eng.use_graph = true;
  int rc = b->eng.init(device);
  if (rc) {
    delete b;
    return rc;
  }
  for (int k = 0; k < ncases; k++) {
    Case *c = new Case();
    rc = c->init(dirs[k], device, b->eng.stream);
    if (rc) {
      b->err = "case " + std::to_string(k) + ": " + c->err;
      delete c;
      for (Case *x : b->slots) delete x;
      b->slots.clear();
      *out = b;  // handle returned so that the error can be read; destroy it
      return rc;
    }
    b->slots.push_back(c);
    b->steps.push_back(0);
  }
  rc = b->rebuild();
  *out = b;
  return rc;
}

int piafs_batch_destroy(piafs_batch *b) {
  if (!b) return 0;
  if (b->eng.stream) b->eng.sync();  // an engine that never started has nothing to wait for
  b->eng.cases.clear();
  for (Case *c : b->slots) delete c;
  delete b;
  return 0;
}

int piafs_batch_use_graph(piafs_batch *b, int on) {
  if (!b) return 1;  // no batch
  b->eng.sync();
  b->eng.use_graph = (on != 0);
  return b->rebuild();
}

int piafs_batch_add(piafs_batch *b, const char *dir, int *slot) {
  if (!b) return 1;  // no batch
  b->eng.sync();
  Case *c = new Case();
  int rc = c->init(dir ? dir : "", b->eng.device, b->eng.stream);
  if (rc) {
    b->err = c->err;
    delete c;
    return rc;
  }
  int k = 0;
  while (k < (int)b->slots.size() && b->slots[k]) k++;
  if (k == (int)b->slots.size()) {
    b->slots.push_back(nullptr);
    b->steps.push_back(0);
  }
  b->slots[k] = c;
  b->steps[k] = 0;
  if (slot) *slot = k;
  return b->rebuild();
}

int piafs_batch_remove(piafs_batch *b, int k) {
  if (!b) return 1;  // no batch
  Case *c = slot_case(b, k);
  if (!c) return 1;
  b->eng.sync();
  b->slots[k] = nullptr;
  int rc = b->rebuild();
  delete c;
  return rc;
}

int piafs_batch_step(piafs_batch *b, int nsteps) {
  if (!b) return 1;  // no batch
  int rc = b->eng.step(nsteps);
  if (rc) {
    b->err = b->eng.err;
    return rc;
  }
  for (size_t k = 0; k < b->slots.size(); k++)
    if (b->slots[k]) b->steps[k] += nsteps;
  return 0;
}

int piafs_batch_sync(piafs_batch *b) {
  if (!b) return 1;  // no batch
  int rc = b->eng.sync();
  if (rc) b->err = b->eng.err;
  return rc;
}

int piafs_batch_download(piafs_batch *b, int k, double *u) {
  if (!b) return 1;  // no batch
  Case *c = slot_case(b, k);
  if (!c) return 1;
  int rc = c->download(u, b->eng.stream);
  if (rc) b->err = c->err;
  return rc;
}

int piafs_batch_info(piafs_batch *b, int k, piafs_info *info) {
  if (!b) return 1;  // no batch
  Case *c = slot_case(b, k);
  if (!c || !info) return 1;
  fill_info(c, info);
  return 0;
}

int piafs_batch_autotune(piafs_batch *b, int *rows_per_lane, double *recon_ms) {
  if (!b) return 1;  // no batch
  static const int cands[] = {4, 8, 12, 16};
  b->eng.sync();
  int rc = b->eng.autotune(cands, 4, 3, rows_per_lane, recon_ms);
  if (rc) b->err = b->eng.err;
  return rc;
}

int piafs_batch_steps_done(piafs_batch *b, int k, int *steps) {
  if (!b) return 1;  // no batch
  if (!slot_case(b, k) || !steps) return 1;
  *steps = b->steps[k];
  return 0;
}

const char *piafs_batch_last_error(const piafs_batch *b) { return b ? b->err.c_str() : ""; }

int piafs_ensemble_create(const char *dir, int device, piafs_batch **out) {
  if (!out) return 1;
  *out = nullptr;
  const std::string d = dir ? dir : "";
  int nsims = 0;
  std::string err;
  const int r = read_ensemble(d, nsims, err);
  piafs_batch *b = new piafs_batch();
  if (r) {
    b->err = r == 1 ? "simulation.inp not found" : err;
    *out = b;  // handle returned so that the error can be read; destroy it
    return 1;
  }
  b->eng.owns_cases = false;
  b->eng.use_graph = true;
  int rc = b->eng.init(device);
  if (rc) {
    b->err = b->eng.err;
    *out = b;
    return rc;
  }
  for (int k = 0; k < nsims; k++) {  // members in index order: slot k is member k
    Case *c = new Case();
    c->cfg.member = k;
    c->cfg.nsims = nsims;
    rc = c->init(d, device, b->eng.stream);
    if (rc) {
      b->err = "domain " + std::to_string(k) + ": " + c->err;
      delete c;
      for (Case *x : b->slots) delete x;
      b->slots.clear();
      b->steps.clear();
      *out = b;
      return rc;
    }
    b->slots.push_back(c);
    b->steps.push_back(0);
  }
  rc = b->rebuild();
  *out = b;
  return rc;
}

int piafs_ensemble_solve(piafs_batch *b, const char *dir, int verbose, double *seconds) {
  if (!b) return 1;  // no batch
  std::vector<Case *> cs;
  for (Case *c : b->slots)
    if (c) cs.push_back(c);
  if (cs.empty()) {
    b->err = "the batch has no members";
    return 1;
  }
  for (size_t k = 0; k < cs.size(); k++)
    if (cs[k]->cfg.nsims != (int)cs.size() || cs[k]->cfg.member != (int)k) {
      b->err = "the batch is not the ensemble of one directory (members added or removed)";
      return 1;
    }
  const int rc = run_solve(b->eng, cs, dir ? dir : "", verbose, seconds, b->err);
  if (!rc)
    for (size_t k = 0; k < b->slots.size(); k++)
      if (b->slots[k]) b->steps[k] += cs[0]->cfg.n_iter - cs[0]->cfg.restart_iter;
  return rc;
}

}  // extern "C"
