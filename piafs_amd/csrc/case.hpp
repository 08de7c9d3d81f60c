// A PIAFS case resident on one GPU: parsed configuration, device buffers and
// the stream its time steps are queued on.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "config.hpp"
#include "kernels.hpp"

namespace piafs {

constexpr int NCLASS = 10;  // kernel classes timed by the profiling mode
const char *class_name(int c);
Grid make_grid(const Config &c);

struct Case {
  Config cfg;
  Grid G{};
  SweepGeom Sx{}, Sy{};
  BCMap bc{};
  PhysParams P{};
  ReconParams R{};
  int device = -1;
  hipStream_t stream = nullptr;
  // second stream for the y-sweep and viscous derivatives (joined by events)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool dual = false;
  double t = 0.0;
  std::string err;
  size_t bytes = 0;
  std::vector<void *> allocs;

  // state (natural planar layout), stage state and its transposed copy
  double *u = nullptr, *U = nullptr, *UT = nullptr, *Udot[4] = {nullptr, nullptr, nullptr, nullptr};
  double *flux = nullptr, *ifc[4] = {nullptr, nullptr, nullptr, nullptr}, *fIx = nullptr, *fIy = nullptr;
  double *scr_b = nullptr, *scr_c = nullptr, *hx = nullptr, *hxT = nullptr;
  double *Qd = nullptr, *FV = nullptr, *dxinv = nullptr, *imap = nullptr;
  double *partials = nullptr, *uprev = nullptr;
  // y-sweep buffers when the sweeps run concurrently
  double *flux_y = nullptr, *ifcy[4] = {nullptr, nullptr, nullptr, nullptr}, *scr_by = nullptr, *scr_cy = nullptr;
  int nparts = 0;

  // profiling: events bracketing kernel classes
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<int> ev_cls;
  size_t ev_used = 0;

  ~Case();
  int init(const std::string &dir, int dev);
  int fail(int code, const std::string &msg);
  double *dalloc(size_t n);
  int upload(const double *uh);
  int download(double *uh, const double *src = nullptr);
  void mark(int cls, bool begin, hipStream_t st = nullptr);
  void rhs_from_stage(double time, double *out);
  void enqueue_step();
  int step(int n, double *cfl, double *norm);
  int rhs_host(const double *uh, double time, double *rh);
  int write_binary(const std::string &path);
  static size_t plan_bytes(const Config &c);
};

}  // namespace piafs
