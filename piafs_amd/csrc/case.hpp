// Cases resident on one GPU and the engine that advances an ensemble of them.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/piafs_amd.h"
#include "config.hpp"
#include "kernels.hpp"

namespace piafs {

Grid make_grid(const Config &c);
std::string output_extension(const Config &c);  // ".bin", ".dat" or "" (op_file_format none)

// One PIAFS simulation: parsed configuration, device buffers and the
// descriptor the kernels read.
struct Case {
  Config cfg;
  CaseDev h{};          // host copy of the device descriptor
  int device = -1;
  size_t bytes = 0;
  double t = 0.0;       // host mirror of the simulation time (t += dt per step)
  std::string err;
  std::vector<void *> allocs;
  double *uprev = nullptr, *partials = nullptr;
  int nparts = 0;

  ~Case();
  // parse the decks in `dir` (host only), then allocate and upload
  int init(const std::string &dir, int dev, hipStream_t st);
  // device half of init for a configuration already parsed into cfg
  int setup(int dev, hipStream_t st);
  int fail(int code, const std::string &msg);
  double *dalloc(size_t n, hipStream_t st);
  int upload(const double *uh, hipStream_t st);
  int download(double *uh, hipStream_t st, const double *src = nullptr);
  int write_binary(const std::string &path, hipStream_t st);
  // run outputs (output.cpp): the solution in the deck's op_file_format, the
  // species files, the exact-solution errors and errors.dat & co.
  int write_solution(const std::string &path, hipStream_t st);
  int write_species(const std::string &dir, const std::string &suffix, double stage_time, hipStream_t st);
  int exact_errors(const std::string &dir, double err3[3], hipStream_t st);
  int write_run_files(const std::string &dir, const double err3[3], long steps, double solver_rt, double main_rt,
                      hipStream_t st);
  // conservation errors per variable (CalculateConservationError.c:20-40); -1 when unchecked
  int conservation_errors(std::vector<double> &ce, hipStream_t st);
  // lanes per line system: segments of at most max_rows rows (descriptor only;
  // the engine's plan must be rebuilt afterwards)
  int set_team_rows(int max_rows);
  static size_t plan_bytes(const Config &c);
};

// An ensemble of cases advanced together: one launch per kernel class per RK
// stage covering every case, optionally replayed from a captured HIP graph.
struct Engine {
  std::vector<Case *> cases;
  bool owns_cases = true;
  int device = 0;
  hipStream_t stream = nullptr;
  StageFork fork;              // side stream of the viscous branch (null: one stream)
  CaseDev *d_cases = nullptr;
  int *d_status = nullptr;  // one status word per member case
  Plan plan;
  std::vector<void *> plan_allocs;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::string err;

  // profiling: events bracketing kernel classes (eager launches only)
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<int> ev_cls;
  size_t ev_used = 0;

  ~Engine();
  int init(int dev);
  int add(Case *c);
  bool built = false;          // plan and descriptor table valid for the member set
  const CaseDev *override_desc = nullptr;  // one-member engine: descriptor to use instead of the case's
  int build();                 // descriptor table and plans for the current member set
  int prepare();               // capture the step graph now (when graphs are used)
  void enqueue_step();         // eager launches of one full step
  int step(int n);             // n steps (graph replay when enabled)
  int stage_only(int s, int last = K_NCLASS);  // one stage, or its classes up to `last`
  double time_class(int cls, int reps, int cls2 = -1);
  int autotune(const int *cands, int ncand, int reps, int *best_rows, double *best_ms);
  int sync();
  void mark(int cls, bool begin, hipStream_t st);
  int fail(int code, const std::string &msg);
  void release_plan();
};

// The device-buffer operators of one case (devops.cpp): a one-member engine
// whose descriptor is the case's in split mode (uL, uR, fL, fR kept apart),
// with a state array of its own for the staged interior state.
struct UnitOps {
  Engine eng;
  CaseDev desc{};
  double *u = nullptr;                         // staged state, planar and ghosted
  double *io[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // fL, fR of each sweep when the case is paired
};

}  // namespace piafs

// C ABI handle of a single case: an engine with one member.
struct piafs_case {
  piafs::Engine eng;
  piafs::Case *c = nullptr;
  std::string err;
  std::unique_ptr<piafs::UnitOps> unit;  // built at the first device-buffer operator call
};

namespace piafs {
void fill_info(const Case *c, piafs_info *info);
// One step of every member of `eng` (the members `cs`) with the largest CFL
// number and the norm of the change over all of them; err on failure.
int step_with_diagnostics(Engine &eng, const std::vector<Case *> &cs, double *cfl, double *norm,
                          std::string &err);
// The reference's run (n_iter steps, screen and file output, errors and the
// run files) over the members `cs` of `eng`, in the directory d.
int run_solve(Engine &eng, const std::vector<Case *> &cs, const std::string &d, int verbose, double *seconds,
              std::string &err);
}  // namespace piafs
