// Allocation-free variant of the batched tridiagonal entry point, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Its staging buffers and
// zero-pivot flag live in one file-scope device array of fixed size; each
// device the library runs on holds its own copy until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/piafs_amd.h"
#include "kernels.hpp"

namespace {

// capacity of the scratch array in doubles: the four staged arrays a, b, c, x
// (n*ns doubles each) and the flag (one double slot) must fit
constexpr size_t kScratchDoubles = PIAFS_NOALLOC_SCRATCH_DOUBLES;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

// The scratch array's address on the current device, looked up once per
// device (each device holds its own copy at its own address).
int scratch_on_current_device(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    g_scratch[dev] = (double *)p;
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int piafs_tridiag_host_noalloc(const double *a, const double *b, const double *c, double *x, int n,
                                          int ns) {
  if (n <= 0 || ns <= 0) return 0;
  if (!a || !b || !c || !x) return 1;
  const size_t m = (size_t)n * ns;
  if (m > (kScratchDoubles - 1) / 4) return 2;  // beyond the scratch array
  double *s = nullptr;
  if (scratch_on_current_device(&s)) return -1;
  double *da = s, *db = s + m, *dc = s + 2 * m, *dx = s + 3 * m;
  int *flag = (int *)(s + 4 * m);
  if (hipMemcpy(da, a, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dc, c, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx, x, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(flag, 0, sizeof(int)) != hipSuccess)  // the last call's flag must not carry over
    return -1;
  if (!piafs::launch_tridiag(da, db, dc, dx, n, ns, flag, 0)) return 1;
  if (hipGetLastError() != hipSuccess) return -1;
  int hflag = 0;
  if (hipMemcpy(x, dx, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  return hflag ? -1 : 0;
}
