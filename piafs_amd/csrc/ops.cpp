// Per-operator entry points of include/piafs_amd.h (host arrays in and out).
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/piafs_amd.h"
#include "kernels.hpp"

namespace {

struct DevBuf {
  double *p = nullptr;
  explicit DevBuf(size_t n) {
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) p = nullptr;
  }
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

}  // namespace

// Both entry points are unit-level checks of one device function: they
// allocate their device buffers per call and synchronise (host arrays in and
// out), which the time step never does.
extern "C" int piafs_tridiag_host(const double *a, const double *b, const double *c, double *x, int n, int ns) {
  if (n <= 0 || ns <= 0) return 0;
  if (!a || !b || !c || !x) return 1;
  const size_t m = (size_t)n * ns;
  DevBuf da(m), db(m), dc(m), dx(m), dflag(1);
  if (!da.p || !db.p || !dc.p || !dx.p || !dflag.p) return -2;
  int *flag = (int *)dflag.p;
  if (hipMemcpy(da.p, a, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db.p, b, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dc.p, c, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx.p, x, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(flag, 0, sizeof(int)) != hipSuccess)
    return -1;
  if (!piafs::launch_tridiag(da.p, db.p, dc.p, dx.p, n, ns, flag, 0)) return 1;
  if (hipGetLastError() != hipSuccess) return -1;
  int hflag = 0;
  if (hipMemcpy(x, dx.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  return hflag ? -1 : 0;
}

extern "C" int piafs_weno_weights_host(int kind, const double *copt, const double *stencils, int count, double eps,
                                       double *w_out) {
  if (count <= 0) return 0;
  if (kind < 0 || kind > 3 || !copt || !stencils || !w_out) return 1;
  DevBuf ds((size_t)count * 5), dw((size_t)count * 3);
  if (!ds.p || !dw.p) return -2;
  if (hipMemcpy(ds.p, stencils, (size_t)count * 5 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return -1;
  piafs::launch_weno_weights(kind, copt, ds.p, count, eps, dw.p, 0);
  if (hipGetLastError() != hipSuccess) return -1;
  if (hipMemcpy(w_out, dw.p, (size_t)count * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return 0;
}
