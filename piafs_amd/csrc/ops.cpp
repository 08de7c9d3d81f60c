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

extern "C" int piafs_tridiag_host(const double *a, const double *b, const double *c, double *x, int n, int ns) {
  if (n <= 0 || ns <= 0) return 0;
  const size_t m = (size_t)n * ns;
  DevBuf da(m), db(m), dc(m), dx(m);
  int *flag = nullptr;
  if (!da.p || !db.p || !dc.p || !dx.p || hipMalloc(&flag, sizeof(int)) != hipSuccess) return -2;
  hipMemcpy(da.p, a, m * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(db.p, b, m * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(dc.p, c, m * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(dx.p, x, m * sizeof(double), hipMemcpyHostToDevice);
  hipMemset(flag, 0, sizeof(int));
  if (!piafs::launch_tridiag(da.p, db.p, dc.p, dx.p, n, ns, flag, 0)) {
    hipFree(flag);
    return 1;  // systems longer than one team covers
  }
  int hflag = 0;
  hipMemcpy(x, dx.p, m * sizeof(double), hipMemcpyDeviceToHost);
  hipMemcpy(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost);
  hipFree(flag);
  if (hipGetLastError() != hipSuccess) return -1;
  return hflag ? -1 : 0;
}

extern "C" int piafs_weno_weights_host(int kind, const double *copt, const double *stencils, int count, double eps,
                                       double *w_out) {
  if (count <= 0) return 0;
  if (kind < 0 || kind > 3) return 1;
  DevBuf ds((size_t)count * 5), dw((size_t)count * 3);
  if (!ds.p || !dw.p) return -2;
  hipMemcpy(ds.p, stencils, (size_t)count * 5 * sizeof(double), hipMemcpyHostToDevice);
  piafs::launch_weno_weights(kind, copt, ds.p, count, eps, dw.p, 0);
  hipMemcpy(w_out, dw.p, (size_t)count * 3 * sizeof(double), hipMemcpyDeviceToHost);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
