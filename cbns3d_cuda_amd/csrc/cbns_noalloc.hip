// cbns_noalloc.hip -- cbns_calc_gradient_noalloc: cbns_calc_gradient for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  cbns_calc_gradient (like
// the reference's calc_gradient, src/calc_gradient.cu:17-51) allocates its
// device staging for f, fx, fy, fz on every call; this variant stages them
// in `scratch`, a file-scope device array of CBNS_NOALLOC_MAX_NODES nodes x 4
// doubles (32 MiB) that every GPU the library runs on holds until the
// process ends, and runs the same gradient kernel.  Its device address is
// looked up once per device (hipGetSymbolAddress with that device current)
// and only that address reaches the runtime.  Calls share the array of
// their GPU, so calls that overlap on one GPU (several host threads) race.
#include <mutex>

#include "cbns_internal.hpp"

namespace cbns {
namespace {

static __device__ __align__(256) double scratch[4 * (long)CBNS_NOALLOC_MAX_NODES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_dev[MAX_DEVICES] = {};

// this device's copy of scratch
int scratch_on_current_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CBNS_ERR_HIP;
  if (dev < 0 || dev >= MAX_DEVICES) return CBNS_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CBNS_ERR_HIP;
    scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = scratch_dev[dev];
  return CBNS_OK;
}

}  // namespace
}  // namespace cbns

extern "C" int cbns_calc_gradient_noalloc(int32_t IM, int32_t JM, int32_t KM, const double* f,
                                          double* fx, double* fy, double* fz) {
  if (IM < 5 || JM < 5 || KM < 5 || !f || !fx || !fy || !fz) return CBNS_ERR_ARG;
  const long NN = (long)IM * JM * KM;
  if (NN > CBNS_NOALLOC_MAX_NODES) return CBNS_ERR_SIZE;
  double* d = nullptr;
  int rc = cbns::scratch_on_current_device(&d);
  if (rc) return rc;
  // every node of fx, fy, fz is written by the kernel and f is copied in
  // whole, so nothing of the previous call's contents is read
  const size_t nb = sizeof(double) * NN;
  if (hipMemcpy(d, f, nb, hipMemcpyHostToDevice) != hipSuccess) return CBNS_ERR_HIP;
  rc = cbns::launch_gradient_nodes(nullptr, IM, JM, KM, d, d + NN, d + 2 * NN, d + 3 * NN);
  if (rc) return rc;
  if (hipMemcpy(fx, d + NN, nb, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(fy, d + 2 * NN, nb, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(fz, d + 3 * NN, nb, hipMemcpyDeviceToHost) != hipSuccess)
    return CBNS_ERR_HIP;
  return CBNS_OK;
}
