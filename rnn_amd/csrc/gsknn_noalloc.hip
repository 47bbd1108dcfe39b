// gsknn_dknn_noalloc_dev: gsknn_dknn_dev for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  No counterpart in the reference library.
//
// The workspace of a call (the partial (distance, index) lists of a launch
// that splits its candidates over workgroups) comes from the file-scope
// device array `scratch` below instead of hipMallocAsync.  Each GPU holds its
// own copy of the array until the process ends; its device address is looked
// up once per device with hipGetSymbolAddress on the first call made with
// that device current, and only that address is ever passed to kernels.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (different streams, or threads) race on it and must be serialised by
// the caller.  Nothing in the array has to start at zero (every partial list
// a merge reads is written in full by the same call), so calls leave its
// contents as they are.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gsknn_hip.h"
#include "gsknn_common.h"

namespace {

constexpr size_t kScratchBytes = GSKNN_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's `scratch` in *out, or an error
// code (gsknn_last_error() set).
int device_scratch(const char *fn, void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return gsknn::api_error(GSKNN_EHIP, fn, hipGetErrorString(e), (int)e);
  if (dev < 0 || dev >= kMaxDevices)
    return gsknn::api_error(GSKNN_ELIMIT, fn, "device index beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return gsknn::api_error(GSKNN_EHIP, fn, "hipGetSymbolAddress(scratch) failed", (int)e);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GSKNN_OK;
}

}  // namespace

extern "C" int gsknn_dknn_noalloc_dev(int m, int n, int d, int k, const double *XA,
                                      const double *XA2, const int *amap, const double *XB,
                                      const double *XB2, const int *bmap, double *D, int *I,
                                      int ldk, const int *hrow, int mode, void *stream) {
  const char *fn = "gsknn_dknn_noalloc_dev";
  void *work = nullptr;
  const int rc = device_scratch(fn, &work);
  if (rc != GSKNN_OK) return rc;
  return gsknn::dknn_dev_fixed_work(fn, m, n, d, k, XA, XA2, amap, XB, XB2, bmap, D, I, ldk, hrow,
                                    mode, work, kScratchBytes, stream);
}
