// C-ABI of libgsknn.so: the reference's public API (include/gsknn.h) and the
// device-array extensions (include/gsknn_hip.h), over the HIP kernels in
// gsknn_kernels.hip.
//
// Host-pointer calls stage the point tables, maps, norms and heap rows into
// device memory, launch, copy the heap rows back and synchronise, so that the
// call behaves like the reference's synchronous CPU call
// (frame/gsknn.c:1106-1184).  Device pointers are used in place.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/gsknn.h"
#include "../../include/gsknn_hip.h"
#include "gsknn_common.h"

using namespace gsknn;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

static int hip_err(hipError_t e, const char *what) {
  return set_err(GSKNN_EHIP, "%s: %s", what, hipGetErrorString(e));
}

extern "C" const char *gsknn_last_error(void) { return g_err; }

static std::atomic<int> g_device{0};

extern "C" int gsknn_set_device(int device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess) return hip_err(e, "hipGetDeviceCount");
  if (device < 0 || device >= count)
    return set_err(GSKNN_EINVAL, "gsknn_set_device: device %d of %d", device, count);
  g_device = device;
  return GSKNN_OK;
}

// ---------------------------------------------------------------------------
// Device-array API
// ---------------------------------------------------------------------------
static int check_sizes(const char *fn, int m, int n, int d, int k, int ldk) {
  if (m < 0 || n < 0 || d < 0 || k < 0)
    return set_err(GSKNN_EINVAL, "%s: negative size (m=%d n=%d d=%d k=%d)", fn, m, n, d, k);
  if (ldk < k) return set_err(GSKNN_EINVAL, "%s: ldk=%d < k=%d", fn, ldk, k);
  return GSKNN_OK;
}

// allow_alloc = false: the caller's workspace must suffice (no allocation of
// any kind happens in the call).
template <typename T>
static int knn_dev(const char *fn, int m, int n, int d, int k, const T *XA, const T *XA2,
                   const int *amap, const T *XB, const T *XB2, const int *bmap, T *D, int *I,
                   int ldk, const int *hrow, int mode, void *work, size_t work_bytes,
                   void *stream, bool allow_alloc = true) {
  int rc = check_sizes(fn, m, n, d, k, ldk);
  if (rc) return rc;
  if (m == 0 || n == 0 || d == 0 || k == 0) return GSKNN_OK;  // gsknn.c:909-913
  if (!XA || !XA2 || !amap || !XB || !XB2 || !bmap || !D || !I)
    return set_err(GSKNN_EINVAL, "%s: null array", fn);
  hipStream_t s = (hipStream_t)stream;
  size_t need = knn_work_bytes(m, n, d, k, (int)sizeof(T), mode);
  void *own = nullptr;
  if (!allow_alloc && need > work_bytes) {
    need = knn_work_bytes_min(m, n, d, k, (int)sizeof(T), mode);
    if (need > work_bytes)
      return set_err(GSKNN_ELIMIT, "%s: the call needs %zu bytes of scratch, more than the %zu available",
                     fn, need, work_bytes);
  }
  // A caller's buffer of at least the least workspace is used as it is (the
  // FP64 MFMA per-wave forms then run in one stage).
  if (need > 0 && work != nullptr && work_bytes < need) {
    const size_t least = knn_work_bytes_min(m, n, d, k, (int)sizeof(T), mode);
    if (work_bytes < least)
      return set_err(GSKNN_EINVAL, "%s: workspace %zu < %zu bytes", fn, work_bytes, least);
  }
  if (need > 0 && work == nullptr) {
    hipError_t e = hipMallocAsync(&own, need, s);
    if (e != hipSuccess) return hip_err(e, "hipMallocAsync(workspace)");
    work = own;
    work_bytes = need;
  }
  KnnArgs<T> a{};
  a.m = m; a.n = n; a.d = d; a.k = k; a.ldk = ldk;
  a.XA = XA; a.XA2 = XA2; a.amap = amap;
  a.XB = XB; a.XB2 = XB2; a.bmap = bmap;
  a.D = D; a.I = I; a.hrow = hrow;
  hipError_t e;
  if constexpr (sizeof(T) == 8)
    e = launch_knn_f64(a, mode, s, work, work_bytes);
  else
    e = launch_knn_f32(a, mode, s, work, work_bytes);
  if (own) (void)hipFreeAsync(own, s);
  if (e == hipErrorNotSupported)
    return set_err(GSKNN_ELIMIT, "%s: no kernel configuration fits d=%d k=%d", fn, d, k);
  if (e != hipSuccess) return hip_err(e, fn);
  return GSKNN_OK;
}

extern "C" size_t gsknn_dknn_work_bytes(int m, int n, int d, int k, int mode) {
  return knn_work_bytes(m, n, d, k, 8, mode);
}

extern "C" size_t gsknn_sknn_work_bytes(int m, int n, int d, int k, int mode) {
  return knn_work_bytes(m, n, d, k, 4, mode);
}

extern "C" size_t gsknn_sknn_overflow_offset(int m, int n, int d, int k, int mode) {
  if (m <= 0 || n <= 0 || d <= 0 || k <= 0 || !knn_f32_takes_mfma(m, d, k, mode)) return (size_t)-1;
  return f32_overflow_offset(m, n, d, k);
}

extern "C" size_t gsknn_dknn_fallback_offset(int m, int n, int d, int k, int mode) {
  return knn_two_stage_count_offset(m, n, d, k, mode);
}

extern "C" int gsknn_dknn_dev(int m, int n, int d, int k, const double *XA, const double *XA2,
                              const int *amap, const double *XB, const double *XB2,
                              const int *bmap, double *D, int *I, int ldk, const int *hrow,
                              int mode, void *work, size_t work_bytes, void *stream) {
  if (mode < 0 || mode > 3) return set_err(GSKNN_EINVAL, "gsknn_dknn_dev: mode %d", mode);
  return knn_dev("gsknn_dknn_dev", m, n, d, k, XA, XA2, amap, XB, XB2, bmap, D, I, ldk, hrow,
                 mode, work, work_bytes, stream);
}

extern "C" int gsknn_sknn_dev(int m, int n, int d, int k, const float *XA, const float *XA2,
                              const int *amap, const float *XB, const float *XB2,
                              const int *bmap, float *D, int *I, int ldk, const int *hrow,
                              int mode, void *work, size_t work_bytes, void *stream) {
  if (mode < 0 || mode > 3) return set_err(GSKNN_EINVAL, "gsknn_sknn_dev: mode %d", mode);
  return knn_dev("gsknn_sknn_dev", m, n, d, k, XA, XA2, amap, XB, XB2, bmap, D, I, ldk, hrow,
                 mode, work, work_bytes, stream);
}

int gsknn::dknn_dev_fixed_work(const char *fn, int m, int n, int d, int k, const double *XA,
                               const double *XA2, const int *amap, const double *XB,
                               const double *XB2, const int *bmap, double *D, int *I, int ldk,
                               const int *hrow, int mode, void *work, size_t work_bytes,
                               void *stream) {
  g_err[0] = 0;
  if (mode < 0 || mode > 3) return set_err(GSKNN_EINVAL, "%s: mode %d", fn, mode);
  return knn_dev(fn, m, n, d, k, XA, XA2, amap, XB, XB2, bmap, D, I, ldk, hrow, mode, work,
                 work_bytes, stream, false);
}

int gsknn::api_error(int code, const char *fn, const char *what, int value) {
  return set_err(code, "%s: %s (%d)", fn, what, value);
}

template <typename T>
static int knn1_dev(const char *fn, int m, int n, int d, int k, const T *XA, const int *amap,
                    const T *XB, const int *bmap, T *D, int *I, int ldk, const int *hrow,
                    void *work, size_t work_bytes, void *stream) {
  int rc = check_sizes(fn, m, n, d, k, ldk);
  if (rc) return rc;
  if (m == 0 || n == 0 || d == 0 || k == 0) return GSKNN_OK;
  if (!XA || !amap || !XB || !bmap || !D || !I) return set_err(GSKNN_EINVAL, "%s: null array", fn);
  hipStream_t s = (hipStream_t)stream;
  const size_t need = knn_l1_work_bytes(m, n, d, k, (int)sizeof(T));
  void *own = nullptr;
  if (need > 0 && (work == nullptr || work_bytes < need)) {
    if (work != nullptr)
      return set_err(GSKNN_EINVAL, "%s: workspace %zu < %zu bytes", fn, work_bytes, need);
    hipError_t e = hipMallocAsync(&own, need, s);
    if (e != hipSuccess) return hip_err(e, "hipMallocAsync(workspace)");
    work = own;
    work_bytes = need;
  }
  KnnArgs<T> a{};
  a.m = m; a.n = n; a.d = d; a.k = k; a.ldk = ldk;
  a.XA = XA; a.amap = amap;
  a.XB = XB; a.bmap = bmap;
  a.D = D; a.I = I; a.hrow = hrow;
  hipError_t e;
  if constexpr (sizeof(T) == 8)
    e = launch_knn_l1_f64(a, s, work, work_bytes);
  else
    e = launch_knn_l1_f32(a, s, work, work_bytes);
  if (own) (void)hipFreeAsync(own, s);
  if (e == hipErrorNotSupported)
    return set_err(GSKNN_ELIMIT, "%s: no kernel configuration fits d=%d k=%d", fn, d, k);
  if (e != hipSuccess) return hip_err(e, fn);
  return GSKNN_OK;
}

extern "C" size_t gsknn_dknn1_work_bytes(int m, int n, int d, int k) {
  return knn_l1_work_bytes(m, n, d, k, 8);
}

extern "C" size_t gsknn_sknn1_work_bytes(int m, int n, int d, int k) {
  return knn_l1_work_bytes(m, n, d, k, 4);
}

extern "C" int gsknn_dknn1_dev(int m, int n, int d, int k, const double *XA, const int *amap,
                               const double *XB, const int *bmap, double *D, int *I, int ldk,
                               const int *hrow, void *work, size_t work_bytes, void *stream) {
  return knn1_dev("gsknn_dknn1_dev", m, n, d, k, XA, amap, XB, bmap, D, I, ldk, hrow, work,
                  work_bytes, stream);
}

extern "C" int gsknn_sknn1_dev(int m, int n, int d, int k, const float *XA, const int *amap,
                               const float *XB, const int *bmap, float *D, int *I, int ldk,
                               const int *hrow, void *work, size_t work_bytes, void *stream) {
  return knn1_dev("gsknn_sknn1_dev", m, n, d, k, XA, amap, XB, bmap, D, I, ldk, hrow, work,
                  work_bytes, stream);
}

extern "C" int gsknn_dsqnorm_dev(int N, int d, const double *X, double *X2, void *stream) {
  if (N < 0 || d < 0) return set_err(GSKNN_EINVAL, "gsknn_dsqnorm_dev: negative size");
  hipError_t e = launch_sqnorm_f64(N, d, X, X2, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "gsknn_dsqnorm_dev");
}

extern "C" int gsknn_ssqnorm_dev(int N, int d, const float *X, float *X2, void *stream) {
  if (N < 0 || d < 0) return set_err(GSKNN_EINVAL, "gsknn_ssqnorm_dev: negative size");
  hipError_t e = launch_sqnorm_f32(N, d, X, X2, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "gsknn_ssqnorm_dev");
}

extern "C" int gsknn_rows_ascending_ddev(int n, int k, double *D, int *I, int ldk, void *stream) {
  if (n < 0 || k < 0 || (n > 0 && k > 0 && (ldk < k || !D || !I)))
    return set_err(GSKNN_EINVAL, "gsknn_rows_ascending_ddev: n=%d k=%d ldk=%d", n, k, ldk);
  if (k > kRowsAscendingMaxK)
    return set_err(GSKNN_ELIMIT, "gsknn_rows_ascending_ddev: k=%d > %d", k, kRowsAscendingMaxK);
  hipError_t e = launch_rows_ascending_f64(n, k, D, I, ldk, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "gsknn_rows_ascending_ddev");
}

extern "C" int gsknn_rows_ascending_sdev(int n, int k, float *D, int *I, int ldk, void *stream) {
  if (n < 0 || k < 0 || (n > 0 && k > 0 && (ldk < k || !D || !I)))
    return set_err(GSKNN_EINVAL, "gsknn_rows_ascending_sdev: n=%d k=%d ldk=%d", n, k, ldk);
  if (k > kRowsAscendingMaxK)
    return set_err(GSKNN_ELIMIT, "gsknn_rows_ascending_sdev: k=%d > %d", k, kRowsAscendingMaxK);
  hipError_t e = launch_rows_ascending_f32(n, k, D, I, ldk, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "gsknn_rows_ascending_sdev");
}

extern "C" int gsknn_tree_level_keys_dev(long long total, int nnodes, int d, int id_bits,
                                         const float *X, const long long *perm,
                                         const long long *starts, const long long *offs,
                                         const float *dirs, long long *pos, long long *keys,
                                         void *stream) {
  // node index, 32 order bits and the id share 63 bits
  const int node_bits = 63 - 32 - id_bits;
  if (total < 0 || nnodes < 0 || d <= 0 || id_bits <= 0 || (total > 0 && nnodes == 0) ||
      node_bits <= 0 || (nnodes > 0 && ((long long)(nnodes - 1) >> node_bits) != 0))
    return set_err(GSKNN_EINVAL, "gsknn_tree_level_keys_dev: total=%lld nnodes=%d d=%d id_bits=%d",
                   total, nnodes, d, id_bits);
  hipError_t e = launch_tree_level_keys(total, nnodes, d, id_bits, X, perm, starts, offs, dirs,
                                        pos, keys, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "gsknn_tree_level_keys_dev");
}

extern "C" size_t gsknn_bf16_prep_bytes(int N, int d) {
  return (N > 0 && d > 0) ? bf16_prep_bytes(N, d) : 0;
}

extern "C" int gsknn_bf16_prep_dev(int N, int d, const float *X, void *prep, void *stream) {
  if (N < 0 || d <= 0 || d > 256)
    return set_err(GSKNN_EINVAL, "gsknn_bf16_prep_dev: N=%d d=%d (need 0 < d <= 256)", N, d);
  if (N > 0 && (!X || !prep)) return set_err(GSKNN_EINVAL, "gsknn_bf16_prep_dev: null array");
  hipError_t e = launch_bf16_prep(N, d, X, prep, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "gsknn_bf16_prep_dev");
}

extern "C" size_t gsknn_bf16_work_bytes(int m, int n, int d, int k) {
  return (m > 0 && n > 0 && d > 0 && k > 0) ? bf16_work_bytes(m, n, d, k) : 0;
}

extern "C" size_t gsknn_bf16_count_offset(int m, int n, int d, int k, int *nsplit) {
  if (nsplit) *nsplit = 0;
  return (m > 0 && n > 0 && d > 0 && k > 0) ? bf16_count_offset(m, n, d, k, nsplit) : 0;
}

extern "C" int gsknn_bf16_knn_dev(int m, int n, int d, int k, int NA, const float *XA,
                                  const void *prepA, const int *amap, int NB, const float *XB,
                                  const void *prepB, const int *bmap, double *D, int *I, int ldk,
                                  const int *hrow, void *work, size_t work_bytes, int *overflow,
                                  void *stream) {
  const char *fn = "gsknn_bf16_knn_dev";
  int rc = check_sizes(fn, m, n, d, k, ldk);
  if (rc) return rc;
  if (m == 0 || n == 0 || d == 0 || k == 0) return GSKNN_OK;
  if (d > 256) return set_err(GSKNN_ELIMIT, "%s: d=%d > 256", fn, d);
  if (!XA || !prepA || !amap || !XB || !prepB || !bmap || !D || !I)
    return set_err(GSKNN_EINVAL, "%s: null array", fn);
  hipStream_t s = (hipStream_t)stream;
  const size_t need = bf16_work_bytes(m, n, d, k);
  if (need == 0) return set_err(GSKNN_ELIMIT, "%s: no kernel configuration fits d=%d k=%d", fn, d, k);
  void *own = nullptr;
  if (work == nullptr || work_bytes < need) {
    if (work != nullptr)
      return set_err(GSKNN_EINVAL, "%s: workspace %zu < %zu bytes", fn, work_bytes, need);
    hipError_t e = hipMallocAsync(&own, need, s);
    if (e != hipSuccess) return hip_err(e, "hipMallocAsync(workspace)");
    work = own;
    work_bytes = need;
  }
  hipError_t e = launch_knn_bf16(m, n, d, k, NA, NB, XA, prepA, amap, XB, prepB, bmap, D, I, ldk,
                                 hrow, work, work_bytes, overflow, s);
  if (own) (void)hipFreeAsync(own, s);
  if (e == hipErrorNotSupported)
    return set_err(GSKNN_ELIMIT, "%s: no kernel configuration fits d=%d k=%d", fn, d, k);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, fn);
}

extern "C" size_t gsknn_sweep_lds_bytes(int elem_bytes, int d, int k, int qpb, int ct) {
  return leaf_sweep_lds_bytes(elem_bytes, d, k, qpb, ct);
}

template <typename T, typename F>
static int sweep_dev(const char *fn, F launch, int d, int k, const T *X, const T *X2,
                     const int *perm, const int *items, const int *nitems, T *D, int *I, int ldk,
                     int qpb, int ct, int grid, void *stream) {
  if (d <= 0 || k <= 0 || ldk < k) return set_err(GSKNN_EINVAL, "%s: d=%d k=%d ldk=%d", fn, d, k, ldk);
  if (!X || !X2 || !perm || !items || !nitems || !D || !I)
    return set_err(GSKNN_EINVAL, "%s: null array", fn);
  hipError_t e = launch(d, k, X, X2, perm, items, nitems, D, I, ldk, qpb, ct, grid,
                        (hipStream_t)stream);
  if (e == hipErrorNotSupported)
    return set_err(GSKNN_ELIMIT, "%s: qpb=%d ct=%d grid=%d does not fit (d=%d k=%d)", fn, qpb, ct,
                   grid, d, k);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, fn);
}

extern "C" int gsknn_sweep_leaves_sdev(int d, int k, const float *X, const float *X2,
                                       const int *perm, const int *items, const int *nitems,
                                       float *D, int *I, int ldk, int qpb, int ct, int grid,
                                       void *stream) {
  return sweep_dev("gsknn_sweep_leaves_sdev", launch_leaf_sweep_f32, d, k, X, X2, perm, items,
                   nitems, D, I, ldk, qpb, ct, grid, stream);
}

extern "C" int gsknn_sweep_norms_sdev(int N, int d, const float *X, const float *X2, float *xn,
                                      float *mx, void *stream) {
  const char *fn = "gsknn_sweep_norms_sdev";
  if (N < 0 || d <= 0) return set_err(GSKNN_EINVAL, "%s: N=%d d=%d", fn, N, d);
  if (N > 0 && (!X || !X2 || !xn || !mx)) return set_err(GSKNN_EINVAL, "%s: null array", fn);
  hipError_t e = launch_sweep_norms(N, d, X, X2, xn, mx, (hipStream_t)stream);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, fn);
}

extern "C" int gsknn_sweep_mfma_supported(int d, int k, int qpb) {
  return leaf_sweep_mfma_ok(d, k, qpb) ? 1 : 0;
}

extern "C" int gsknn_sweep_leaves_mfma_sdev(int d, int k, const float *X, const float *X2,
                                            const float *xn, const float *mx, const int *perm,
                                            const int *items, const int *nitems, float *D, int *I,
                                            int ldk, int qpb, int grid, void *stream) {
  const char *fn = "gsknn_sweep_leaves_mfma_sdev";
  if (d <= 0 || k <= 0 || ldk < k) return set_err(GSKNN_EINVAL, "%s: d=%d k=%d ldk=%d", fn, d, k, ldk);
  if (!X || !X2 || !xn || !mx || !perm || !items || !nitems || !D || !I)
    return set_err(GSKNN_EINVAL, "%s: null array", fn);
  hipError_t e = launch_leaf_sweep_mfma(d, k, X, X2, xn, mx, perm, items, nitems, D, I, ldk, qpb,
                                        grid, (hipStream_t)stream);
  if (e == hipErrorNotSupported)
    return set_err(GSKNN_ELIMIT, "%s: d=%d k=%d qpb=%d grid=%d not supported", fn, d, k, qpb, grid);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, fn);
}

extern "C" int gsknn_sweep_leaves_ddev(int d, int k, const double *X, const double *X2,
                                       const int *perm, const int *items, const int *nitems,
                                       double *D, int *I, int ldk, int qpb, int ct, int grid,
                                       void *stream) {
  return sweep_dev("gsknn_sweep_leaves_ddev", launch_leaf_sweep_f64, d, k, X, X2, perm, items,
                   nitems, D, I, ldk, qpb, ct, grid, stream);
}

// ---------------------------------------------------------------------------
// Host-pointer staging
// ---------------------------------------------------------------------------
namespace {

bool is_device_ptr(const void *p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}

// Grow-only device buffers and a stream per device, for the host-pointer
// entry points, which serialise on g_stage_mu.  begin_call points g_buf at
// the buffers of the device selected by gsknn_set_device, so a device switch
// between calls never reuses another device's stream or memory.
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};
struct DevState {
  DevBuf buf[10];
  hipStream_t stream = nullptr;
};
std::mutex g_stage_mu;
std::map<int, DevState> g_state;
DevBuf *g_buf = nullptr;

int ensure(DevBuf &b, size_t bytes, void **out) {
  if (bytes == 0) bytes = 16;
  if (b.n < bytes) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return set_err(GSKNN_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    b.n = bytes;
  }
  *out = b.p;
  return GSKNN_OK;
}

int get_max_index(const int *map, int len, bool dev, int *mx, const char *fn) {
  std::vector<int> tmp;
  const int *h = map;
  if (dev) {
    tmp.resize((size_t)len);
    hipError_t e = hipMemcpy(tmp.data(), map, sizeof(int) * (size_t)len, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err(e, "hipMemcpy(map)");
    h = tmp.data();
  }
  int v = -1;
  for (int i = 0; i < len; ++i) {
    if (h[i] < 0) return set_err(GSKNN_EINVAL, "%s: negative index %d at %d", fn, h[i], i);
    v = std::max(v, h[i]);
  }
  *mx = v;
  return GSKNN_OK;
}

int g_mode_env = -1;
int env_mode() {
  if (g_mode_env < 0) {
    const char *s = getenv("GSKNN_MODE");
    g_mode_env = MODE_AUTO;
    if (s && !strcmp(s, "exact")) g_mode_env = MODE_EXACT;
    if (s && !strcmp(s, "mfma")) g_mode_env = MODE_MFMA;
  }
  return g_mode_env;
}

// Device views of the candidate/query inputs of one host-API call; arrays
// that already live on the device are used in place.
template <typename T>
struct Inputs {
  const T *XA, *XA2, *XB, *XB2;
  const int *amap, *bmap;
};

auto upload(DevBuf &b, const void *src, size_t bytes, void **dst, hipStream_t s) -> int {
  int r = ensure(b, bytes, dst);
  if (r) return r;
  hipError_t e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, "hipMemcpyAsync(H2D)");
}

// Caller holds g_stage_mu.  Point tables are copied up to the largest row
// the maps touch; one copy serves both roles when XA == XB.
// With norms = false (L1 calls) XA2/XB2 are neither read nor staged.
template <typename T>
int stage_inputs(const char *fn, int m, int n, int d, const T *XA, const T *XA2,
                 const int *amap, const T *XB, const T *XB2, const int *bmap, hipStream_t s,
                 Inputs<T> *out, bool norms = true) {
  if (!XA || !amap || !XB || !bmap || (norms && (!XA2 || !XB2)))
    return set_err(GSKNN_EINVAL, "%s: null array", fn);
  const bool dA = is_device_ptr(XA), dam = is_device_ptr(amap);
  const bool dB = is_device_ptr(XB), dbm = is_device_ptr(bmap);
  const bool dA2 = !norms || is_device_ptr(XA2), dB2 = !norms || is_device_ptr(XB2);
  int rc;
  int maxa = -1, maxb = -1;
  if (!dA || !dA2)
    if ((rc = get_max_index(amap, m, dam, &maxa, fn))) return rc;
  if (!dB || !dB2)
    if ((rc = get_max_index(bmap, n, dbm, &maxb, fn))) return rc;
  const bool same = (XA == XB) && (XA2 == XB2) && !dA && (!norms || !dA2);
  if (same) maxa = maxb = std::max(maxa, maxb);
  *out = Inputs<T>{XA, XA2, XB, XB2, amap, bmap};
  void *p;
  if (!dA) { if ((rc = upload(g_buf[0], XA, sizeof(T) * (size_t)(maxa + 1) * d, &p, s))) return rc; out->XA = (T *)p; }
  if (!dA2) { if ((rc = upload(g_buf[1], XA2, sizeof(T) * (size_t)(maxa + 1), &p, s))) return rc; out->XA2 = (T *)p; }
  if (!norms) out->XA2 = out->XB2 = nullptr;
  if (!dam) { if ((rc = upload(g_buf[2], amap, sizeof(int) * (size_t)m, &p, s))) return rc; out->amap = (int *)p; }
  if (same) {
    out->XB = out->XA;
    out->XB2 = norms ? out->XA2 : nullptr;
  } else {
    if (!dB) { if ((rc = upload(g_buf[3], XB, sizeof(T) * (size_t)(maxb + 1) * d, &p, s))) return rc; out->XB = (T *)p; }
    if (!dB2) { if ((rc = upload(g_buf[4], XB2, sizeof(T) * (size_t)(maxb + 1), &p, s))) return rc; out->XB2 = (T *)p; }
  }
  if (!dbm) { if ((rc = upload(g_buf[5], bmap, sizeof(int) * (size_t)n, &p, s))) return rc; out->bmap = (int *)p; }
  return GSKNN_OK;
}

// Caller holds g_stage_mu.
int begin_call(hipStream_t *s) {
  const int dev = g_device.load();
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return hip_err(e, "hipSetDevice");
  DevState &st = g_state[dev];
  if (!st.stream) {
    e = hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_err(e, "hipStreamCreate");
  }
  g_buf = st.buf;
  *s = st.stream;
  return GSKNN_OK;
}

// Heap rows in: n rows of stride ldk, host or device.
template <typename T>
int stage_rows(T *D, int *I, int n, int ldk, hipStream_t s, T **gD, int **gI, bool *dD, bool *dI) {
  *dD = is_device_ptr(D);
  *dI = is_device_ptr(I);
  *gD = D;
  *gI = I;
  const size_t ents = (size_t)n * ldk;
  void *p;
  int rc;
  if (!*dD) { if ((rc = upload(g_buf[6], D, sizeof(T) * ents, &p, s))) return rc; *gD = (T *)p; }
  if (!*dI) { if ((rc = upload(g_buf[7], I, sizeof(int) * ents, &p, s))) return rc; *gI = (int *)p; }
  return GSKNN_OK;
}

template <typename T>
int finish_rows(T *D, int *I, const T *gD, const int *gI, bool dD, bool dI, size_t ents,
                hipStream_t s, const char *fn) {
  hipError_t e;
  if (!dD && (e = hipMemcpyAsync(D, gD, sizeof(T) * ents, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return hip_err(e, "hipMemcpyAsync(D2H)");
  if (!dI && (e = hipMemcpyAsync(I, gI, sizeof(int) * ents, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return hip_err(e, "hipMemcpyAsync(D2H)");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_err(e, fn);
  return GSKNN_OK;
}

// mode >= 0: squared-2-norm kNN in that mode; kModeL1: L1 kNN.
constexpr int kModeL1 = -1;

template <typename T>
hipError_t launch_any(const KnnArgs<T> &a, int mode, hipStream_t s, void *work, size_t wb) {
  if (mode == kModeL1) {
    if constexpr (sizeof(T) == 8)
      return launch_knn_l1_f64(a, s, work, wb);
    else
      return launch_knn_l1_f32(a, s, work, wb);
  }
  if constexpr (sizeof(T) == 8)
    return launch_knn_f64(a, mode, s, work, wb);
  else
    return launch_knn_f32(a, mode, s, work, wb);
}

size_t any_work_bytes(int m, int n, int d, int k, int elem, int mode) {
  return mode == kModeL1 ? knn_l1_work_bytes(m, n, d, k, elem)
                         : knn_work_bytes(m, n, d, k, elem, mode);
}

// One staged kNN call: host or device pointers in, result in D/I (row
// stride ldk) when the call returns.  mode = kModeL1 selects the L1 norm.
template <typename T>
int knn_host(const char *fn, int m, int n, int d, int k, const T *XA, const T *XA2,
             const int *amap, const T *XB, const T *XB2, const int *bmap, T *D, int *I,
             int ldk, int mode) {
  if (m < 0 || n < 0 || d < 0 || k < 0) return set_err(GSKNN_EINVAL, "%s: negative size", fn);
  if (m == 0 || n == 0 || d == 0 || k == 0) return GSKNN_OK;  // frame/gsknn.c:909-913
  if (!D || !I) return set_err(GSKNN_EINVAL, "%s: null heap arrays", fn);
  std::lock_guard<std::mutex> lock(g_stage_mu);
  hipStream_t s;
  int rc;
  if ((rc = begin_call(&s))) return rc;
  Inputs<T> in;
  if ((rc = stage_inputs(fn, m, n, d, XA, XA2, amap, XB, XB2, bmap, s, &in, mode != kModeL1)))
    return rc;
  T *gD;
  int *gI;
  bool dD, dI;
  if ((rc = stage_rows(D, I, n, ldk, s, &gD, &gI, &dD, &dI))) return rc;
  const size_t wb = any_work_bytes(m, n, d, k, (int)sizeof(T), mode);
  void *work = nullptr;
  if (wb && (rc = ensure(g_buf[8], wb, &work))) return rc;
  KnnArgs<T> a{};
  a.m = m; a.n = n; a.d = d; a.k = k; a.ldk = ldk;
  a.XA = in.XA; a.XA2 = in.XA2; a.amap = in.amap;
  a.XB = in.XB; a.XB2 = in.XB2; a.bmap = in.bmap;
  a.D = gD; a.I = gI; a.hrow = nullptr;
  hipError_t e = launch_any(a, mode, s, work, wb);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    if (e == hipErrorNotSupported)
      return set_err(GSKNN_ELIMIT, "%s: no kernel configuration fits d=%d k=%d", fn, d, k);
    return hip_err(e, fn);
  }
  return finish_rows(D, I, gD, gI, dD, dI, (size_t)n * ldk, s, fn);
}

// dgsknn_var3 (frame/gsknn.c:597-659) on heapCreate_d rows with the header
// layout: the filled part of each row is unpacked into a plain row padded
// with (+inf, INT_MAX) placeholders, the kNN runs on the plain rows, and the
// result is packed back behind the header with the new fill count.
int var3_host(int m, int n, int d, int r, const double *XA, const double *XA2,
              const int *amap, const double *XB, const double *XB2, const int *bmap,
              heap_t *heap) {
  const char *fn = "dgsknn_var3";
  if (m < 0 || n < 0 || d < 0 || r < 0) return set_err(GSKNN_EINVAL, "%s: negative size", fn);
  if (m == 0 || n == 0 || d == 0 || r == 0) return GSKNN_OK;
  if (heap->ldk < r + KNN_HEAP_OFFSET)
    return set_err(GSKNN_EINVAL, "%s: heap ldk %d has no header room for r=%d", fn, heap->ldk, r);
  std::lock_guard<std::mutex> lock(g_stage_mu);
  hipStream_t s;
  int rc;
  if ((rc = begin_call(&s))) return rc;
  Inputs<double> in;
  if ((rc = stage_inputs(fn, m, n, d, XA, XA2, amap, XB, XB2, bmap, s, &in))) return rc;
  double *gD;
  int *gI;
  bool dD, dI;
  const int ldk = heap->ldk;
  if ((rc = stage_rows(heap->D, heap->I, n, ldk, s, &gD, &gI, &dD, &dI))) return rc;
  void *tD, *tI;
  if ((rc = ensure(g_buf[9], (size_t)n * r * (sizeof(double) + sizeof(int)) + 16, &tD))) return rc;
  tI = (char *)tD + ((size_t)n * r * sizeof(double) + 15) / 16 * 16;
  hipError_t e = launch_hdr_to_rows(n, r, ldk, gD, gI, (double *)tD, (int *)tI, s);
  const size_t wb = knn_work_bytes(m, n, d, r, 8, env_mode());
  void *work = nullptr;
  if (e == hipSuccess && wb && (rc = ensure(g_buf[8], wb, &work))) return rc;
  if (e == hipSuccess) {
    KnnArgs<double> a{};
    a.m = m; a.n = n; a.d = d; a.k = r; a.ldk = r;
    a.XA = in.XA; a.XA2 = in.XA2; a.amap = in.amap;
    a.XB = in.XB; a.XB2 = in.XB2; a.bmap = in.bmap;
    a.D = (double *)tD; a.I = (int *)tI; a.hrow = nullptr;
    e = launch_knn_f64(a, env_mode(), s, work, wb);
  }
  if (e == hipSuccess) e = launch_rows_to_hdr(n, r, ldk, (const double *)tD, (const int *)tI, gD, gI, s);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    if (e == hipErrorNotSupported)
      return set_err(GSKNN_ELIMIT, "%s: no kernel configuration fits d=%d r=%d", fn, d, r);
    return hip_err(e, fn);
  }
  return finish_rows(heap->D, heap->I, gD, gI, dD, dI, (size_t)n * ldk, s, fn);
}

// dgssq2nrm (frame/gsknn.c:481-594): the full m x n block of distances,
// C[i + j*ldc] for candidate i and query j.
int sq2nrm_host(int m, int n, int d, const double *XA, const double *XA2, const int *amap,
                const double *XB, const double *XB2, const int *bmap, double *C, int ldc) {
  const char *fn = "dgssq2nrm";
  if (m < 0 || n < 0 || d < 0) return set_err(GSKNN_EINVAL, "%s: negative size", fn);
  if (m == 0 || n == 0 || d == 0) return GSKNN_OK;
  if (!C || ldc < m) return set_err(GSKNN_EINVAL, "%s: C null or ldc %d < m %d", fn, ldc, m);
  std::lock_guard<std::mutex> lock(g_stage_mu);
  hipStream_t s;
  int rc;
  if ((rc = begin_call(&s))) return rc;
  Inputs<double> in;
  if ((rc = stage_inputs(fn, m, n, d, XA, XA2, amap, XB, XB2, bmap, s, &in))) return rc;
  const bool dC = is_device_ptr(C);
  double *gC = C;
  const size_t ents = (size_t)ldc * (n - 1) + m;
  void *p;
  if (!dC) {
    if ((rc = ensure(g_buf[6], sizeof(double) * ents, &p))) return rc;
    gC = (double *)p;
  }
  hipError_t e = launch_sq2nrm_f64(m, n, d, in.XA, in.XA2, in.amap, in.XB, in.XB2, in.bmap, gC, ldc, s);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(s);
    return hip_err(e, fn);
  }
  if (!dC) {
    // Copy column by column so entries between m and ldc keep the caller's values.
    e = hipMemcpy2DAsync(C, sizeof(double) * ldc, gC, sizeof(double) * ldc, sizeof(double) * m, n,
                         hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_err(e, "hipMemcpy2DAsync(D2H)");
  }
  e = hipStreamSynchronize(s);
  return e == hipSuccess ? GSKNN_OK : hip_err(e, fn);
}

}  // namespace

// ---------------------------------------------------------------------------
// Reference API: kNN entry points
// ---------------------------------------------------------------------------
extern "C" void dgsknn_var1(int m, int n, int k, int r, double *XA, double *XA2, int *amap,
                            double *XB, double *XB2, int *bmap, double *D, int *I) {
  // frame/gsknn.c:883-1103: A = queries (m, heap rows), B = candidates (n),
  // row stride r (:930).
  g_err[0] = 0;
  (void)knn_host("dgsknn_var1", n, m, k, r, XB, XB2, bmap, XA, XA2, amap, D, I, r, env_mode());
}

extern "C" void sgsknn_var1(int m, int n, int k, int r, float *XA, float *XA2, int *amap,
                            float *XB, float *XB2, int *bmap, float *D, int *I) {
  g_err[0] = 0;
  (void)knn_host("sgsknn_var1", n, m, k, r, XB, XB2, bmap, XA, XA2, amap, D, I, r, env_mode());
}

extern "C" void sgsknn(int m, int n, int k, int r, float *XA, float *XA2, int *alpha, float *XB,
                       float *XB2, int *beta, heap_t *heap) {
  // frame/gsknn.c:1106-1135: swap the roles and call var1 on heap->D_s.
  if (!heap) {
    set_err(GSKNN_EINVAL, "sgsknn: null heap");
    return;
  }
  sgsknn_var1(n, m, k, r, XB, XB2, beta, XA, XA2, alpha, heap->D_s, heap->I);
}

extern "C" void dgsknn(int m, int n, int k, int r, double *XA, double *XA2, int *alpha,
                       double *XB, double *XB2, int *beta, heap_t *heap) {
  // frame/gsknn.c:1138-1184
  if (!heap) {
    set_err(GSKNN_EINVAL, "dgsknn: null heap");
    return;
  }
  if (r > KNN_VAR_THRES) {
    dgsknn_var3(m, n, k, r, XA, XA2, alpha, XB, XB2, beta, heap);
  } else {
    dgsknn_var1(n, m, k, r, XB, XB2, beta, XA, XA2, alpha, heap->D, heap->I);
  }
}

// L1 (KNN_1NORM) on heap rows of stride r, the dgsknn_var1 layout the
// reference's L1 kernels run in (frame/gsknn.c:930); the reference has no L1
// kernel for the r > 512 header layout.
template <typename T>
static void knn1_heap(const char *fn, int m, int n, int k, int r, T *XA, int *alpha, T *XB,
                      int *beta, T *D, int *I) {
  g_err[0] = 0;
  if (r > KNN_VAR_THRES) {
    set_err(GSKNN_ELIMIT, "%s: r=%d > %d (the large-r path has no L1 form)", fn, r,
            KNN_VAR_THRES);
    return;
  }
  (void)knn_host<T>(fn, m, n, k, r, XA, nullptr, alpha, XB, nullptr, beta, D, I, r, kModeL1);
}

extern "C" void dgsknn_1norm(int m, int n, int k, int r, double *XA, int *alpha, double *XB,
                             int *beta, heap_t *heap) {
  if (!heap) {
    set_err(GSKNN_EINVAL, "dgsknn_1norm: null heap");
    return;
  }
  knn1_heap("dgsknn_1norm", m, n, k, r, XA, alpha, XB, beta, heap->D, heap->I);
}

extern "C" void sgsknn_1norm(int m, int n, int k, int r, float *XA, int *alpha, float *XB,
                             int *beta, heap_t *heap) {
  if (!heap) {
    set_err(GSKNN_EINVAL, "sgsknn_1norm: null heap");
    return;
  }
  knn1_heap("sgsknn_1norm", m, n, k, r, XA, alpha, XB, beta, heap->D_s, heap->I);
}

// ---------------------------------------------------------------------------
// Heap utilities (frame/gsknn_heap.c, frame/gsknn_util.c), host side.
// ---------------------------------------------------------------------------
extern "C" double *gsknn_malloc_aligned(int m, int n, int size) {
  // frame/gsknn_util.c:119-136, without the exit(1) on failure.
  void *ptr = nullptr;
  const size_t bytes = (size_t)size * (size_t)m * (size_t)n;
  if (posix_memalign(&ptr, 32, bytes ? bytes : 32) != 0) {
    set_err(GSKNN_ENOMEM, "gsknn_malloc_aligned: %zu bytes", bytes);
    return nullptr;
  }
  return (double *)ptr;
}

extern "C" void swap_float(float *x, int i, int j) { std::swap(x[i], x[j]); }
extern "C" void swap_double(double *x, int i, int j) { std::swap(x[i], x[j]); }
extern "C" void swap_int(int *I, int i, int j) { std::swap(I[i], I[j]); }

// Selection sort by ascending distance: same result as the reference's
// bubbleSort_[sd] (frame/gsknn_util.c:77-111) for distinct keys.
template <typename T>
static void sort_pairs(int n, T *D, int *I) {
  for (int i = 0; i < n; ++i)
    for (int j = n - 1; j > i; --j)
      if (D[j] < D[j - 1]) {
        std::swap(D[j], D[j - 1]);
        std::swap(I[j], I[j - 1]);
      }
}
extern "C" void bubbleSort_s(int n, float *D, int *I) { sort_pairs(n, D, I); }
extern "C" void bubbleSort_d(int n, double *D, int *I) { sort_pairs(n, D, I); }

// Max-heap sift-down by distance (frame/gsknn_heap.c:52-96).
template <typename T>
static void sift_host(T *D, int s, int n, int *I) {
  while (2 * s + 1 < n) {
    int j = 2 * s + 1;
    if (j + 1 < n && D[j] < D[j + 1]) ++j;
    if (!(D[s] < D[j])) break;
    std::swap(D[s], D[j]);
    std::swap(I[s], I[j]);
    s = j;
  }
}

extern "C" void heapSelect_d(int m, int r, double *x, int *alpha, double *D, int *I) {
  // frame/gsknn_heap.c:133-162: replace the root when x <= root.
  for (int i = 0; i < m; ++i) {
    if (x[i] > D[0]) continue;
    D[0] = x[i];
    I[0] = alpha[i];
    sift_host(D, 0, r, I);
  }
}

extern "C" void heapSelect_s(int m, int r, float *x, int *alpha, float *D, int *I) {
  for (int i = 0; i < m; ++i) {
    if (x[i] > D[0]) continue;
    D[0] = x[i];
    I[0] = alpha[i];
    sift_host(D, 0, r, I);
  }
}

extern "C" heap_t *heapAttach_s(int m, int k, float *D, int *I) {
  // frame/gsknn_heap.c:169-186
  heap_t *h = (heap_t *)calloc(1, sizeof(heap_t));
  if (!h) return nullptr;
  h->m = m; h->k = k; h->d = 2; h->ro_s = 0.0f; h->ldk = k;
  h->D_s = D; h->I = I; h->type = KNN_2NORM; h->prec = KNN_SINGLE;
  return h;
}

extern "C" heap_t *heapAttach_d(int m, int k, double *D, int *I) {
  // frame/gsknn_heap.c:188-207
  heap_t *h = (heap_t *)calloc(1, sizeof(heap_t));
  if (!h) return nullptr;
  h->m = m; h->k = k; h->d = 2; h->ro = 0.0; h->ldk = k;
  h->D = D; h->I = I; h->type = KNN_2NORM; h->prec = KNN_DOUBLE;
  return h;
}

extern "C" heap_t *heapCreate_s(int m, int k, float ro) {
  // frame/gsknn_heap.c:215-245
  heap_t *h = (heap_t *)calloc(1, sizeof(heap_t));
  if (!h) return nullptr;
  h->m = m; h->k = k; h->d = 2; h->ro_s = ro; h->ldk = k;
  h->type = KNN_2NORM; h->prec = KNN_SINGLE;
  h->D_s = (float *)malloc(sizeof(float) * (size_t)k * m + 4);
  h->I = (int *)malloc(sizeof(int) * (size_t)k * m + 4);
  if (!h->D_s || !h->I) {
    free(h->D_s); free(h->I); free(h);
    set_err(GSKNN_ENOMEM, "heapCreate_s");
    return nullptr;
  }
  for (size_t t = 0; t < (size_t)k * m; ++t) {
    h->D_s[t] = ro;
    h->I[t] = -1;
  }
  return h;
}

extern "C" heap_t *heapCreate_d(int m, int k, double ro) {
  // frame/gsknn_heap.c:247-319: k > 512 rows get a 3-slot header
  // (D[0] = ro, I[0] = k, I[1] = fill count, I[2] = ldk) before a 4-ary heap.
  heap_t *h = (heap_t *)calloc(1, sizeof(heap_t));
  if (!h) return nullptr;
  int ldk;
  if (k > KNN_VAR_THRES) {
    ldk = ((k + KNN_HEAP_OFFSET - 1) / 4 + 1) * 4;
    h->d = 4;
  } else {
    ldk = k;
    h->d = 2;
  }
  h->m = m; h->k = k; h->ro = ro; h->ldk = ldk;
  h->type = KNN_2NORM; h->prec = KNN_DOUBLE;
  h->D = gsknn_malloc_aligned(ldk, m, sizeof(double));
  h->I = (int *)gsknn_malloc_aligned(ldk, m, sizeof(int));
  if (!h->D || !h->I) {
    free(h->D); free(h->I); free(h);
    return nullptr;
  }
  for (int i = 0; i < m; ++i) {
    double *Dr = h->D + (size_t)i * ldk;
    int *Ir = h->I + (size_t)i * ldk;
    if (k > KNN_VAR_THRES) {
      Dr[0] = ro; Dr[1] = 0.0; Dr[2] = 0.0;
      Ir[0] = k; Ir[1] = 0; Ir[2] = ldk;
      for (int j = 0; j < k; ++j) {
        Dr[KNN_HEAP_OFFSET + j] = ro;
        Ir[KNN_HEAP_OFFSET + j] = -1;
      }
    } else {
      for (int j = 0; j < k; ++j) {
        Dr[j] = ro;
        Ir[j] = -1;
      }
    }
  }
  return h;
}

extern "C" void heapFree_s(heap_t *heap) {
  // frame/gsknn_heap.c:321-328 (frees D_s and I, also for attached heaps)
  if (!heap) return;
  free(heap->D_s);
  free(heap->I);
  free(heap);
}

extern "C" void heapFree_d(heap_t *heap) {
  if (!heap) return;
  free(heap->D);
  free(heap->I);
  free(heap);
}

// ---------------------------------------------------------------------------
// Large-r path and the distance-matrix routine
// ---------------------------------------------------------------------------
extern "C" void dgsknn_var3(int m, int n, int k, int r, double *XA, double *XA2, int *amap,
                            double *XB, double *XB2, int *bmap, heap_t *heap) {
  g_err[0] = 0;
  if (!heap) {
    set_err(GSKNN_EINVAL, "dgsknn_var3: null heap");
    return;
  }
  (void)var3_host(m, n, k, r, XA, XA2, amap, XB, XB2, bmap, heap);
}

extern "C" void dgssq2nrm(int m, int n, int k, double *XA, double *XA2, int *amap, double *XB,
                          double *XB2, int *bmap, double *C, int ldc) {
  g_err[0] = 0;
  (void)sq2nrm_host(m, n, k, XA, XA2, amap, XB, XB2, bmap, C, ldc);
}
