// Batched all-nearest-neighbour sweep over the leaves of a randomized-
// projection tree: one persistent launch per tree iteration.
//
// No counterpart in the reference; each leaf is the reference's gsknn call
// with amap = bmap = the leaf's point ids and heap rows = the same ids
// (merging into global heap rows, frame/gsknn.c:1138-1184 semantics), so the
// result equals the composition of per-leaf dgsknn_ref/sgsknn_ref calls.
//
// Work list: items[i] = {leaf start in perm, leaf size, first query of the
// block, 0}; the item count is read from device memory so that a launch
// captured once in a hipGraph serves every iteration.  Each workgroup walks
// the list with stride gridDim.x (items sorted by leaf size, largest first,
// balance the persistent grid).  Within an item the exact block routine runs
// one thread per query with the oracle's arithmetic (the scalar-candidate
// form for d in {8, 16, 32, 64}), so distances are bit-identical to the
// oracle's.  Leaves of one iteration are disjoint, so no
// heap row is touched by two items of the same launch.
#include "gsknn_device.h"

namespace gsknn {

// D > 0: d == D in {8, 16, 32, 64}, the scalar-candidate block routine
// (CT unused); D == 0: the LDS-tile routine with DQ.
template <typename T, int QPB, int DQ, int D>
__global__ __launch_bounds__(QPB) void knn_leaf_sweep_kernel(KnnArgs<T> base, const int4 *items,
                                                             const int *nitems, int CT) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int count = *nitems;
  for (int it = blockIdx.x; it < count; it += gridDim.x) {
    const int4 w = items[it];
    KnnArgs<T> a = base;
    a.m = w.y;
    a.n = w.y;
    a.amap = base.amap + w.x;
    a.bmap = base.amap + w.x;
    a.hrow = base.amap + w.x;
    a.nsplit = 1;
    a.cps = w.y;
    if constexpr (D > 0)
      exact_block_s<T, QPB, D>(a, w.z / QPB, 0, smem);
    else
      exact_block<T, QPB, DQ>(a, CT, w.z / QPB, 0, smem);
    __syncthreads();  // LDS is reused by the next item
  }
}

// One level of the randomized-projection tree: for element i of the nodes
// being split (node n holds perm[starts[n], starts[n] + sizes[n]) and its
// elements are numbered from offs[n]), the position in perm and the sort key
// (n, projection order, point id); sorting the keys orders every node by
// (projection, id), ties by id, so the median split is deterministic.  The
// projection is sum_p x_p dir_p in p order, products and sums rounded
// separately; its FP32 bits map to an unsigned order (negative values
// reversed).  Blocks of 256 consecutive elements: one thread finds the node
// of the block's first element by binary search and the others step forward
// from it (a block spans few nodes); rows are read 16 bytes at a time when d
// is a multiple of 4 and the tables are 16-byte aligned.
__global__ __launch_bounds__(256) void tree_level_keys_kernel(
    long long total, int nnodes, int d, int id_bits, const float *X, const long long *perm,
    const long long *starts, const long long *offs, const float *dirs, long long *pos_out,
    long long *key_out) {
  __shared__ int base_node;
  const long long i0 = (long long)blockIdx.x * blockDim.x;
  const long long i = i0 + threadIdx.x;
  if (threadIdx.x == 0) {
    int lo = 0, hi = nnodes - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offs[mid] <= i0) lo = mid;
      else hi = mid - 1;
    }
    base_node = lo;
  }
  __syncthreads();
  if (i >= total) return;
  int lo = base_node;
  while (lo + 1 < nnodes && offs[lo + 1] <= i) ++lo;
  const long long pos = starts[lo] + (i - offs[lo]);
  const long long pt = perm[pos];
  const float *x = X + pt * d;
  const float *dv = dirs + (size_t)lo * d;
  float s = 0.0f;
  if ((d & 3) == 0 && (((uintptr_t)X | (uintptr_t)dirs) & 15) == 0) {
    const float4 *x4 = (const float4 *)x;
    const float4 *d4 = (const float4 *)dv;
    for (int p = 0; p < d / 4; ++p) {
      const float4 a = x4[p], w = d4[p];
      s = __fadd_rn(s, __fmul_rn(a.x, w.x));
      s = __fadd_rn(s, __fmul_rn(a.y, w.y));
      s = __fadd_rn(s, __fmul_rn(a.z, w.z));
      s = __fadd_rn(s, __fmul_rn(a.w, w.w));
    }
  } else {
    for (int p = 0; p < d; ++p) s = __fadd_rn(s, __fmul_rn(x[p], dv[p]));
  }
  int b = __float_as_int(s);
  if (b < 0) b ^= 0x7FFFFFFF;
  const long long ord = (long long)b + (1LL << 31);
  pos_out[i] = pos;
  key_out[i] = ((long long)lo << (32 + id_bits)) | (ord << id_bits) | pt;
}

hipError_t launch_tree_level_keys(long long total, int nnodes, int d, int id_bits,
                                  const float *X, const long long *perm,
                                  const long long *starts, const long long *offs,
                                  const float *dirs, long long *pos, long long *keys,
                                  hipStream_t s) {
  if (total <= 0) return hipSuccess;
  const long long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(tree_level_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, s, total,
                     nnodes, d, id_bits, X, perm, starts, offs, dirs, pos, keys);
  return hipGetLastError();
}

template <typename T>
size_t leaf_sweep_lds(int d, int k, int qpb, int ct) {
  if (scalar_dim(d)) return ExactScalarLayout<T>(k, qpb).bytes;
  return ExactLayout<T>(k, qpb, d, ct, d <= kExactDQ).bytes;
}

template <typename T>
hipError_t launch_leaf_sweep_t(int d, int k, const T *X, const T *X2, const int *perm,
                               const int *items, const int *nitems, T *D, int *I, int ldk,
                               int qpb, int ct, int grid, hipStream_t s) {
  KnnArgs<T> a{};
  a.d = d; a.k = k; a.ldk = ldk;
  a.XA = X; a.XA2 = X2; a.XB = X; a.XB2 = X2;
  a.amap = perm; a.bmap = perm;
  a.D = D; a.I = I;
  a.nsplit = 1;
  const size_t lds = leaf_sweep_lds<T>(d, k, qpb, ct);
  if (lds > 160 * 1024 || ct < 4 || (ct & 3) || grid < 1) return hipErrorNotSupported;
  hipError_t e = hipSuccess;
#define GSKNN_LEAF_CASE(QPB_)                                                             \
  if (qpb == QPB_) {                                                                      \
    auto kern = d == 8    ? knn_leaf_sweep_kernel<T, QPB_, 0, 8>                          \
                : d == 16 ? knn_leaf_sweep_kernel<T, QPB_, 0, 16>                         \
                : d == 32 ? knn_leaf_sweep_kernel<T, QPB_, 0, 32>                         \
                : d == 64 ? knn_leaf_sweep_kernel<T, QPB_, 0, 64>                         \
                : d <= kExactDQ ? knn_leaf_sweep_kernel<T, QPB_, kExactDQ, 0>             \
                                : knn_leaf_sweep_kernel<T, QPB_, 0, 0>;                   \
    e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            (int)lds);                                                    \
    if (e != hipSuccess) return e;                                                        \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(QPB_), lds, s, a, (const int4 *)items, nitems, ct); \
    return hipGetLastError();                                                             \
  }
  GSKNN_LEAF_CASE(64)
  GSKNN_LEAF_CASE(32)
  GSKNN_LEAF_CASE(16)
#undef GSKNN_LEAF_CASE
  return hipErrorNotSupported;
}

hipError_t launch_leaf_sweep_f32(int d, int k, const float *X, const float *X2, const int *perm,
                                 const int *items, const int *nitems, float *D, int *I, int ldk,
                                 int qpb, int ct, int grid, hipStream_t s) {
  return launch_leaf_sweep_t<float>(d, k, X, X2, perm, items, nitems, D, I, ldk, qpb, ct, grid, s);
}

hipError_t launch_leaf_sweep_f64(int d, int k, const double *X, const double *X2, const int *perm,
                                 const int *items, const int *nitems, double *D, int *I, int ldk,
                                 int qpb, int ct, int grid, hipStream_t s) {
  return launch_leaf_sweep_t<double>(d, k, X, X2, perm, items, nitems, D, I, ldk, qpb, ct, grid, s);
}

size_t leaf_sweep_lds_bytes(int elem, int d, int k, int qpb, int ct) {
  return elem == 8 ? leaf_sweep_lds<double>(d, k, qpb, ct) : leaf_sweep_lds<float>(d, k, qpb, ct);
}

}  // namespace gsknn
