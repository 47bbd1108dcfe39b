// Shared declarations of the HIP exact-kNN library (device and host side).
//
// Terminology follows the GSKNN C API (ChenhanYu/rnn include/gsknn.h:90-116):
// m candidates gathered from table XA through amap, n queries gathered from
// table XB through bmap; heap row j (stride ldk) belongs to query bmap[j] and
// stores distances D and candidate indices I = amap values.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gsknn {

// Pass threshold of the FP32 and BF16 candidate filters.  A pair whose MFMA
// value t = fl(-2 dot~ + |c|^2) satisfies |(t + nq) - dist| <= delta <= dmx
// for the distance dist it brackets can reach the bound U only if
// t <= (U - nq) + dmx in exact arithmetic; the FP32 evaluation is raised by
// 2^-18 (|U| + |nq| + dmx), which covers its own three roundings and those
// of the pair's bounds (2^-20 of |U| + 3 dmx), so every such pair passes.
__device__ __forceinline__ float f32_pass_thr(float U, float nq, float dmx) {
  return __builtin_fmaf(0x1p-18f, fabsf(U) + fabsf(nq) + dmx, (U - nq) + dmx);
}

// v[bb] for a per-lane bb in [0, N): a tree of selects on the bits of bb
// (N - 1 selects).  The values must be scalars (FP32 results): a select
// between two elements of one vector would become an element at a run-time
// index, a chain of compares per level.
template <int LO, int N, int NV>
__device__ __forceinline__ float select_tree(const float (&v)[NV], int bb) {
  if constexpr (N == 1) {
    return v[LO];
  } else {
    constexpr int H = N / 2;
    const float lo = select_tree<LO, H, NV>(v, bb);
    const float hi = select_tree<LO + H, H, NV>(v, bb);
    return (bb & H) ? hi : lo;
  }
}

// Kernel arguments of one kNN launch.  All pointers are device pointers.
template <typename T>
struct KnnArgs {
  int m, n, d, k, ldk;
  const T *XA, *XA2;
  const int *amap;
  const T *XB, *XB2;
  const int *bmap;
  T *D;
  int *I;
  const int *hrow;  // heap row of query j (nullptr: row j)
  int nsplit;       // candidate splits (grid.y); > 1 writes partial lists
  int cps;          // candidates per split
  T *partD;         // ascending partial lists (nsplit > 1): [nsplit][k][n], or
  int *partI;       // [n][nsplit][k] when part_qmajor
  int part_qmajor;  // partial lists query-major (read by the wave merge)
  const T *SN;      // FP64 MFMA: candidate position i's norm XA2[amap[i]]
                    // shifted down by the MFMA error margin (mfma_shift_norm,
                    // from the row's squared norm computed by the library)
  // FP64 MFMA two-stage form (unsplit launches): the streaming kernel keeps
  // the k2 = k + slack smallest candidates of each query under their MFMA
  // key, [n][k2] in hk/hid (heap order, root first), and the finalize kernel
  // forms their oracle distances.
  int k2;
  T *hk;
  int *hid;
  // Launches over a device-side query list: when set, the query count is
  // *ncount (a.n only bounds it), read by the kernel.
  const int *ncount;
};

enum Mode {
  MODE_AUTO = 0,   // FP64: MFMA when d >= 32 or the call is large (auto_mfma), else exact VALU
  MODE_EXACT = 1,  // VALU, the reference oracle's arithmetic (bit-exact)
  MODE_MFMA = 2,   // FP64 MFMA contraction, the direct form wherever it fits
  MODE_MFMA_TILE = 3,  // FP64 MFMA, the LDS-tile kernel only (FP32: as MODE_MFMA)
};

// Launchers (gsknn_kernels.hip).  Return hipSuccess or the launch error.
hipError_t launch_knn_f64(const KnnArgs<double> &a, int mode, hipStream_t s,
                          void *work, size_t work_bytes);
hipError_t launch_knn_f32(const KnnArgs<float> &a, int mode, hipStream_t s,
                          void *work, size_t work_bytes);
// Workspace bytes the launch above needs (shifted norms, partial lists of
// split launches, the two-stage form's key heaps), and the least it can run
// with (the FP64 MFMA per-wave forms then run in one stage).
size_t knn_work_bytes(int m, int n, int d, int k, int elem_bytes, int mode);
size_t knn_work_bytes_min(int m, int n, int d, int k, int elem_bytes, int mode);
// Byte offset in an FP64 call's workspace of the two-stage form's fallback
// count ((size_t)-1 without that form).
size_t knn_two_stage_count_offset(int m, int n, int d, int k, int mode);

// FP64 MFMA operand-ring kernel (gsknn_ring.hip), d <= 256: queries per
// workgroup, candidates per tile, LDS bytes, workgroups per CU that its
// registers allow (0 when the instantiation spills; vec = 16-byte row
// loads), and the launch (a.nsplit/a.cps set, grid gx x nsplit).
bool ring_supported(int d);
int ring_queries_per_block(int d);
int ring_candidates_per_tile(int d);
size_t ring_lds_bytes(int d, int k, bool two);
int ring_reg_blocks(int d, bool vec);
int ring_terms(int d);  // MFMA terms per dot product (8 x pairs)
hipError_t launch_knn_ring(const KnnArgs<double> &a, bool vec, int gx, size_t lds, hipStream_t s,
                           bool two);

// Merge of split partial lists into the heap rows (gsknn_kernels.hip);
// supported when a workgroup's heaps fit in LDS.
bool knn_merge_supported(int k, int elem_bytes);
// Whether a split launch's lists are merged one wave per query; its partial
// lists are then laid out query-major (KnnArgs::part_qmajor).
bool knn_merge_query_major(int n, int k, int nsplit);
hipError_t launch_knn_merge(const KnnArgs<double> &a, hipStream_t s);
hipError_t launch_knn_merge(const KnnArgs<float> &a, hipStream_t s);

// L1-norm (KNN_1NORM) kNN (gsknn_l1.hip); XA2/XB2 are not read.
size_t knn_l1_work_bytes(int m, int n, int d, int k, int elem_bytes);
hipError_t launch_knn_l1_f64(const KnnArgs<double> &a, hipStream_t s, void *work,
                             size_t work_bytes);
hipError_t launch_knn_l1_f32(const KnnArgs<float> &a, hipStream_t s, void *work,
                             size_t work_bytes);

// dgssq2nrm: C[i + j*ldc] for candidate i < m, query j < n.
hipError_t launch_sq2nrm_f64(int m, int n, int d, const double *XA, const double *XA2,
                             const int *amap, const double *XB, const double *XB2,
                             const int *bmap, double *C, int ldc, hipStream_t s);
// Header-layout rows (k > 512, 3-slot header + 4-ary heap) <-> plain rows.
hipError_t launch_hdr_to_rows(int n, int k, int ldk, const double *Dh, const int *Ih,
                              double *Dt, int *It, hipStream_t s);
hipError_t launch_rows_to_hdr(int n, int k, int ldk, const double *Dt, const int *It,
                              double *Dh, int *Ih, hipStream_t s);

// Rows to ascending (distance, index) order in place (k <= 4096).
constexpr int kRowsAscendingMaxK = 4096;
hipError_t launch_rows_ascending_f64(int n, int k, double *D, int *I, int ldk, hipStream_t s);
hipError_t launch_rows_ascending_f32(int n, int k, float *D, int *I, int ldk, hipStream_t s);

// FP32 MFMA candidate pass + re-rank in the oracle's FP32 arithmetic
// (gsknn_f32.hip).
size_t f32_work_bytes(int m, int n, int d, int k);
// Byte offset in that workspace of the count of queries whose candidate list
// overflowed and were re-scanned ((size_t)-1 without an FP32 MFMA plan).
size_t f32_overflow_offset(int m, int n, int d, int k);
bool f32_mfma_supported(int d, int k);
// Whether an FP32 call (gsknn_kernels.hip) takes that form.
bool knn_f32_takes_mfma(int m, int d, int k, int mode);
hipError_t launch_knn_f32_mfma(const KnnArgs<float> &a, hipStream_t s, void *work,
                               size_t work_bytes, int *overflow_out);

// Leaf sweep of a tree iteration (gsknn_sweep.hip).
hipError_t launch_leaf_sweep_f32(int d, int k, const float *X, const float *X2, const int *perm,
                                 const int *items, const int *nitems, float *D, int *I, int ldk,
                                 int qpb, int ct, int grid, hipStream_t s);
hipError_t launch_leaf_sweep_f64(int d, int k, const double *X, const double *X2, const int *perm,
                                 const int *items, const int *nitems, double *D, int *I, int ldk,
                                 int qpb, int ct, int grid, hipStream_t s);
size_t leaf_sweep_lds_bytes(int elem, int d, int k, int qpb, int ct);
// FP32 MFMA form of the leaf sweep (gsknn_sweep.hip): point norms, support
// test and launch.
hipError_t launch_sweep_norms(int N, int d, const float *X, const float *X2, float *xn, float *mx,
                              hipStream_t s);
bool leaf_sweep_mfma_ok(int d, int k, int qpb);
hipError_t launch_leaf_sweep_mfma(int d, int k, const float *X, const float *X2, const float *xn,
                                  const float *mx, const int *perm, const int *items,
                                  const int *nitems, float *D, int *I, int ldk, int qpb, int grid,
                                  hipStream_t s);
// One tree level's positions and sort keys (gsknn_sweep.hip).
hipError_t launch_tree_level_keys(long long total, int nnodes, int d, int id_bits,
                                  const float *X, const long long *perm,
                                  const long long *starts, const long long *offs,
                                  const float *dirs, long long *pos, long long *keys,
                                  hipStream_t s);

// BF16 candidate pass + FP64 re-rank (gsknn_rerank.hip).
size_t bf16_prep_bytes(int N, int d);
hipError_t launch_bf16_prep(int N, int d, const float *X, void *prep, hipStream_t s);
size_t bf16_work_bytes(int m, int n, int d, int k);
// Byte offset in the workspace of the [nsplit][n] candidate counts.
size_t bf16_count_offset(int m, int n, int d, int k, int *nsplit);
hipError_t launch_knn_bf16(int m, int n, int d, int k, int NA, int NB, const float *XA,
                           const void *prepA, const int *amap, const float *XB, const void *prepB,
                           const int *bmap, double *D, int *I, int ldk, const int *hrow,
                           void *work, size_t work_bytes, int *overflow_out, hipStream_t s);

// gsknn_dknn_dev with a fixed workspace and no allocation of any kind: a
// call needing more than work_bytes fails with GSKNN_ELIMIT (gsknn_api.cpp).
int dknn_dev_fixed_work(const char *fn, int m, int n, int d, int k, const double *XA,
                        const double *XA2, const int *amap, const double *XB, const double *XB2,
                        const int *bmap, double *D, int *I, int ldk, const int *hrow, int mode,
                        void *work, size_t work_bytes, void *stream);
// Sets gsknn_last_error() to "fn: what (value)" and returns code.
int api_error(int code, const char *fn, const char *what, int value);

hipError_t launch_sqnorm_f64(int N, int d, const double *X, double *X2, hipStream_t s);
hipError_t launch_sqnorm_f32(int N, int d, const float *X, float *X2, hipStream_t s);

}  // namespace gsknn
