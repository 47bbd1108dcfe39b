/*
 * gsknn_hip.h -- extensions of libgsknn.so beyond the reference C API.
 *
 * None of these has a counterpart in ChenhanYu/rnn; they give callers that
 * keep their data on the GPU a stream-ordered form of the same operations,
 * plus error reporting (the reference returns void and exits on failure,
 * frame/gsknn_util.c:128-133).
 *
 * Conventions: every pointer passed to a *_dev function is a device pointer
 * (hipMalloc); `stream` is a hipStream_t passed as void* (NULL = the null
 * stream); the call only enqueues work and returns 0, or a nonzero error
 * code with gsknn_last_error() set.  Naming of m/n/amap/bmap/D/I is that of
 * dgsknn in gsknn.h: m candidates (XA, amap), n queries (XB, bmap), heap row
 * j of stride ldk belongs to query bmap[j] -- or to row hrow[j] when a heap
 * row map is given.  d is the dimension and k the neighbour count.
 */
#ifndef GSKNN_MI355X_GSKNN_HIP_H
#define GSKNN_MI355X_GSKNN_HIP_H

#include <stddef.h>

#include "gsknn.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Contraction selection.  Every mode returns the reference oracle's
 * distance bits and the canonical (distance, lower index) neighbour set.
 * FP64: AUTO = MFMA when d >= 32, and below it (d >= 4) for calls large
 * enough that MFMA measured faster (calls with m <= 2048 and k <= 64 stay
 * exact; d = 8 and 16 from 2^26 pairs, other d from 2^22); EXACT = VALU with the
 * oracle's arithmetic; MFMA = FP64 MFMA wherever a per-wave form fits: for
 * d <= 64 the form that streams candidates into registers per wave, above
 * it the operand-ring form (query operands in registers up to d = 256,
 * streamed beyond), in two stages when the launch is not split (key heaps,
 * then oracle distances of their entries); AUTO takes the direct form only
 * when it keeps at least as many queries resident as the LDS-tile kernel;
 * MFMA_TILE = FP64 MFMA with the LDS-tile kernel only (d <= 256; exact
 * VALU above).  The MFMA value only filters: every pair that can reach a
 * row's k-th distance under a rigorous error margin has its distance formed
 * in the oracle's arithmetic.
 * FP32: AUTO/MFMA = FP32 MFMA candidate pass + re-rank in the oracle's FP32
 * arithmetic when d >= 8 and m >= 64, else the exact VALU kernel; EXACT =
 * always the VALU kernel. */
#define GSKNN_MODE_AUTO 0
#define GSKNN_MODE_EXACT 1
#define GSKNN_MODE_MFMA 2
#define GSKNN_MODE_MFMA_TILE 3

/* Error codes. */
#define GSKNN_OK 0
#define GSKNN_EINVAL 1   /* bad argument */
#define GSKNN_EHIP 2     /* HIP runtime error */
#define GSKNN_ENOMEM 3   /* device allocation failed */
#define GSKNN_ELIMIT 4   /* size beyond what the kernels support */

/* Message of the last failure on this thread ("" if none). */
const char *gsknn_last_error(void);

/* Device used by the host-pointer API of gsknn.h (default 0). */
int gsknn_set_device(int device);

/* FP64 / FP32 kNN on device arrays (dgsknn semantics, canonical ties).
 * `work` must hold gsknn_[ds]knn_work_bytes(...) bytes, or be NULL, in which
 * case the call allocates stream-ordered scratch itself. */
size_t gsknn_dknn_work_bytes(int m, int n, int d, int k, int mode);
int gsknn_dknn_dev(int m, int n, int d, int k, const double *XA,
                   const double *XA2, const int *amap, const double *XB,
                   const double *XB2, const int *bmap, double *D, int *I,
                   int ldk, const int *hrow, int mode, void *work,
                   size_t work_bytes, void *stream);

/* gsknn_dknn_noalloc_dev: gsknn_dknn_dev without the work arguments, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The call allocates nothing: the workspace a launch needs (partial lists
 * when the candidates are split over workgroups, i.e.
 * gsknn_dknn_work_bytes(m, n, d, k, mode) bytes) comes from a fixed device
 * array of GSKNN_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; a call needing more fails with GSKNN_ELIMIT.  Results equal
 * gsknn_dknn_dev's.  All calls on one GPU share its array: calls that
 * overlap on one GPU (other streams or threads) race and must be
 * serialised by the caller. */
#define GSKNN_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int gsknn_dknn_noalloc_dev(int m, int n, int d, int k, const double *XA,
                           const double *XA2, const int *amap,
                           const double *XB, const double *XB2,
                           const int *bmap, double *D, int *I, int ldk,
                           const int *hrow, int mode, void *stream);

size_t gsknn_sknn_work_bytes(int m, int n, int d, int k, int mode);
int gsknn_sknn_dev(int m, int n, int d, int k, const float *XA,
                   const float *XA2, const int *amap, const float *XB,
                   const float *XB2, const int *bmap, float *D, int *I,
                   int ldk, const int *hrow, int mode, void *work,
                   size_t work_bytes, void *stream);

/* For reporting (no counterpart in the reference): byte offset in the
 * workspace of a gsknn_sknn_dev call (mode and sizes as there) of one int,
 * the number of queries whose FP32 MFMA candidate list overflowed its
 * capacity and were re-scanned against every candidate in the oracle's FP32
 * arithmetic; (size_t)-1 when such a call does not take the FP32 MFMA form
 * (d < 8, m < 64, the exact mode, or heaps too large for it). */
size_t gsknn_sknn_overflow_offset(int m, int n, int d, int k, int mode);

/* For reporting (no counterpart in the reference): byte offset in the
 * workspace of a gsknn_dknn_dev call (mode and sizes as there, workspace of
 * gsknn_dknn_work_bytes bytes) of one int, the number of queries the FP64
 * MFMA two-stage form could not settle from its key heaps (more than the
 * slack of candidates within the MFMA error margin of the k-th distance)
 * and re-ran in one stage; (size_t)-1 when such a call does not take the
 * two-stage form (exact mode, candidate splits, LDS-tile kernel). */
size_t gsknn_dknn_fallback_offset(int m, int n, int d, int k, int mode);

/* Exact kNN from a BF16 MFMA candidate pass and an FP64 re-rank.
 *
 * FP32 point tables XA (NA rows) and XB (NB rows) are first prepared once
 * each with gsknn_bf16_prep_dev into a buffer of gsknn_bf16_prep_bytes(N, d)
 * bytes (BF16 copy, FP64 squared norms, rounding-error norms).  The kNN call
 * then bounds every BF16 distance by a rigorous error margin, keeps the
 * candidates whose lower bound can still reach a row's k-th distance, and
 * re-ranks them with FP64 direct differences sum_p (a_p - b_p)^2, merging
 * into FP64 heap rows (D, I; stride ldk; optional row map hrow) under the
 * canonical (distance, index) order.  The k-sets equal those of FP64
 * distances.  `overflow` (device int, may be NULL) receives the number of
 * queries whose candidate list overflowed and were re-scanned exactly.
 * d <= 256. */
size_t gsknn_bf16_prep_bytes(int N, int d);
int gsknn_bf16_prep_dev(int N, int d, const float *X, void *prep, void *stream);
size_t gsknn_bf16_work_bytes(int m, int n, int d, int k);
int gsknn_bf16_knn_dev(int m, int n, int d, int k, int NA, const float *XA,
                       const void *prepA, const int *amap, int NB,
                       const float *XB, const void *prepB, const int *bmap,
                       double *D, int *I, int ldk, const int *hrow,
                       void *work, size_t work_bytes, int *overflow,
                       void *stream);

/* Candidate counts of the last gsknn_bf16_knn_dev call with this `work`
 * (no counterpart in the reference; for reporting): at the returned byte
 * offset, *nsplit x n ints, the number of candidates each query kept in each
 * candidate split.  The re-rank reads up to the list capacity of them and
 * re-scans a query whose count exceeded it. */
size_t gsknn_bf16_count_offset(int m, int n, int d, int k, int *nsplit);

/* Leaf sweep of one randomized-projection-tree iteration, as one persistent
 * launch.  perm holds the point ids grouped by leaf; items is an int array
 * [nitems][4] = {leaf start in perm, leaf size, first query of the block, 0}
 * with the first query a multiple of qpb, and *nitems (device memory) is the
 * number of items, so one launch captured in a hipGraph can be replayed for
 * every iteration.  For each item the queries perm[start + q0 .. + qpb) are
 * matched against the candidates perm[start .. start + size) (gsknn with
 * amap = bmap = the leaf), merging into heap row = point id (stride ldk).
 * qpb in {16, 32, 64}; ct (candidates per LDS tile) a multiple of 4, unused
 * for d in {8, 16, 32, 64}, where candidates reach the ALUs through scalar
 * loads and LDS holds only the heaps; grid = persistent workgroups.
 * Distances use the reference oracle's arithmetic. */
size_t gsknn_sweep_lds_bytes(int elem_bytes, int d, int k, int qpb, int ct);
int gsknn_sweep_leaves_sdev(int d, int k, const float *X, const float *X2,
                            const int *perm, const int *items,
                            const int *nitems, float *D, int *I, int ldk,
                            int qpb, int ct, int grid, void *stream);
int gsknn_sweep_leaves_ddev(int d, int k, const double *X, const double *X2,
                            const int *perm, const int *items,
                            const int *nitems, double *D, int *I, int ldk,
                            int qpb, int ct, int grid, void *stream);

/* FP32 MFMA form of the same sweep (no counterpart in the reference): the
 * same work list, heap rows and results (distances in the reference oracle's
 * arithmetic, bit for bit); v_mfma_f32_16x16x4f32 only filters, under a
 * rigorous error bound, which pairs get their oracle distance formed.
 * gsknn_sweep_norms_sdev computes once per table xn[i] = |X[i]| (rounded
 * up) and mx = {max xn, max |X2|}, which the bound needs.  Supported shapes:
 * gsknn_sweep_mfma_supported(d, k, qpb) (d in {16, 32}, k <= 64, qpb = 32);
 * grid = persistent workgroups of one wave. */
int gsknn_sweep_norms_sdev(int N, int d, const float *X, const float *X2, float *xn,
                           float *mx, void *stream);
int gsknn_sweep_mfma_supported(int d, int k, int qpb);
int gsknn_sweep_leaves_mfma_sdev(int d, int k, const float *X, const float *X2,
                                 const float *xn, const float *mx, const int *perm,
                                 const int *items, const int *nitems, float *D,
                                 int *I, int ldk, int qpb, int grid, void *stream);

/* Sorted output: the n heap rows (row j at D + j*ldk, I + j*ldk, k <= 4096)
 * rewritten in place in ascending (distance, index) order, e.g. after a
 * gsknn call.  No entry point of the reference library does this; its tree
 * codes extract each heap to ascending (distance, index) order on the host
 * (frame/old/dgsrnn_directKQuery.cpp:110-123).  Unfilled (+inf, -1) slots
 * sort after every finite entry. */
int gsknn_rows_ascending_ddev(int n, int k, double *D, int *I, int ldk, void *stream);
int gsknn_rows_ascending_sdev(int n, int k, float *D, int *I, int ldk, void *stream);

/* One level of the randomized-projection tree of the all-NN sweep (no
 * counterpart in the reference library, whose tree codes are outside it):
 * element i < total of the nnodes nodes being split (node n holds perm
 * positions [starts[n], starts[n] + size), elements numbered from offs[n],
 * offs ascending) gets its position pos[i] and the key
 * keys[i] = n << (32 + id_bits) | order(projection on dirs[n*d..]) << id_bits
 * | point id, with node index, 32 order bits and id fitting in 63 bits.
 * Sorting the keys orders each node by (projection, id). */
int gsknn_tree_level_keys_dev(long long total, int nnodes, int d, int id_bits,
                              const float *X, const long long *perm,
                              const long long *starts, const long long *offs,
                              const float *dirs, long long *pos, long long *keys,
                              void *stream);

/* L1-norm kNN (knn_type KNN_1NORM).  The reference registers L1 kernels at
 * index KNN_1NORM of its FP64 kernel tables (kernels/x86_64/sandybridge/
 * gsknn_kernel.h:125-163: rnn_rank_k_abs_int_d8x4, rnn_r_1norm_int_d8x4_row)
 * but its dgsknn always dispatches index 0 (frame/gsknn.c:265, :409), so
 * heap->type never selects them; these entry points are that path, driven
 * as dgsknn_var1 drives the tables (frame/gsknn.c:996-1086): distance =
 * sum_p |a_p - b_p|, summed in p order within blocks of 256 dimensions with
 * the block sums added in order, then the same heap select, ties to the
 * lower index.  No norms are read.
 *
 * dgsknn_1norm / sgsknn_1norm: the dgsknn / sgsknn argument convention
 * (include/gsknn.h:90-116) without XA2/XB2 -- m candidates XA[alpha[i]], n
 * queries XB[beta[j]], k the dimension, heap rows of stride r (r <= 512;
 * the reference's large-r header layout has no L1 form) in heap->D (FP64)
 * or heap->D_s (FP32) and heap->I.  Host or device pointers; errors via
 * gsknn_last_error().  FP32 applies the same rule in float (the reference's
 * FP32 tables hold no L1 row kernel). */
void dgsknn_1norm(int m, int n, int k, int r, double *XA, int *alpha,
                  double *XB, int *beta, heap_t *heap);
void sgsknn_1norm(int m, int n, int k, int r, float *XA, int *alpha,
                  float *XB, int *beta, heap_t *heap);

/* Device-array form (conventions of gsknn_dknn_dev; `work` as there, from
 * gsknn_[ds]knn1_work_bytes). */
size_t gsknn_dknn1_work_bytes(int m, int n, int d, int k);
int gsknn_dknn1_dev(int m, int n, int d, int k, const double *XA,
                    const int *amap, const double *XB, const int *bmap,
                    double *D, int *I, int ldk, const int *hrow, void *work,
                    size_t work_bytes, void *stream);
size_t gsknn_sknn1_work_bytes(int m, int n, int d, int k);
int gsknn_sknn1_dev(int m, int n, int d, int k, const float *XA,
                    const int *amap, const float *XB, const int *bmap,
                    float *D, int *I, int ldk, const int *hrow, void *work,
                    size_t work_bytes, void *stream);

/* Squared norms X2[i] = sum_p X[i*d + p]^2, summed in p order. */
int gsknn_dsqnorm_dev(int N, int d, const double *X, double *X2, void *stream);
int gsknn_ssqnorm_dev(int N, int d, const float *X, float *X2, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GSKNN_MI355X_GSKNN_HIP_H */
