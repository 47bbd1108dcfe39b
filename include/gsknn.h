/*
 * gsknn.h -- drop-in C API of the MI355X exact-kNN library (libgsknn.so).
 *
 * Same symbols, argument order and heap_t layout as GSKNN's public header
 * (ChenhanYu/rnn include/gsknn.h:39-277), so a caller of the reference links
 * against this library unchanged.  Every declaration below cites the
 * reference declaration it replaces.
 *
 * Argument meaning (as in the reference):
 *   m, XA, XA2, amap   m candidate points: rows amap[i] of table XA (row
 *                      stride = dimension), squared norms XA2[amap[i]]
 *   n, XB, XB2, bmap   n query points; heap row j belongs to query bmap[j]
 *   k                  the point dimension
 *   r                  the number of neighbours kept per query
 *   heap->D / heap->I  n rows of r (distance, amap value) pairs, an unordered
 *                      max-heap per row, updated in place across calls
 *
 * Differences a caller can observe:
 *   - The kNN work runs on the current HIP device.  Host pointers are staged
 *     to device memory and back inside the call; device pointers (hipMalloc)
 *     are used in place.  Calls return when the heaps hold the result.
 *   - Ties in distance are resolved by the lower index (the reference keeps
 *     whichever equal candidate it scanned last).  On inputs without exact
 *     ties the neighbour sets are identical.
 *   - Rows are written back sorted descending, which is a valid max-heap.
 *   - Errors never exit the process: the call returns early and
 *     gsknn_last_error() (gsknn_hip.h) describes the failure.
 *   - dgsknn_ref/sgsknn_ref (reference include/gsknn_ref.h) are CPU test
 *     oracles and are not part of this library.
 *   - heapSelect_dheap is declared by the reference but never defined
 *     there (include/gsknn.h:236-243); it is not declared here.
 */
#ifndef GSKNN_MI355X_GSKNN_H
#define GSKNN_MI355X_GSKNN_H

#ifdef __cplusplus
extern "C" {
#endif

/* include/gsknn.h:46-47 */
#define KNN_VAR_THRES 512
#define KNN_HEAP_OFFSET 3

/* include/gsknn.h:49 */
typedef unsigned long long dim_t;

/* include/gsknn.h:51-54 */
typedef enum { KNN_2NORM, KNN_1NORM } knn_type;

/* include/gsknn.h:56-59 */
typedef enum { KNN_DOUBLE, KNN_SINGLE } knn_prec;

/* include/gsknn.h:61-73 (internal to the reference's micro-kernels; kept so
 * that sources naming the type still compile) */
struct aux_s {
  double *b_next;
  float *b_next_s;
  int *I;
  double *D;
  float *D_s;
  int ldr;
  char *flag;
  int pc;
  int m;
  int n;
};
typedef struct aux_s aux_t;

/* include/gsknn.h:75-88, byte-identical layout */
struct heap_s {
  int m;
  int k;
  int ldk;
  int d;
  double ro;
  double *D;
  float ro_s;
  float *D_s;
  int *I;
  knn_type type;
  knn_prec prec;
};
typedef struct heap_s heap_t;

/* include/gsknn.h:90-102: FP32 kNN into heap->D_s / heap->I. */
void sgsknn(int m, int n, int k, int r, float *XA, float *XA2, int *alpha,
            float *XB, float *XB2, int *beta, heap_t *heap);

/* include/gsknn.h:104-116: FP64 kNN into heap->D / heap->I; r > 512 takes
 * the header-layout path of dgsknn_var3. */
void dgsknn(int m, int n, int k, int r, double *XA, double *XA2, int *alpha,
            double *XB, double *XB2, int *beta, heap_t *heap);

/* include/gsknn.h:118-133 and :135-148: the raw-array form.  As in the
 * reference the roles are swapped: here m, XA, amap are the QUERIES (heap
 * rows) and n, XB, bmap the candidates; D and I have m rows of stride r
 * (frame/gsknn.c:1120-1134, :1167-1182, :930). */
void sgsknn_var1(int m, int n, int k, int r, float *XA, float *XA2, int *amap,
                 float *XB, float *XB2, int *bmap, float *D, int *I);
void dgsknn_var1(int m, int n, int k, int r, double *XA, double *XA2,
                 int *amap, double *XB, double *XB2, int *bmap, double *D,
                 int *I);

/* include/gsknn.h:150-162: large-r path (r > 512) on heaps made by
 * heapCreate_d, whose rows carry the 3-slot header of frame/gsknn_heap.c:
 * 257-308 and a 4-ary max-heap behind it. */
void dgsknn_var3(int m, int n, int k, int r, double *XA, double *XA2,
                 int *amap, double *XB, double *XB2, int *bmap, heap_t *heap);

/* include/gsknn.h:164-176: C[i + j*ldc] = -2 XA[amap[i]].XB[bmap[j]]
 * + XA2[amap[i]] + XB2[bmap[j]] for the m x n block (frame/gsknn.c:481-594). */
void dgssq2nrm(int m, int n, int k, double *XA, double *XA2, int *amap,
               double *XB, double *XB2, int *bmap, double *C, int ldc);

/* include/gsknn.h:178-182: 32-byte aligned host allocation of m*n*size
 * bytes.  Returns NULL on failure (the reference exits). */
double *gsknn_malloc_aligned(int m, int n, int size);

/* include/gsknn.h:184-216: heap rows, initialised to (ro, -1). */
heap_t *heapCreate_s(int m, int k, float ro);
heap_t *heapCreate_d(int m, int k, double ro);
void heapFree_s(heap_t *heap);
void heapFree_d(heap_t *heap);
heap_t *heapAttach_s(int m, int k, float *D, int *I);
heap_t *heapAttach_d(int m, int k, double *D, int *I);

/* include/gsknn.h:218-234: host-side heap select over m values into one row,
 * with the reference's rule (replace the root when x <= root). */
void heapSelect_s(int m, int r, float *x, int *alpha, float *D, int *I);
void heapSelect_d(int m, int r, double *x, int *alpha, double *D, int *I);

/* include/gsknn.h:245-273 */
void swap_float(float *x, int i, int j);
void swap_double(double *x, int i, int j);
void swap_int(int *I, int i, int j);
void bubbleSort_s(int n, float *D, int *I);
void bubbleSort_d(int n, double *D, int *I);

#ifdef __cplusplus
}
#endif

#endif /* GSKNN_MI355X_GSKNN_H */
