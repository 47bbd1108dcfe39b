/*
 * gsknn_oracle.c -- CPU restatement of the GSKNN reference oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * rnn_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product library never links or calls it.
 *
 * What it restates (file:line in ChenhanYu/rnn):
 *   - dgsknn_ref / sgsknn_ref                frame/gsknn_ref.c:63-162
 *   - gather, sequential dot in p order,
 *     x(-2), + XA2[alpha[i]], + XB2[beta[j]],
 *     no clamp                               frame/dgsknn_ref_impl.h:41-103
 *                                            frame/sgsknn_ref_impl.h:41-103
 *   - heapSelect_[sd] + HeapAdjust_[sd]      frame/gsknn_heap.c:52-162
 *   - heap row stride = r (k) for k <= 512   frame/gsknn.c:930
 *   - L1 norm (KNN_1NORM kernels, blocked
 *     sum of |a_p - b_p|)                    frame/gsknn.c:996-1086,
 *                                            kernels/x86_64/sandybridge/
 *                                            rnn_rank_k_abs_int_d8x4.c,
 *                                            rnn_r_1norm_int_d8x4_row.c
 *
 * Two deliberate differences from the reference oracle:
 *   1. Blocking.  The reference materialises the m x n matrix Cs
 *      (gsknn_ref.c:141), 34 GB at m = n = 65536.  Here each heap row j scans
 *      candidates i = 0..m-1 in increasing order, computing each distance with
 *      exactly the reference's arithmetic, so the sequence of heap operations
 *      per row -- and therefore the heap array -- is identical.
 *   2. A canonical tie mode.  The reference replaces the heap root when
 *      x <= root (gsknn_heap.c:153), so among equal distances the one scanned
 *      last wins.  Canonical mode orders entries by (distance, index) and
 *      replaces the root only when the new entry is strictly smaller, so the
 *      lower index wins a tie and the result is independent of scan order.
 *
 * Compile with -ffp-contract=off so that a*b + c is never fused: the
 * reference's build (-O3 on baseline x86-64, no FMA unit) rounds the product
 * and the sum separately.
 *
 * Parity pins (tests/test_oracle.py): heap arrays bit-identical to the
 * reference compiled from its own sources (oracle/_ref, when present) and to
 * the committed fixtures in tests/golden/ made from it; brute force in long
 * double on small inputs; integer lattices with exact ties; chunking and
 * permutation invariants; d = 1 reducing to sorting |q - r|.
 * L1 pins (tests/test_oracle_l1.py): heap arrays bit-identical to the
 * reference's own L1 kernels reached through its kernel tables
 * (oracle/_ref/libgsknn_opt.so, live and as fixtures), an unblocked sum
 * failing that fixture, long-double brute force, integer lattices, d = 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Heap primitives.                                                    */
/* ------------------------------------------------------------------ */

/* Compat order: distance only; follows HeapAdjust_d (gsknn_heap.c:75-96). */
#define DEF_ADJUST_COMPAT(T, NAME)                                          \
  static void NAME(T *D, int s, int n, int *I) {                            \
    while (2 * s + 1 < n) {                                                 \
      int j = 2 * s + 1;                                                    \
      if (j + 1 < n && D[j] < D[j + 1]) j++;                                \
      if (D[s] < D[j]) {                                                    \
        T td = D[s]; D[s] = D[j]; D[j] = td;                                \
        int ti = I[s]; I[s] = I[j]; I[j] = ti;                              \
        s = j;                                                              \
      } else {                                                              \
        break;                                                              \
      }                                                                     \
    }                                                                       \
  }
DEF_ADJUST_COMPAT(double, adjust_compat_d)
DEF_ADJUST_COMPAT(float, adjust_compat_s)

/* Canonical order: (distance, index) lexicographic. */
#define CANON_LESS(da, ia, db, ib) ((da) < (db) || ((da) == (db) && (ia) < (ib)))

#define DEF_ADJUST_CANON(T, NAME)                                           \
  static void NAME(T *D, int s, int n, int *I) {                            \
    while (2 * s + 1 < n) {                                                 \
      int j = 2 * s + 1;                                                    \
      if (j + 1 < n && CANON_LESS(D[j], I[j], D[j + 1], I[j + 1])) j++;     \
      if (CANON_LESS(D[s], I[s], D[j], I[j])) {                             \
        T td = D[s]; D[s] = D[j]; D[j] = td;                                \
        int ti = I[s]; I[s] = I[j]; I[j] = ti;                              \
        s = j;                                                              \
      } else {                                                              \
        break;                                                              \
      }                                                                     \
    }                                                                       \
  }
DEF_ADJUST_CANON(double, adjust_canon_d)
DEF_ADJUST_CANON(float, adjust_canon_s)

/* heapSelect_d (gsknn_heap.c:133-162): for each x in scan order, skip when
 * x > root, else overwrite the root and sift down. */
void oracle_heap_select_d(int m, int k, const double *x, const int *alpha,
                          double *D, int *I, int canonical) {
  for (int i = 0; i < m; i++) {
    if (canonical) {
      if (!CANON_LESS(x[i], alpha[i], D[0], I[0])) continue;
      D[0] = x[i]; I[0] = alpha[i];
      adjust_canon_d(D, 0, k, I);
    } else {
      if (x[i] > D[0]) continue;
      D[0] = x[i]; I[0] = alpha[i];
      adjust_compat_d(D, 0, k, I);
    }
  }
}

void oracle_heap_select_s(int m, int k, const float *x, const int *alpha,
                          float *D, int *I, int canonical) {
  for (int i = 0; i < m; i++) {
    if (canonical) {
      if (!CANON_LESS(x[i], alpha[i], D[0], I[0])) continue;
      D[0] = x[i]; I[0] = alpha[i];
      adjust_canon_s(D, 0, k, I);
    } else {
      if (x[i] > D[0]) continue;
      D[0] = x[i]; I[0] = alpha[i];
      adjust_compat_s(D, 0, k, I);
    }
  }
}

/* Bottom-up heapify under the canonical order (a row handed in by a caller
 * may be a max-heap by distance only). */
void oracle_heapify_canon_d(int k, double *D, int *I) {
  for (int s = k / 2 - 1; s >= 0; s--) adjust_canon_d(D, s, k, I);
}
void oracle_heapify_canon_s(int k, float *D, int *I) {
  for (int s = k / 2 - 1; s >= 0; s--) adjust_canon_s(D, s, k, I);
}

/* ------------------------------------------------------------------ */
/* Distances with the reference oracle's arithmetic.                    */
/* ------------------------------------------------------------------ */

/* dgsknn_ref_impl.h:63-72 then :91-103: Cs = sum_p As[i,p]*Bs[j,p] in p order
 * (product rounded, then sum rounded), Cs *= -2, Cs += XA2[alpha[i]],
 * Cs += XB2[beta[j]]. */
static inline double dist_ref_d(const double *a, const double *b, int d,
                                double a2, double b2) {
  double s = 0.0;
  for (int p = 0; p < d; p++) s += a[p] * b[p];
  s *= -2.0;
  s += a2;
  s += b2;
  return s;
}

/* sgsknn_ref_impl.h: the same in float; Cs *= -2.0 promotes to double and
 * rounds back, which is exact. */
static inline float dist_ref_s(const float *a, const float *b, int d, float a2,
                               float b2) {
  float s = 0.0f;
  for (int p = 0; p < d; p++) s += a[p] * b[p];
  s = (float)((double)s * -2.0);
  s += a2;
  s += b2;
  return s;
}

/* ------------------------------------------------------------------ */
/* Blocked dgsknn_ref / sgsknn_ref.                                     */
/* ------------------------------------------------------------------ */

/* m candidates (XA, XA2, amap), n queries (XB, XB2, bmap); heap row j
 * (stride ldk) belongs to query bmap[j]; the stored index is amap[i]. */
void oracle_dgsknn(int m, int n, int d, int k, const double *XA,
                   const double *XA2, const int *amap, const double *XB,
                   const double *XB2, const int *bmap, double *D, int *I,
                   int ldk, int canonical) {
  if (m == 0 || n == 0 || d == 0 || k == 0) return; /* gsknn_ref.c:135 */
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; j++) {
    const double *b = XB + (size_t)bmap[j] * d;
    const double b2 = XB2[bmap[j]];
    double *Dj = D + (size_t)j * ldk;
    int *Ij = I + (size_t)j * ldk;
    if (canonical) oracle_heapify_canon_d(k, Dj, Ij);
    for (int i = 0; i < m; i++) {
      const double *a = XA + (size_t)amap[i] * d;
      double x = dist_ref_d(a, b, d, XA2[amap[i]], b2);
      oracle_heap_select_d(1, k, &x, &amap[i], Dj, Ij, canonical);
    }
  }
}

void oracle_sgsknn(int m, int n, int d, int k, const float *XA,
                   const float *XA2, const int *amap, const float *XB,
                   const float *XB2, const int *bmap, float *D, int *I,
                   int ldk, int canonical) {
  if (m == 0 || n == 0 || d == 0 || k == 0) return;
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; j++) {
    const float *b = XB + (size_t)bmap[j] * d;
    const float b2 = XB2[bmap[j]];
    float *Dj = D + (size_t)j * ldk;
    int *Ij = I + (size_t)j * ldk;
    if (canonical) oracle_heapify_canon_s(k, Dj, Ij);
    for (int i = 0; i < m; i++) {
      const float *a = XA + (size_t)amap[i] * d;
      float x = dist_ref_s(a, b, d, XA2[amap[i]], b2);
      oracle_heap_select_s(1, k, &x, &amap[i], Dj, Ij, canonical);
    }
  }
}

/* ------------------------------------------------------------------ */
/* L1 norm (knn_type KNN_1NORM).                                        */
/* ------------------------------------------------------------------ */

/* The reference registers L1 kernels at index KNN_1NORM of its FP64 kernel
 * tables (kernels/x86_64/sandybridge/gsknn_kernel.h:125-163) but its driver
 * always calls index 0 (frame/gsknn.c:265, :409).  Routed through the
 * tables, dgsknn_var1 computes, per pair, |a_p - b_p| summed in p order
 * within blocks of DKNN_KC = 256 dimensions, each block sum starting from
 * zero and the block sums added in order (frame/gsknn.c:996-1086,
 * rnn_rank_k_abs_int_d8x4.c:62-195 and :226-247, rnn_r_1norm_int_d8x4_row.c
 * :46-275), then max(0, s), a no-op on these sums, and the same row heap
 * select as the 2-norm.  dgsknn_ref_impl.h:74-85 keeps the unblocked sum as
 * a comment; it equals this for d <= 256.  The FP32 tables hold no L1 row
 * kernel; the FP32 form here applies the same rule in float. */
#define ORACLE_L1_KC 256

static inline double dist_l1_d(const double *q, const double *c, int d) {
  double s = 0.0;
  for (int p0 = 0; p0 < d; p0 += ORACLE_L1_KC) {
    const int pe = p0 + ORACLE_L1_KC < d ? p0 + ORACLE_L1_KC : d;
    double t = 0.0;
    for (int p = p0; p < pe; p++) t += fabs(q[p] - c[p]);
    s = p0 == 0 ? t : s + t;
  }
  return s;
}

static inline float dist_l1_s(const float *q, const float *c, int d) {
  float s = 0.0f;
  for (int p0 = 0; p0 < d; p0 += ORACLE_L1_KC) {
    const int pe = p0 + ORACLE_L1_KC < d ? p0 + ORACLE_L1_KC : d;
    float t = 0.0f;
    for (int p = p0; p < pe; p++) t += fabsf(q[p] - c[p]);
    s = p0 == 0 ? t : s + t;
  }
  return s;
}

/* Same roles and heap rows as oracle_dgsknn: m candidates (XA, amap), n
 * queries (XB, bmap), heap row j of stride ldk; no norms. */
void oracle_dgsknn_l1(int m, int n, int d, int k, const double *XA,
                      const int *amap, const double *XB, const int *bmap,
                      double *D, int *I, int ldk, int canonical) {
  if (m == 0 || n == 0 || d == 0 || k == 0) return;
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; j++) {
    const double *b = XB + (size_t)bmap[j] * d;
    double *Dj = D + (size_t)j * ldk;
    int *Ij = I + (size_t)j * ldk;
    if (canonical) oracle_heapify_canon_d(k, Dj, Ij);
    for (int i = 0; i < m; i++) {
      double x = dist_l1_d(b, XA + (size_t)amap[i] * d, d);
      oracle_heap_select_d(1, k, &x, &amap[i], Dj, Ij, canonical);
    }
  }
}

void oracle_sgsknn_l1(int m, int n, int d, int k, const float *XA,
                      const int *amap, const float *XB, const int *bmap,
                      float *D, int *I, int ldk, int canonical) {
  if (m == 0 || n == 0 || d == 0 || k == 0) return;
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; j++) {
    const float *b = XB + (size_t)bmap[j] * d;
    float *Dj = D + (size_t)j * ldk;
    int *Ij = I + (size_t)j * ldk;
    if (canonical) oracle_heapify_canon_s(k, Dj, Ij);
    for (int i = 0; i < m; i++) {
      float x = dist_l1_s(b, XA + (size_t)amap[i] * d, d);
      oracle_heap_select_s(1, k, &x, &amap[i], Dj, Ij, canonical);
    }
  }
}

/* The reference's own entry points (gsknn_ref.h:40-68), compat mode, row
 * stride r.  Argument k is the dimension and r the neighbour count, as in
 * the reference. */
void dgsknn_ref(int m, int n, int k, int r, double *XA, double *XA2,
                int *alpha, double *XB, double *XB2, int *beta, double *D,
                int *I) {
  if (m == 0 || n == 0 || k == 0 || r == 0) return;
  oracle_dgsknn(m, n, k, r, XA, XA2, alpha, XB, XB2, beta, D, I, r, 0);
}

void sgsknn_ref(int m, int n, int k, int r, float *XA, float *XA2, int *alpha,
                float *XB, float *XB2, int *beta, float *D, int *I) {
  if (m == 0 || n == 0 || k == 0 || r == 0) return;
  oracle_sgsknn(m, n, k, r, XA, XA2, alpha, XB, XB2, beta, D, I, r, 0);
}

/* ------------------------------------------------------------------ */
/* Helpers for comparisons.                                             */
/* ------------------------------------------------------------------ */

typedef struct { double d; int i; } pair_d;
typedef struct { float d; int i; } pair_s;

static int cmp_pair_d(const void *x, const void *y) {
  const pair_d *a = (const pair_d *)x, *b = (const pair_d *)y;
  if (a->d < b->d) return -1;
  if (a->d > b->d) return 1;
  return (a->i > b->i) - (a->i < b->i);
}
static int cmp_pair_s(const void *x, const void *y) {
  const pair_s *a = (const pair_s *)x, *b = (const pair_s *)y;
  if (a->d < b->d) return -1;
  if (a->d > b->d) return 1;
  return (a->i > b->i) - (a->i < b->i);
}

/* Sort every row ascending by (distance, index), in place. */
void oracle_sort_rows_d(int n, int k, double *D, int *I, int ldk) {
#pragma omp parallel
  {
    pair_d *buf = (pair_d *)malloc(sizeof(pair_d) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(static)
    for (int j = 0; j < n; j++) {
      for (int t = 0; t < k; t++) {
        buf[t].d = D[(size_t)j * ldk + t];
        buf[t].i = I[(size_t)j * ldk + t];
      }
      qsort(buf, (size_t)k, sizeof(pair_d), cmp_pair_d);
      for (int t = 0; t < k; t++) {
        D[(size_t)j * ldk + t] = buf[t].d;
        I[(size_t)j * ldk + t] = buf[t].i;
      }
    }
    free(buf);
  }
}

void oracle_sort_rows_s(int n, int k, float *D, int *I, int ldk) {
#pragma omp parallel
  {
    pair_s *buf = (pair_s *)malloc(sizeof(pair_s) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(static)
    for (int j = 0; j < n; j++) {
      for (int t = 0; t < k; t++) {
        buf[t].d = D[(size_t)j * ldk + t];
        buf[t].i = I[(size_t)j * ldk + t];
      }
      qsort(buf, (size_t)k, sizeof(pair_s), cmp_pair_s);
      for (int t = 0; t < k; t++) {
        D[(size_t)j * ldk + t] = buf[t].d;
        I[(size_t)j * ldk + t] = buf[t].i;
      }
    }
    free(buf);
  }
}

/* 1 if every row is a valid max-heap under distance order (canonical = 0) or
 * under (distance, index) order (canonical = 1). */
int oracle_rows_are_heaps_d(int n, int k, const double *D, const int *I,
                            int ldk, int canonical) {
  for (int j = 0; j < n; j++) {
    const double *Dj = D + (size_t)j * ldk;
    const int *Ij = I + (size_t)j * ldk;
    for (int c = 1; c < k; c++) {
      int p = (c - 1) / 2;
      if (canonical ? CANON_LESS(Dj[p], Ij[p], Dj[c], Ij[c]) : (Dj[p] < Dj[c]))
        return 0;
    }
  }
  return 1;
}

int oracle_rows_are_heaps_s(int n, int k, const float *D, const int *I,
                            int ldk, int canonical) {
  for (int j = 0; j < n; j++) {
    const float *Dj = D + (size_t)j * ldk;
    const int *Ij = I + (size_t)j * ldk;
    for (int c = 1; c < k; c++) {
      int p = (c - 1) / 2;
      if (canonical ? CANON_LESS(Dj[p], Ij[p], Dj[c], Ij[c]) : (Dj[p] < Dj[c]))
        return 0;
    }
  }
  return 1;
}

/* FP64 direct-difference distances sum_p (a_p - b_p)^2 under the same
 * canonical selection: the exact-rerank definition used by the BF16 form.
 * Points are FP32 (upcast exactly); the sum runs in p order. */
void oracle_direct_knn_f32in(int m, int n, int d, int k, const float *XA,
                             const int *amap, const float *XB,
                             const int *bmap, double *D, int *I, int ldk) {
  if (m == 0 || n == 0 || d == 0 || k == 0) return;
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; j++) {
    const float *b = XB + (size_t)bmap[j] * d;
    double *Dj = D + (size_t)j * ldk;
    int *Ij = I + (size_t)j * ldk;
    oracle_heapify_canon_d(k, Dj, Ij);
    for (int i = 0; i < m; i++) {
      const float *a = XA + (size_t)amap[i] * d;
      double s = 0.0;
      for (int p = 0; p < d; p++) {
        double t = (double)a[p] - (double)b[p];
        s += t * t;
      }
      oracle_heap_select_d(1, k, &s, &amap[i], Dj, Ij, 1);
    }
  }
}

/* Brute force in long double by direct differences, then a full sort by
 * (distance, index); writes the k smallest ascending.  For tiny inputs. */
void oracle_brute_d(int m, int n, int d, int k, const double *XA,
                    const int *amap, const double *XB, const int *bmap,
                    double *D, int *I, int ldk) {
  pair_d *buf = (pair_d *)malloc(sizeof(pair_d) * (size_t)(m > 0 ? m : 1));
  for (int j = 0; j < n; j++) {
    const double *b = XB + (size_t)bmap[j] * d;
    for (int i = 0; i < m; i++) {
      const double *a = XA + (size_t)amap[i] * d;
      long double s = 0.0L;
      for (int p = 0; p < d; p++) {
        long double t = (long double)a[p] - (long double)b[p];
        s += t * t;
      }
      buf[i].d = (double)s;
      buf[i].i = amap[i];
    }
    qsort(buf, (size_t)m, sizeof(pair_d), cmp_pair_d);
    for (int t = 0; t < k && t < m; t++) {
      D[(size_t)j * ldk + t] = buf[t].d;
      I[(size_t)j * ldk + t] = buf[t].i;
    }
  }
  free(buf);
}

/* L1 brute force in long double (sum of |a_p - b_p| with one final
 * rounding for d < 2^11), full sort by (distance, index); tiny inputs. */
void oracle_brute_l1_d(int m, int n, int d, int k, const double *XA,
                       const int *amap, const double *XB, const int *bmap,
                       double *D, int *I, int ldk) {
  pair_d *buf = (pair_d *)malloc(sizeof(pair_d) * (size_t)(m > 0 ? m : 1));
  for (int j = 0; j < n; j++) {
    const double *b = XB + (size_t)bmap[j] * d;
    for (int i = 0; i < m; i++) {
      const double *a = XA + (size_t)amap[i] * d;
      long double s = 0.0L;
      for (int p = 0; p < d; p++) s += fabsl((long double)a[p] - (long double)b[p]);
      buf[i].d = (double)s;
      buf[i].i = amap[i];
    }
    qsort(buf, (size_t)m, sizeof(pair_d), cmp_pair_d);
    for (int t = 0; t < k && t < m; t++) {
      D[(size_t)j * ldk + t] = buf[t].d;
      I[(size_t)j * ldk + t] = buf[t].i;
    }
  }
  free(buf);
}

/* Squared norms X2[i] = sum_p X[i][p]^2 in p order (test/test_dgsknn.c:152-159). */
void oracle_sqnorm_d(int N, int d, const double *X, double *X2) {
  for (int i = 0; i < N; i++) {
    double s = 0.0;
    for (int p = 0; p < d; p++) s += X[(size_t)i * d + p] * X[(size_t)i * d + p];
    X2[i] = s;
  }
}
void oracle_sqnorm_s(int N, int d, const float *X, float *X2) {
  for (int i = 0; i < N; i++) {
    float s = 0.0f;
    for (int p = 0; p < d; p++) s += X[(size_t)i * d + p] * X[(size_t)i * d + p];
    X2[i] = s;
  }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
