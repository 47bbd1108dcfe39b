/*
 * C caller of libgsknn.so through include/gsknn.h, the way the reference's
 * own drivers call its library (test/test_dgsknn.c:133-191): heapCreate_d,
 * dgsknn over one or more candidate maps into the same heap rows, read the
 * rows, heapFree_d.
 *
 * usage: capi_driver IN OUT
 *   IN  (binary, native endianness): int32 N, d, n, k, ncalls; double X[N*d];
 *       double X2[N]; int32 bmap[n]; then per call: int32 m; int32 amap[m].
 *   OUT (binary): for each call, the n rows of k (double D, then int32 I),
 *       read from heap->D / heap->I with row stride heap->ldk.
 * Exit status 0 on success.
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>

#include "gsknn.h"

/* heap_t must stay byte-identical to the reference's (include/gsknn.h:75-88). */
_Static_assert(sizeof(heap_t) == 64, "heap_t size");
_Static_assert(offsetof(heap_t, m) == 0, "heap_t.m");
_Static_assert(offsetof(heap_t, k) == 4, "heap_t.k");
_Static_assert(offsetof(heap_t, ldk) == 8, "heap_t.ldk");
_Static_assert(offsetof(heap_t, d) == 12, "heap_t.d");
_Static_assert(offsetof(heap_t, ro) == 16, "heap_t.ro");
_Static_assert(offsetof(heap_t, D) == 24, "heap_t.D");
_Static_assert(offsetof(heap_t, ro_s) == 32, "heap_t.ro_s");
_Static_assert(offsetof(heap_t, D_s) == 40, "heap_t.D_s");
_Static_assert(offsetof(heap_t, I) == 48, "heap_t.I");
_Static_assert(offsetof(heap_t, type) == 56, "heap_t.type");
_Static_assert(offsetof(heap_t, prec) == 60, "heap_t.prec");

static int read_all(FILE *f, void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  FILE *in = fopen(argv[1], "rb");
  if (!in) return 3;
  int hdr[5];
  if (!read_all(in, hdr, sizeof hdr)) return 4;
  const int N = hdr[0], d = hdr[1], n = hdr[2], k = hdr[3], ncalls = hdr[4];
  double *X = (double *)malloc(sizeof(double) * (size_t)N * d);
  double *X2 = (double *)malloc(sizeof(double) * (size_t)N);
  int *bmap = (int *)malloc(sizeof(int) * (size_t)n);
  if (!X || !X2 || !bmap) return 5;
  if (!read_all(in, X, sizeof(double) * (size_t)N * d) || !read_all(in, X2, sizeof(double) * (size_t)N) ||
      !read_all(in, bmap, sizeof(int) * (size_t)n))
    return 6;

  heap_t *heap = heapCreate_d(n, k, 1.79E+308);
  if (!heap || heap->m != n || heap->k != k || heap->ldk != k) return 7;
  FILE *out = fopen(argv[2], "wb");
  if (!out) return 8;
  for (int c = 0; c < ncalls; ++c) {
    int m = 0;
    if (!read_all(in, &m, sizeof m)) return 9;
    int *amap = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    if (!amap || !read_all(in, amap, sizeof(int) * (size_t)m)) return 10;
    /* dgsknn(m, n, k = dimension, r = neighbours, ...): frame/gsknn.c:1138-1184 */
    dgsknn(m, n, d, k, X, X2, amap, X, X2, bmap, heap);
    for (int j = 0; j < n; ++j) {
      fwrite(heap->D + (size_t)j * heap->ldk, sizeof(double), (size_t)k, out);
      fwrite(heap->I + (size_t)j * heap->ldk, sizeof(int), (size_t)k, out);
    }
    free(amap);
  }
  fclose(out);
  fclose(in);
  heapFree_d(heap);
  free(X);
  free(X2);
  free(bmap);
  return 0;
}
