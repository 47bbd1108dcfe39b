// Device helpers shared by the kNN kernels: canonical order, exact
// arithmetic, LDS heaps and the exact (oracle-arithmetic) block routine.
//
// Reference behaviour (ChenhanYu/rnn):
//   distance  = (-2 * sum_p a_p b_p + XA2[amap[i]]) + XB2[bmap[j]], no clamp
//               (frame/dgsknn_ref_impl.h:63-72, :91-103)
//   selection = per heap row, keep the k smallest in a max-heap whose root is
//               the current k-th distance (frame/gsknn_heap.c:104-162)
#pragma once

#include <limits.h>

#include <type_traits>

#include "gsknn_common.h"

namespace gsknn {

__host__ __device__ constexpr size_t al16(size_t x) { return (x + 15) & ~size_t(15); }

template <typename T>
__device__ __forceinline__ bool cless(T da, int ia, T db, int ib) {
  return da < db || (da == db && ia < ib);
}

// Rounded product and sum, never fused (the library is built with
// -ffp-contract=off as well).
__device__ __forceinline__ double xmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double xadd(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float xmul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float xadd(float a, float b) { return __fadd_rn(a, b); }

__device__ __forceinline__ double xsub(double a, double b) { return __dsub_rn(a, b); }
__device__ __forceinline__ float xsub(float a, float b) { return __fsub_rn(a, b); }

// The reference oracle's distance between candidate row c and query row q:
// s = sum_p c_p q_p in p order from 0.0, every product and sum rounded; then
// (s x (-2) + cn) + qn, no clamp (frame/dgsknn_ref_impl.h:63-72, :91-103).
__device__ __forceinline__ double oracle_dist(const double *c, const double *q, int d, double cn,
                                              double qn) {
  double s = 0.0;
  int p = 0;
  for (; p + 4 <= d; p += 4) {
    const double c0 = c[p], c1 = c[p + 1], c2 = c[p + 2], c3 = c[p + 3];
    const double q0 = q[p], q1 = q[p + 1], q2 = q[p + 2], q3 = q[p + 3];
    s = xadd(s, xmul(c0, q0));
    s = xadd(s, xmul(c1, q1));
    s = xadd(s, xmul(c2, q2));
    s = xadd(s, xmul(c3, q3));
  }
  for (; p < d; ++p) s = xadd(s, xmul(c[p], q[p]));
  return xadd(xadd(xmul(s, -2.0), cn), qn);
}

typedef double xdbl2 __attribute__((ext_vector_type(2)));

// Oracle distances of R pairs at once (candidate row c[r], query row q[r]):
// each sum in p order from 0.0, every product and sum rounded; the R
// independent sums are interleaved so their loads overlap (UNR dim pairs of
// loads in flight per sum).  VEC: rows 16-byte aligned, d even.
template <int R, bool VEC, int UNR>
__device__ __forceinline__ void oracle_dist_batch(const double *const (&c)[R],
                                                  const double *const (&q)[R], int d,
                                                  const double (&cn)[R], const double (&qn)[R],
                                                  double (&out)[R]) {
  double s[R];
#pragma unroll
  for (int r = 0; r < R; ++r) s[r] = 0.0;
  int p = 0;
  if (VEC) {
#pragma unroll UNR
    for (; p + 2 <= d; p += 2) {
      xdbl2 cv[R], qv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cv[r] = *(const xdbl2 *)(c[r] + p);
        qv[r] = *(const xdbl2 *)(q[r] + p);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        s[r] = xadd(s[r], xmul(cv[r].x, qv[r].x));
        s[r] = xadd(s[r], xmul(cv[r].y, qv[r].y));
      }
    }
  }
  for (; p < d; ++p) {
#pragma unroll
    for (int r = 0; r < R; ++r) s[r] = xadd(s[r], xmul(c[r][p], q[r][p]));
  }
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = xadd(xadd(xmul(s[r], -2.0), cn[r]), qn[r]);
}

// FP64 MFMA filter margins.  The MFMA dot product s_m of a pair (a, b) of
// dimension d, summed over T = 4 KS terms (zero padding included) in any
// order, and the oracle's sequential sum s_o both lie within
// gamma_n sum_p |a_p b_p| of the exact dot product (gamma_n = n u / (1 - n u),
// u = 2^-53, n = T + 8 and d), so
//   |2 s_m - 2 s_o| <= c (|a|^2 + |b|^2) / 2,  c = 2.5 (T + 8 + d) u
// (Cauchy-Schwarz, 2 |a||b| <= |a|^2 + |b|^2, 25 % slack for the roundings of
// the margins themselves and of the squared norms, which are computed from
// the operands, not taken from the caller).  A kernel then shifts the
// candidate's norm down by c |a|^2 / 2 (mfma_shift_norm) and the query's
// threshold up by c |b|^2 / 2 (mfma_threshold): every pair whose oracle
// distance can reach the row's root passes the filter, and a passing pair's
// exact distance is formed by oracle_dist.
__device__ __forceinline__ double mfma_margin_c(int terms, int d) {
  return 0x1p-53 * 2.5 * (double)(terms + 8 + d);
}

// fl(cn - e) with e = 1.5 max(c a2 / 2, 4 u |cn|): at least c a2 / 2 below cn
// after its own rounding, whatever the caller's norm cn.  NaN stays NaN
// (empty candidate slots carry NaN norms and never pass).
__device__ __forceinline__ double mfma_shift_norm(double cn, double a2, double c) {
  const double e = xmul(1.5, fmax(xmul(xmul(0.5, c), a2), xmul(0x1p-51, __builtin_fabs(cn))));
  return xsub(cn, e);
}

// Threshold on t = fl(-2 s_m + shifted cn) for a row of root r and query norm
// qn: a pair whose oracle distance fl(fl(-2 s_o + cn) + qn) is <= r has
// t <= fl(fl(r - qn) + 2^-48 (|r| + |qn|) + c b2 / 2 + 2^-1000).  The relative
// term covers the roundings of the distance, of t and of this sum; +inf
// roots give +inf.
__device__ __forceinline__ double mfma_threshold(double r, double qn, double half_c_b2) {
  const double m = xadd(xadd(xmul(0x1p-48, xadd(__builtin_fabs(r), __builtin_fabs(qn))), half_c_b2),
                        0x1p-1000);
  return xadd(xsub(r, qn), m);
}

// ---------------------------------------------------------------------------
// FP64 MFMA two-stage form.  Stage 1 (the streaming kernel) keeps, per query,
// the k2 = k + slack smallest candidates under the canonical order of
// (t, id), t = fl(-2 s_m + shifted cn) the pair's MFMA key, in an LDS
// max-heap (no row is read for a passing pair), and writes it out with its
// root first.  A candidate leaves the heap, or never enters it, only with a
// key at or above the heap's root at that moment, and roots only fall, so
// every candidate outside the final heap has t >= T, the final root key
// (when the heap never filled, no candidate was left out by it).
// Candidates the filter drops under the resident row's root r0 have
// t > mfma_threshold(r0), so their oracle distance exceeds r0.
// Stage 2 (knn_two_stage_finalize_kernel) forms the oracle distances of the
// heap's entries, merges them with the resident row and takes the k smallest
// entries; E, the k-th, is then at most r0.  If the heap filled and
// T > mfma_threshold(E), every candidate left out has t > mfma_threshold(E)
// and so an oracle distance above E: the row is the oracle's.  Otherwise the
// query is re-run by the one-stage kernel (oracle distance for every passing
// pair), which happens only when more than `slack` candidates lie within the
// MFMA margin of the k-th distance (exact ties, duplicated points).
// ---------------------------------------------------------------------------
struct MfmaHeapLds {
  size_t hd, hi, r0, bytes;
  __host__ __device__ MfmaHeapLds(int Q, int kh, bool two) {
    hd = 0;
    hi = al16(sizeof(double) * (size_t)kh * Q);
    r0 = al16(hi + sizeof(int) * (size_t)kh * Q);
    bytes = two ? r0 + sizeof(double) * (size_t)Q : hi + sizeof(int) * (size_t)kh * Q;
  }
};

// Heap entries per query of the two-stage form.
__host__ __device__ constexpr int two_stage_k2(int k) { return k + (k / 8 > 8 ? k / 8 : 8); }

// Entries the finalize kernel sorts per query: the next power of two
// >= k + k2 (at least 64); the form is planned only up to kTwoMaxP.
__host__ __device__ inline int two_stage_sort_size(int k, int k2) {
  int P = 64;
  while (P < k + k2) P *= 2;
  return P;
}
constexpr int kTwoMaxP = 2048;

// Empty key heap of slot q (kh entries of (+inf, INT_MAX)); r0[q] = the
// resident row's largest distance.
template <typename T>
__device__ void key_heap_init(const KnnArgs<T> &a, T *hd, int *hi, T *r0, int QS, int q, int j,
                              int kh) {
  for (int t = 0; t < kh; ++t) {
    hd[t * QS + q] = __builtin_inf();
    hi[t * QS + q] = INT_MAX;
  }
  const T *Dr = a.D + (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
  const int *Ir = a.I + (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
  T md = Dr[0];
  int mi = Ir[0];
  for (int t = 1; t < a.k; ++t)
    if (cless(md, mi, Dr[t], Ir[t])) {
      md = Dr[t];
      mi = Ir[t];
    }
  r0[q] = md;
}

// Filter threshold of the two-stage form: the resident row's threshold, or
// the key heap's root once it is below it.  A NaN r0 admits nothing, as in
// the one-stage form.
__device__ __forceinline__ double key_threshold(double r0, double hroot, double qn,
                                                double half_c_b2) {
  const double t0 = mfma_threshold(r0, qn, half_c_b2);
  return hroot < t0 ? hroot : t0;
}

// Key heap of slot q out to row j of a.hk/a.hid, heap order (root first).
template <typename T>
__device__ void key_heap_store(const KnnArgs<T> &a, const T *hd, const int *hi, int QS, int q,
                               int j, int kh) {
  T *ok = a.hk + (size_t)j * kh;
  int *oi = a.hid + (size_t)j * kh;
  for (int t = 0; t < kh; ++t) {
    ok[t] = hd[t * QS + q];
    oi[t] = hi[t * QS + q];
  }
}

// Constant-address-space view of a global array: uniform reads through it
// become scalar loads (SGPR operands shared by the whole wave).  Only arrays
// a kernel never writes are viewed so.
template <typename T>
using sptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ sptr<T> scalar_view(const T *p) {
  return (sptr<T>)p;
}

// The same pointer, which the compiler must treat as computed from the TC
// values: loads through it cannot be issued before every value exists.  The
// asm statements are empty.  Used to keep one chunk's scalar loads from
// being hoisted above the previous chunk's arithmetic (hoisting them all
// would need more SGPRs than a wave has).
template <typename T, int TC>
__device__ __forceinline__ sptr<T> after(sptr<T> p, const T (&v)[TC]) {
#pragma unroll
  for (int u = 0; u < TC; ++u) asm volatile("" : "+s"(p) : "v"(v[u]));
  return p;
}

// Sift (xd, xi) down from slot s of the heap of query q (k entries, LDS,
// stride QS between entries).
template <typename T>
__device__ void sift_down(T *hd, int *hi, int QS, int q, int s, int k, T xd, int xi) {
  while (true) {
    int c = 2 * s + 1;
    if (c >= k) break;
    T cd = hd[c * QS + q];
    int ci = hi[c * QS + q];
    if (c + 1 < k) {
      T rd = hd[(c + 1) * QS + q];
      int ri = hi[(c + 1) * QS + q];
      if (cless(cd, ci, rd, ri)) {
        ++c;
        cd = rd;
        ci = ri;
      }
    }
    if (!cless(xd, xi, cd, ci)) break;
    hd[s * QS + q] = cd;
    hi[s * QS + q] = ci;
    s = c;
  }
  hd[s * QS + q] = xd;
  hi[s * QS + q] = xi;
}

template <typename T>
__device__ __forceinline__ int heap_row(const KnnArgs<T> &a, int j) {
  return a.hrow ? a.hrow[j] : j;
}

// Bring the resident row of query j into the LDS heap of slot q.  A launch
// over all candidates heapifies the row under the canonical order; a split
// launch starts from k copies of the row's canonical maximum instead, so only
// candidates that beat the resident row end up in its partial list.
template <typename T>
__device__ void heap_load(const KnnArgs<T> &a, T *hd, int *hi, int QS, int q, int j,
                          bool sentinel) {
  const int k = a.k;
  const size_t row = (size_t)heap_row(a, j) * a.ldk;
  const T *Dr = a.D + row;
  const int *Ir = a.I + row;
  if (!sentinel) {
    for (int t = 0; t < k; ++t) {
      hd[t * QS + q] = Dr[t];
      hi[t * QS + q] = Ir[t];
    }
    for (int s = k / 2 - 1; s >= 0; --s) sift_down(hd, hi, QS, q, s, k, hd[s * QS + q], hi[s * QS + q]);
  } else {
    T md = Dr[0];
    int mi = Ir[0];
    for (int t = 1; t < k; ++t) {
      T x = Dr[t];
      int xi = Ir[t];
      if (cless(md, mi, x, xi)) {
        md = x;
        mi = xi;
      }
    }
    for (int t = 0; t < k; ++t) {
      hd[t * QS + q] = md;
      hi[t * QS + q] = mi;
    }
  }
}

// Heapsort the LDS heap ascending in place, then write it out: descending
// into the resident row, or ascending into the partial list of a split
// (query-major lists: each thread writes its k entries contiguously).
template <typename T>
__device__ void heap_store(const KnnArgs<T> &a, T *hd, int *hi, int QS, int q, int j, int split) {
  const int k = a.k;
  for (int e = k - 1; e > 0; --e) {
    T xd = hd[e * QS + q];
    int xi = hi[e * QS + q];
    hd[e * QS + q] = hd[q];
    hi[e * QS + q] = hi[q];
    sift_down(hd, hi, QS, q, 0, e, xd, xi);
  }
  if (a.nsplit == 1) {
    const size_t row = (size_t)heap_row(a, j) * a.ldk;
    for (int t = 0; t < k; ++t) {
      a.D[row + t] = hd[(k - 1 - t) * QS + q];
      a.I[row + t] = hi[(k - 1 - t) * QS + q];
    }
  } else {
    for (int t = 0; t < k; ++t) {
      const size_t o = a.part_qmajor ? ((size_t)j * a.nsplit + split) * k + t
                                     : ((size_t)split * k + t) * a.n + j;
      a.partD[o] = hd[t * QS + q];
      a.partI[o] = hi[t * QS + q];
    }
  }
}

// ---------------------------------------------------------------------------
// Per-lane survivor queue for thread-per-query heaps.  A candidate that beats
// the lane's root as of the last flush waits in the queue; when some lane's
// queue would overflow, every lane applies its queue (and the group at hand)
// to its own heap at once, so heap updates run side by side instead of one
// divergent lane at a time.  A stale root only lets more candidates through;
// each is tested exactly against the live root when applied.  The heap is
// ordered canonically, so its final contents do not depend on the order of
// application.
// ---------------------------------------------------------------------------
template <typename T, int QR>
struct LaneQueue {
  T d[QR];
  int i[QR];
  int n;
};

template <typename T, int QR>
__device__ __forceinline__ void queue_init(LaneQueue<T, QR> &Q) {
#pragma unroll
  for (int r = 0; r < QR; ++r) {
    Q.d[r] = T(0);
    Q.i[r] = 0;
  }
  Q.n = 0;
}

template <typename T, int QR, int TC>
__device__ __forceinline__ void queue_apply(LaneQueue<T, QR> &Q, const T (&dist)[TC],
                                            const int (&id)[TC], unsigned pass, T *hd, int *hi,
                                            int QS, int q, int k, T &rd, int &ri) {
#pragma unroll
  for (int r = 0; r < QR; ++r)
    if (r < Q.n && cless(Q.d[r], Q.i[r], hd[q], hi[q])) sift_down(hd, hi, QS, q, 0, k, Q.d[r], Q.i[r]);
#pragma unroll
  for (int u = 0; u < TC; ++u)
    if (((pass >> u) & 1u) && cless(dist[u], id[u], hd[q], hi[q]))
      sift_down(hd, hi, QS, q, 0, k, dist[u], id[u]);
  Q.n = 0;
  rd = hd[q];
  ri = hi[q];
}

// Offer a group of TC distances (the first cnt valid) of lane q.  Every lane
// of the wave must call it (the overflow test is wave-wide).
template <typename T, int QR, int TC>
__device__ __forceinline__ void queue_offer(LaneQueue<T, QR> &Q, const T (&dist)[TC],
                                            const int (&id)[TC], int cnt, bool valid, T *hd,
                                            int *hi, int QS, int q, int k, T &rd, int &ri) {
  unsigned pass = 0u;
#pragma unroll
  for (int u = 0; u < TC; ++u)
    if (valid && u < cnt && cless(dist[u], id[u], rd, ri)) pass |= 1u << u;
  if (__any(Q.n + __builtin_popcount(pass) > QR)) {
    if (valid) queue_apply<T, QR, TC>(Q, dist, id, pass, hd, hi, QS, q, k, rd, ri);
  } else {
#pragma unroll
    for (int u = 0; u < TC; ++u) {
      if ((pass >> u) & 1u) {
#pragma unroll
        for (int r = 0; r < QR; ++r)
          if (Q.n == r) {
            Q.d[r] = dist[u];
            Q.i[r] = id[u];
          }
        ++Q.n;
      }
    }
  }
}

// Apply what is left in the queues (every lane of the wave calls it).
template <typename T, int QR>
__device__ __forceinline__ void queue_drain(LaneQueue<T, QR> &Q, bool valid, T *hd, int *hi,
                                            int QS, int q, int k, T &rd, int &ri) {
  if (__any(Q.n > 0)) {
    const T nd[1] = {T(0)};
    const int ni[1] = {0};
    if (valid) queue_apply<T, QR, 1>(Q, nd, ni, 0u, hd, hi, QS, q, k, rd, ri);
  }
}

constexpr int kLaneQueue = 8;

// ---------------------------------------------------------------------------
// Exact kernel: one thread per query, the reference oracle's arithmetic.
// LDS: heap [k][QPB] | query rows [QPB][d+1] | candidate tile [d][CT] |
//      candidate norms [CT] | candidate ids [CT]
// ---------------------------------------------------------------------------
template <typename T>
struct ExactLayout {
  size_t hd, hi, sq, sc, cn, ci, bytes;
  // qreg: query rows live in registers (no LDS copy).
  __host__ __device__ ExactLayout(int k, int QPB, int d, int CT, bool qreg = false) {
    hd = 0;
    hi = al16(hd + sizeof(T) * (size_t)k * QPB);
    sq = al16(hi + sizeof(int) * (size_t)k * QPB);
    sc = al16(sq + (qreg ? 0 : sizeof(T) * (size_t)QPB * (d + 1)));
    cn = al16(sc + sizeof(T) * (size_t)d * CT);
    ci = al16(cn + sizeof(T) * (size_t)CT);
    bytes = al16(ci + sizeof(int) * (size_t)CT);
  }
};

// One workgroup's share of a launch: queries [qblock*QPB, +QPB) against the
// candidates of split `split`.  Every thread of the workgroup must call it.
// DQ > 0 (requires d <= DQ): each thread keeps its query row in registers and
// scores 8 candidates per pass; DQ = 0 reads the row from LDS.  Either way a
// pair's dot product is summed in p order with products and sums rounded
// separately, the oracle's arithmetic.
template <typename T, int QPB, int DQ = 0>
__device__ void exact_block(const KnnArgs<T> &a, int CT, int qblock, int split,
                            unsigned char *smem) {
  const int k = a.k, d = a.d;
  const ExactLayout<T> L(k, QPB, d, CT, DQ > 0);
  T *hd = (T *)(smem + L.hd);
  int *hi = (int *)(smem + L.hi);
  T *sq = (T *)(smem + L.sq);
  T *sc = (T *)(smem + L.sc);
  T *cn = (T *)(smem + L.cn);
  int *ci = (int *)(smem + L.ci);

  const int q = threadIdx.x;
  const int j = qblock * QPB + q;
  const bool valid = j < a.n;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);

  T qn = T(0);
  T rd = T(0);
  int ri = 0;
  T *myq = sq + (size_t)q * (d + 1);
  T qreg[DQ > 0 ? DQ : 1];
#pragma unroll
  for (int p = 0; p < (DQ > 0 ? DQ : 1); ++p) qreg[p] = T(0);
  if (valid) {
    const T *xq = a.XB + (size_t)a.bmap[j] * d;
    if (DQ > 0) {
#pragma unroll
      for (int p = 0; p < (DQ > 0 ? DQ : 1); ++p) qreg[p] = p < d ? xq[p] : T(0);
    } else {
      for (int p = 0; p < d; ++p) myq[p] = xq[p];
    }
    qn = a.XB2[a.bmap[j]];
    heap_load(a, hd, hi, QPB, q, j, a.nsplit > 1);
    rd = hd[q];
    ri = hi[q];
  }

  // Cs *= -2; Cs += XA2[alpha[i]]; Cs += XB2[beta[j]], then the lane queue.
  // Every lane takes part (the queue's overflow test is wave-wide); lanes
  // without a query compute on stale rows and offer nothing.
  LaneQueue<T, kLaneQueue> Q;
  queue_init(Q);
  auto offer = [&](auto tc, const T *s, int c) __attribute__((always_inline)) {
    constexpr int TC = decltype(tc)::value;
    T dist[TC];
    int id[TC];
#pragma unroll
    for (int u = 0; u < TC; ++u) {
      dist[u] = xadd(xadd(xmul(s[u], T(-2)), cn[c + u]), qn);
      id[u] = ci[c + u];
    }
    queue_offer<T, kLaneQueue, TC>(Q, dist, id, TC, valid, hd, hi, QPB, q, k, rd, ri);
  };
  typedef std::integral_constant<int, 1> I1;
  typedef std::integral_constant<int, 4> I4;
  typedef std::integral_constant<int, 8> I8;

  for (int c0 = cbeg; c0 < cend; c0 += CT) {
    const int cnt = min(CT, cend - c0);
    __syncthreads();
    for (int e = q; e < cnt * d; e += QPB) {
      const int c = e / d, p = e - c * d;
      sc[p * CT + c] = a.XA[(size_t)a.amap[c0 + c] * d + p];
    }
    for (int e = q; e < cnt; e += QPB) {
      const int id = a.amap[c0 + e];
      cn[e] = a.XA2[id];
      ci[e] = id;
    }
    __syncthreads();
    int c = 0;
    if constexpr (DQ > 0) {
      for (; c + 8 <= cnt; c += 8) {
        T s[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] = T(0);
#pragma unroll
        for (int p = 0; p < DQ; ++p) {
          if (p < d) {
            const T qp = qreg[p];
            const T *cp = sc + p * CT + c;
#pragma unroll
            for (int u = 0; u < 8; ++u) s[u] = xadd(s[u], xmul(cp[u], qp));
          }
        }
        offer(I8{}, s, c);
      }
      for (; c < cnt; ++c) {
        T s0[1] = {T(0)};
#pragma unroll
        for (int p = 0; p < DQ; ++p)
          if (p < d) s0[0] = xadd(s0[0], xmul(sc[p * CT + c], qreg[p]));
        offer(I1{}, s0, c);
      }
      continue;
    }
    for (; c + 4 <= cnt; c += 4) {
      T s[4] = {T(0), T(0), T(0), T(0)};
      for (int p = 0; p < d; ++p) {
        const T qp = myq[p];
        const T *cp = sc + p * CT + c;
        s[0] = xadd(s[0], xmul(cp[0], qp));
        s[1] = xadd(s[1], xmul(cp[1], qp));
        s[2] = xadd(s[2], xmul(cp[2], qp));
        s[3] = xadd(s[3], xmul(cp[3], qp));
      }
      offer(I4{}, s, c);
    }
    for (; c < cnt; ++c) {
      T s[1] = {T(0)};
      for (int p = 0; p < d; ++p) s[0] = xadd(s[0], xmul(sc[p * CT + c], myq[p]));
      offer(I1{}, s, c);
    }
  }
  queue_drain(Q, valid, hd, hi, QPB, q, k, rd, ri);
  if (valid) heap_store(a, hd, hi, QPB, q, j, split);
}


// ---------------------------------------------------------------------------
// Exact block routine, scalar-candidate form (d == D in {8, 16, 32, 64}).
// Same work, arithmetic and result as exact_block; the query row sits in
// registers and candidates are walked by the wave in groups of kScalarTC,
// their ids, coordinates and norms read with scalar loads and used as SGPR
// operands, so each pair-dimension is one v_mul plus one v_add and
// candidates cost no LDS or vector-memory traffic.  LDS holds only the heaps
// (ExactScalarLayout).
// ---------------------------------------------------------------------------
constexpr int kScalarTC = 4;

__host__ __device__ constexpr bool scalar_dim(int d) {
  return d == 8 || d == 16 || d == 32 || d == 64;
}

template <typename T>
struct ExactScalarLayout {
  size_t hd, hi, bytes;
  __host__ __device__ ExactScalarLayout(int k, int QPB) {
    hd = 0;
    hi = al16(sizeof(T) * (size_t)k * QPB);
    bytes = al16(hi + sizeof(int) * (size_t)k * QPB);
  }
};

template <typename T, int QPB, int D, bool FULL>
__device__ __forceinline__ void exact_group_s(sptr<int> am, sptr<T> xs, sptr<T> x2, int c0,
                                              int cend, const T (&qreg)[D], T qn, bool valid,
                                              LaneQueue<T, kLaneQueue> &Q, T *hd, int *hi, int q,
                                              int k, T &rd, int &ri) {
  constexpr int TC = kScalarTC;
  constexpr int P0 = 64 / (int)sizeof(T);  // one s_load_dwordx16 per candidate
  constexpr int P = P0 < D ? P0 : D;
  int id[TC];
#pragma unroll
  for (int u = 0; u < TC; ++u) id[u] = am[FULL ? c0 + u : min(c0 + u, cend - 1)];
  T s[TC];
#pragma unroll
  for (int u = 0; u < TC; ++u) s[u] = T(0);
#pragma unroll
  for (int p0 = 0; p0 < D; p0 += P) {
    const sptr<T> base = p0 == 0 ? xs : after<T, TC>(xs, s);
    T cv[TC][P];
#pragma unroll
    for (int u = 0; u < TC; ++u)
#pragma unroll
      for (int i = 0; i < P; ++i) cv[u][i] = base[(size_t)id[u] * D + p0 + i];
    // Cs += As[i,p] * Bs[j,p], in p order, product and sum rounded separately
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
      for (int u = 0; u < TC; ++u) s[u] = xadd(s[u], xmul(cv[u][i], qreg[p0 + i]));
  }
  // Cs *= -2; Cs += XA2[alpha[i]]; Cs += XB2[beta[j]]
  T dist[TC];
#pragma unroll
  for (int u = 0; u < TC; ++u) dist[u] = xadd(xadd(xmul(s[u], T(-2)), x2[id[u]]), qn);
  queue_offer<T, kLaneQueue, TC>(Q, dist, id, FULL ? TC : cend - c0, valid, hd, hi, QPB, q, k,
                                 rd, ri);
}

template <typename T, int QPB, int D>
__device__ void exact_block_s(const KnnArgs<T> &a, int qblock, int split, unsigned char *smem) {
  const int k = a.k;
  const ExactScalarLayout<T> L(k, QPB);
  T *hd = (T *)(smem + L.hd);
  int *hi = (int *)(smem + L.hi);
  const int q = threadIdx.x;
  const int j = qblock * QPB + q;
  const bool valid = j < a.n;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);

  T qreg[D];
  T qn = T(0), rd = T(0);
  int ri = 0;
  {
    const T *xq = a.XB + (size_t)(valid ? a.bmap[j] : 0) * D;
#pragma unroll
    for (int p = 0; p < D; ++p) qreg[p] = valid ? xq[p] : T(0);
  }
  if (valid) {
    qn = a.XB2[a.bmap[j]];
    heap_load(a, hd, hi, QPB, q, j, a.nsplit > 1);
    rd = hd[q];
    ri = hi[q];
  }
  const sptr<int> am = scalar_view(a.amap);
  const sptr<T> xs = scalar_view(a.XA);
  const sptr<T> x2 = scalar_view(a.XA2);
  LaneQueue<T, kLaneQueue> Q;
  queue_init(Q);
  int c0 = cbeg;
  for (; c0 + kScalarTC <= cend; c0 += kScalarTC)
    exact_group_s<T, QPB, D, true>(am, xs, x2, c0, cend, qreg, qn, valid, Q, hd, hi, q, k, rd, ri);
  if (c0 < cend)
    exact_group_s<T, QPB, D, false>(am, xs, x2, c0, cend, qreg, qn, valid, Q, hd, hi, q, k, rd,
                                    ri);
  queue_drain(Q, valid, hd, hi, QPB, q, k, rd, ri);
  if (valid) heap_store(a, hd, hi, QPB, q, j, split);
}

template <typename T, int QPB, int D>
__global__ __launch_bounds__(QPB) void knn_exact_s_kernel(KnnArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  exact_block_s<T, QPB, D>(a, blockIdx.x, blockIdx.y, smem);
}

// Query rows in registers when d <= kExactDQ.
constexpr int kExactDQ = 32;

template <typename T, int QPB, int DQ>
__global__ __launch_bounds__(QPB) void knn_exact_kernel(KnnArgs<T> a, int CT) {
  extern __shared__ __align__(16) unsigned char smem[];
  exact_block<T, QPB, DQ>(a, CT, blockIdx.x, blockIdx.y, smem);
}


}  // namespace gsknn
