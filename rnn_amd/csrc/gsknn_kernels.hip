// HIP kernels of the exact-kNN path for gfx950 (CDNA4).
//
// Reference behaviour (ChenhanYu/rnn):
//   distance  = (-2 * sum_p a_p b_p + XA2[amap[i]]) + XB2[bmap[j]], no clamp
//               (frame/dgsknn_ref_impl.h:63-72, :91-103)
//   selection = per heap row, keep the k smallest in a max-heap whose root is
//               the current k-th distance (frame/gsknn_heap.c:104-162)
//
// Selection here uses the canonical total order (distance, index): a candidate
// enters a row only when it is strictly smaller than the row's root, so a tie
// in distance goes to the lower index and the result does not depend on the
// order in which candidates are visited.  That is what lets a launch split the
// candidates over workgroups and merge the partial lists afterwards.
//
// Kernels
//   knn_exact_kernel<T>     one thread per query, VALU, the reference oracle's
//                           arithmetic in p order (product and sum rounded
//                           separately): distances bit-identical to it.
//   knn_mfma_f64_kernel     FP64 MFMA (v_mfma_f64_16x16x4_f64) contraction,
//                           wave-owned queries, LDS heaps.
//   knn_merge_kernel<T>     folds the partial lists of a candidate-split
//                           launch into the resident heap rows.
//   sqnorm_kernel<T>        X2[i] = sum_p X[i][p]^2 in p order.
//
// Heaps live in LDS as [entry][query] so that lanes touching the same entry
// of different queries hit different banks.  Rows are written back sorted
// in descending order, which is a valid max-heap.
#include <limits.h>

#include "gsknn_device.h"

namespace gsknn {

// ---------------------------------------------------------------------------
// FP64 MFMA kernel.
//
// A workgroup of W waves owns Q = W * 16 * QB queries; wave w owns 16 * QB of
// them.  Candidates stream through LDS in tiles of C = 16 * CB, double
// buffered, stored in MFMA fragment order so that each A-operand read is 64
// consecutive doubles.  v_mfma_f64_16x16x4_f64 takes A = 16 candidates x 4
// dims (lane l: candidate l%16, dim l/16) and B = 4 dims x 16 queries (lane l:
// dim l/16, query l%16); lane l of the 16x16 result holds query l%16 against
// candidates l/16 + 4*v, v = 0..3 (measured on gfx950).  The query fragments
// stay in registers for the whole launch (d <= 4*KS).
//
// Epilogue per lane: dist = (-2 * dot + XA2[amap[i]]) + XB2[bmap[j]], test
// against the row's root.  A wave-wide ballot skips the insertion step when
// no lane has a survivor, which is the common case once the rows are warm.
// Otherwise the four lane groups (l/16 = 0..3, each covering the 16 queries
// once) insert in turn, so no two lanes ever touch the same heap at once.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

struct MfmaLayout {
  size_t tA, tn, ti, hd, hi, bytes;
  __host__ __device__ MfmaLayout(int KS, int CB, int Q, int k) {
    const size_t tile = (size_t)KS * CB * 64;
    const size_t C = 16 * (size_t)CB;
    tA = 0;
    tn = al16(tA + sizeof(double) * tile);
    ti = al16(tn + sizeof(double) * C);
    hd = al16(ti + sizeof(int) * C);
    hi = al16(hd + sizeof(double) * (size_t)k * Q);
    bytes = al16(hi + sizeof(int) * (size_t)k * Q);
  }
};

// Candidate tiles go global -> registers -> LDS: thread t stages dims
// [p0, p0 + DPT) of candidate slot t / TPC.  The amap entries are read two
// tiles ahead and the point rows one tile ahead, so no load waits on another
// inside the loop and the loads overlap the MFMA work of the current tile.
// One LDS tile buffer (two barriers per tile) keeps the C=32, k=32 workgroup
// at 64 KB so that two workgroups share a CU.
template <int KS, int QB, int CB, int W>
__global__ __launch_bounds__(64 * W) void knn_mfma_f64_kernel(KnnArgs<double> a) {
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB, DP = 4 * KS, NT = 64 * W;
  static_assert(NT % C == 0, "threads per candidate");
  constexpr int TPC = NT / C;
  static_assert(DP % TPC == 0, "dims per thread");
  constexpr int DPT = DP / TPC;
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k, d = a.d;
  const MfmaLayout L(KS, CB, Q, k);
  double *tA = (double *)(smem + L.tA);
  double *tn = (double *)(smem + L.tn);
  int *ti = (int *)(smem + L.ti);
  double *hd = (double *)(smem + L.hd);
  int *hi = (int *)(smem + L.hi);

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, grp = lane >> 4;
  const int qbase = blockIdx.x * Q;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const bool sentinel = a.nsplit > 1;

  if (tid < Q && qbase + tid < a.n) heap_load(a, hd, hi, Q, tid, qbase + tid, sentinel);

  double bq[QB][KS];
  double qn[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = w * QW + qb * 16 + (lane & 15);
    const int j = qbase + ql[qb];
    qv[qb] = j < a.n;
    const int jj = qv[qb] ? j : a.n - 1;
    const double *xq = a.XB + (size_t)a.bmap[jj] * d;
    qn[qb] = qv[qb] ? a.XB2[a.bmap[jj]] : 0.0;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int p = 4 * s + grp;
      const double v = xq[p < d ? p : d - 1];
      bq[qb][s] = (qv[qb] && p < d) ? v : 0.0;
    }
  }

  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  const int cs = tid / TPC;
  const int p0 = (tid % TPC) * DPT;
  double stage[DPT];
  double stage_n = 0.0;
  auto cand_id = [&](int t) -> int {
    const int c = min(cbeg + t * C + cs, cend - 1);
    return a.amap[c];
  };
  auto load_x = [&](int id) {
    const double *row = a.XA + (size_t)id * d;
#pragma unroll
    for (int e = 0; e < DPT; ++e) {
      const int p = p0 + e;
      stage[e] = row[p < d ? p : d - 1];
    }
    stage_n = a.XA2[id];
  };
  auto store_x = [&](int t, int id) {
    const bool ok = cbeg + t * C + cs < cend;
    const int cb = cs >> 4, i = cs & 15;
#pragma unroll
    for (int e = 0; e < DPT; ++e) {
      const int p = p0 + e;
      const int s = p >> 2, kk = p & 3;
      tA[(s * CB + cb) * 64 + i + 16 * kk] = (ok && p < d) ? stage[e] : 0.0;
    }
    if (tid % TPC == 0) {
      // An empty slot gets distance +inf and index INT_MAX: it loses to
      // every row entry, including (+inf, -1) placeholders.
      tn[cs] = ok ? stage_n : __builtin_inf();
      ti[cs] = ok ? id : INT_MAX;
    }
  };

  int id_n = 0;
  if (ntiles > 0) {
    const int id0 = cand_id(0);
    load_x(id0);
    id_n = ntiles > 1 ? cand_id(1) : id0;
    __syncthreads();  // heaps loaded
    store_x(0, id0);
  }
  __syncthreads();
  double rd[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) rd[qb] = qv[qb] ? hd[ql[qb]] : -__builtin_inf();

  constexpr int QR = 4;
  double pq_d[QR];
  int pq_i[QR], pq_q[QR];
  int pq_n = 0;
#pragma unroll
  for (int r = 0; r < QR; ++r) {
    pq_d[r] = 0.0;
    pq_i[r] = 0;
    pq_q[r] = 0;
  }
  // Apply the queued survivors and this tile's (pass bits over acc) to the
  // heaps.  The four lane groups cover the wave's queries once each, so
  // within a group every lane owns a distinct heap; groups take turns.
  auto flush = [&](const unsigned *pass, const dbl4 (*acc)[CB]) {
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
      if (grp == g) {
#pragma unroll
        for (int r = 0; r < QR; ++r) {
          if (r < pq_n) {
            int slot = ql[0];
#pragma unroll
            for (int qb = 1; qb < QB; ++qb)
              if (pq_q[r] == qb) slot = ql[qb];
            if (cless(pq_d[r], pq_i[r], hd[slot], hi[slot]))
              sift_down(hd, hi, Q, slot, 0, k, pq_d[r], pq_i[r]);
          }
        }
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) {
          unsigned bits = pass[qb];
          const int slot = ql[qb];
          while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            double dist = acc[qb][0][0];
#pragma unroll
            for (int bb = 1; bb < CB * 4; ++bb)
              if (b == bb) dist = acc[qb][bb >> 2][bb & 3];
            const int idx = ti[(b >> 2) * 16 + grp + 4 * (b & 3)];
            if (cless(dist, idx, hd[slot], hi[slot])) sift_down(hd, hi, Q, slot, 0, k, dist, idx);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    pq_n = 0;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) rd[qb] = qv[qb] ? hd[ql[qb]] : -__builtin_inf();
  };

  for (int t = 0; t < ntiles; ++t) {
    const bool more = t + 1 < ntiles;
    const int id_nn = (t + 2 < ntiles) ? cand_id(t + 2) : id_n;
    if (more) load_x(id_n);

    dbl4 acc[QB][CB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const double av = tA[(s * CB + cb) * 64 + lane];
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bq[qb][s], acc[qb][cb], 0, 0, 0);
      }
    }

    // Hot path: one FP64 compare per pair against the row's current root;
    // the exact (distance, index) test runs only for the rare survivors.
    unsigned pass[QB];
    bool anyp = false;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      pass[qb] = 0u;
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = cb * 16 + grp + 4 * v;
          // (-2 * dot + XA2[amap[i]]) + XB2[bmap[j]]; -2 * dot is exact.
          const double dist = xadd(__builtin_fma(acc[qb][cb][v], -2.0, tn[c]), qn[qb]);
          acc[qb][cb][v] = dist;
          if (dist <= rd[qb]) {
            pass[qb] |= 1u << (cb * 4 + v);
            anyp = true;
          }
        }
    }
    // Survivors wait in a per-lane queue of QR entries; the heaps (and the
    // roots used by the filter) are updated only when some lane's queue
    // would overflow, which keeps the slow path off most tiles.  A stale
    // root only lets more pairs through the filter; every queued pair is
    // tested exactly against the live root when it is applied.
    int npass = 0;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) npass += __builtin_popcount(pass[qb]);
    if (__any(pq_n + npass > QR)) {
      flush(pass, acc);
    } else if (__any(anyp)) {
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        unsigned bits = pass[qb];
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1;
          double dist = acc[qb][0][0];
#pragma unroll
          for (int bb = 1; bb < CB * 4; ++bb)
            if (b == bb) dist = acc[qb][bb >> 2][bb & 3];
          const int idx = ti[(b >> 2) * 16 + grp + 4 * (b & 3)];
#pragma unroll
          for (int r = 0; r < QR; ++r)
            if (pq_n == r) {
              pq_d[r] = dist;
              pq_i[r] = idx;
              pq_q[r] = qb;
            }
          ++pq_n;
        }
      }
    }
    __syncthreads();  // tile consumed
    if (more) store_x(t + 1, id_n);
    __syncthreads();  // next tile visible
    id_n = id_nn;
  }

  {
    const unsigned none[QB] = {};
    const dbl4 zero[QB][CB] = {};
    if (__any(pq_n > 0)) flush(none, zero);
  }
  __syncthreads();
  if (tid < Q && qbase + tid < a.n) heap_store(a, hd, hi, Q, tid, qbase + tid, split);
}

// ---------------------------------------------------------------------------
// Merge of split partial lists into the resident rows: one thread per query.
// Each partial list is ascending, so the scan of a list stops at its first
// entry that does not beat the root.  Sentinel entries (copies of the
// resident maximum) never beat the root and are skipped by that rule.
// ---------------------------------------------------------------------------
template <typename T, int QPB>
__global__ __launch_bounds__(QPB) void knn_merge_kernel(KnnArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k;
  T *hd = (T *)smem;
  int *hi = (int *)(smem + al16(sizeof(T) * (size_t)k * QPB));
  const int q = threadIdx.x;
  const int j = blockIdx.x * QPB + q;
  if (j >= a.n) return;
  heap_load(a, hd, hi, QPB, q, j, false);
  T rd = hd[q];
  int ri = hi[q];
  for (int s = 0; s < a.nsplit; ++s) {
    for (int t = 0; t < k; ++t) {
      const size_t o = ((size_t)s * k + t) * a.n + j;
      const T xd = a.partD[o];
      const int xi = a.partI[o];
      if (!cless(xd, xi, rd, ri)) break;
      sift_down(hd, hi, QPB, q, 0, k, xd, xi);
      rd = hd[q];
      ri = hi[q];
    }
  }
  KnnArgs<T> b = a;
  b.nsplit = 1;
  heap_store(b, hd, hi, QPB, q, j, 0);
}

template <typename T>
__global__ void sqnorm_kernel(int N, int d, const T *X, T *X2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const T *x = X + (size_t)i * d;
  T s = T(0);
  for (int p = 0; p < d; ++p) s = xadd(s, xmul(x[p], x[p]));
  X2[i] = s;
}

// dgssq2nrm: one thread per (candidate i, query j), the oracle's arithmetic.
__global__ void sq2nrm_kernel(int m, int n, int d, const double *XA, const double *XA2,
                              const int *amap, const double *XB, const double *XB2,
                              const int *bmap, double *C, int ldc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  const double *a = XA + (size_t)amap[i] * d;
  const double *b = XB + (size_t)bmap[j] * d;
  double s = 0.0;
  for (int p = 0; p < d; ++p) s = xadd(s, xmul(a[p], b[p]));
  C[(size_t)j * ldc + i] = xadd(xadd(xmul(s, -2.0), XA2[amap[i]]), XB2[bmap[j]]);
}

// Header rows (frame/gsknn_heap.c:294-308): D[0] = ro, I[0] = k, I[1] = fill
// count, I[2] = ldk, entries from offset 3.  While a row is filling, its
// filled entries occupy the last `fill` slots (gsknn_heapselect_int_d4.c:
// 66-75).  Unfilled slots become (+inf, INT_MAX) placeholders, which every
// finite candidate beats.
__global__ void hdr_to_rows_kernel(int n, int k, int ldk, const double *Dh, const int *Ih,
                                   double *Dt, int *It) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (t >= k || j >= n) return;
  const size_t h = (size_t)j * ldk;
  const int fill = Ih[h + 1];
  const bool real = fill >= k || t >= k - fill;
  Dt[(size_t)j * k + t] = real ? Dh[h + 3 + t] : __builtin_inf();
  It[(size_t)j * k + t] = real ? Ih[h + 3 + t] : INT_MAX;
}

// Rows come back sorted descending, placeholders first; the real entries go
// to the last `count` slots, a valid 4-ary max-heap behind the header.
__global__ void rows_to_hdr_kernel(int n, int k, int ldk, const double *Dt, const int *It,
                                   double *Dh, int *Ih) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t h = (size_t)j * ldk;
  const double *dt = Dt + (size_t)j * k;
  const int *it = It + (size_t)j * k;
  int count = 0;
  for (int t = 0; t < k; ++t) count += !(it[t] == INT_MAX && dt[t] == __builtin_inf());
  for (int t = k - count; t < k; ++t) {
    Dh[h + 3 + t] = dt[t];
    Ih[h + 3 + t] = it[t];
  }
  Ih[h + 1] = count;
}

// ---------------------------------------------------------------------------
// Host-side planning and launch.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kLdsMax = 160 * 1024;  // gfx950 LDS per workgroup
constexpr int kTargetBlocks = 512;      // 2 workgroups per CU on 256 CUs

struct Plan {
  int kind = 0;  // 0 exact, 1 mfma
  int qpb = 64;  // queries per workgroup
  int ct = 0;    // exact: candidates per tile
  int ks = 0, qb = 0, w = 0;  // mfma
  size_t lds = 0;
  int gx = 0, nsplit = 1, cps = 0;
};

int exact_ct(int d, int elem) {
  int ct = (int)((16384 / elem) / (d > 0 ? d : 1));
  ct = ct < 4 ? 4 : (ct > 64 ? 64 : ct);
  return ct & ~3;
}

int mfma_ks(int d) {
  const int need = (d + 3) / 4;
  for (int ks : {4, 8, 16, 32, 64})
    if (need <= ks) return ks;
  return 0;
}

// Candidate blocks of 16 per tile: two for d <= 64, one above, where the
// query fragments already take most of the registers.
constexpr int mfma_cb(int ks) { return ks >= 32 ? 1 : 2; }

size_t merge_lds(int k, int elem, int qpb) {
  return al16((size_t)elem * k * qpb) + sizeof(int) * (size_t)k * qpb;
}

int merge_qpb(int k, int elem) {
  for (int qpb : {64, 32, 16})
    if (merge_lds(k, elem, qpb) <= kLdsMax) return qpb;
  return 0;
}

Plan make_plan(int m, int n, int d, int k, int elem, int mode) {
  Plan p;
  const bool f64 = elem == 8;
  const int ks = f64 ? mfma_ks(d) : 0;
  const bool mfma = f64 && ks > 0 && (mode == MODE_MFMA || (mode == MODE_AUTO && d >= 32));
  int tile_c = 0;
  if (mfma) {
    p.kind = 1;
    p.ks = ks;
    const int opts[4][2] = {{2, 4}, {1, 4}, {1, 2}, {1, 1}};
    for (auto &o : opts) {
      const int Q = o[1] * 16 * o[0];
      const size_t b = MfmaLayout(ks, mfma_cb(ks), Q, k).bytes;
      if (b <= kLdsMax) {
        p.qb = o[0];
        p.w = o[1];
        p.qpb = Q;
        p.lds = b;
        break;
      }
    }
    if (p.qb == 0) p.kind = 0;  // heap too large for LDS: exact kernel
    tile_c = 16 * mfma_cb(ks);
  }
  if (p.kind == 0) {
    p.ct = exact_ct(d, elem);
    p.qpb = 0;
    for (int qpb : {64, 32, 16}) {
      const size_t b = f64 ? ExactLayout<double>(k, qpb, d, p.ct).bytes
                           : ExactLayout<float>(k, qpb, d, p.ct).bytes;
      if (b <= kLdsMax) {
        p.qpb = qpb;
        p.lds = b;
        break;
      }
    }
    tile_c = p.ct;
  }
  if (p.qpb == 0) return p;
  p.gx = (n + p.qpb - 1) / p.qpb;
  // Split the candidates when the query blocks alone cannot fill the chip;
  // each split keeps at least 8 tiles of work.
  int ns = 1;
  if (p.gx < kTargetBlocks && merge_qpb(k, elem) > 0) {
    ns = (kTargetBlocks + p.gx - 1) / p.gx;
    const int maxs = m / (8 * tile_c);
    if (ns > maxs) ns = maxs;
    if (ns < 1) ns = 1;
  }
  if (ns > 1) {
    int cps = (m + ns - 1) / ns;
    cps = (cps + tile_c - 1) / tile_c * tile_c;
    ns = (m + cps - 1) / cps;
    p.cps = cps;
    p.nsplit = ns;
  }
  if (p.nsplit <= 1) {
    p.nsplit = 1;
    p.cps = m;
  }
  return p;
}

template <typename K>
hipError_t set_lds(K kernel, size_t bytes) {
  return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

template <int KS>
hipError_t launch_mfma_ks(const Plan &p, const KnnArgs<double> &a, hipStream_t s) {
  const dim3 grid(p.gx, p.nsplit);
  hipError_t e = hipSuccess;
#define GSKNN_MFMA_CASE(QB_, W_)                                              \
  if (p.qb == QB_ && p.w == W_) {                                             \
    auto kern = knn_mfma_f64_kernel<KS, QB_, mfma_cb(KS), W_>;                \
    if ((e = set_lds(kern, p.lds)) != hipSuccess) return e;                   \
    hipLaunchKernelGGL(kern, grid, dim3(64 * W_), p.lds, s, a);               \
    return hipGetLastError();                                                 \
  }
  GSKNN_MFMA_CASE(2, 4)
  GSKNN_MFMA_CASE(1, 4)
  GSKNN_MFMA_CASE(1, 2)
  GSKNN_MFMA_CASE(1, 1)
#undef GSKNN_MFMA_CASE
  return hipErrorNotSupported;
}

template <typename T>
hipError_t launch_exact(const Plan &p, const KnnArgs<T> &a, hipStream_t s) {
  const dim3 grid(p.gx, p.nsplit);
  hipError_t e = hipSuccess;
#define GSKNN_EXACT_CASE(QPB_)                                                \
  if (p.qpb == QPB_) {                                                        \
    auto kern = knn_exact_kernel<T, QPB_>;                                    \
    if ((e = set_lds(kern, p.lds)) != hipSuccess) return e;                   \
    hipLaunchKernelGGL(kern, grid, dim3(QPB_), p.lds, s, a, p.ct);            \
    return hipGetLastError();                                                 \
  }
  GSKNN_EXACT_CASE(64)
  GSKNN_EXACT_CASE(32)
  GSKNN_EXACT_CASE(16)
#undef GSKNN_EXACT_CASE
  return hipErrorNotSupported;
}

template <typename T>
hipError_t launch_merge(const KnnArgs<T> &a, hipStream_t s) {
  const int qpb = merge_qpb(a.k, (int)sizeof(T));
  const size_t lds = merge_lds(a.k, (int)sizeof(T), qpb);
  hipError_t e = hipSuccess;
#define GSKNN_MERGE_CASE(QPB_)                                                \
  if (qpb == QPB_) {                                                          \
    auto kern = knn_merge_kernel<T, QPB_>;                                    \
    if ((e = set_lds(kern, lds)) != hipSuccess) return e;                     \
    hipLaunchKernelGGL(kern, dim3((a.n + QPB_ - 1) / QPB_), dim3(QPB_), lds, s, a); \
    return hipGetLastError();                                                 \
  }
  GSKNN_MERGE_CASE(64)
  GSKNN_MERGE_CASE(32)
  GSKNN_MERGE_CASE(16)
#undef GSKNN_MERGE_CASE
  return hipErrorNotSupported;
}

size_t part_bytes(const Plan &p, int n, int k, int elem) {
  if (p.nsplit <= 1) return 0;
  const size_t ents = (size_t)p.nsplit * k * n;
  return al16(ents * elem) + al16(ents * sizeof(int));
}

// FP32 calls with d >= 8 and at least 64 candidates take the FP32 MFMA
// candidate pass + oracle-arithmetic re-rank; the rest the exact kernel.
bool use_f32_mfma(int m, int d, int k, int mode) {
  return mode != MODE_EXACT && d >= 8 && m >= 64 && f32_mfma_supported(d, k);
}

template <typename T>
hipError_t launch_knn(KnnArgs<T> a, int mode, hipStream_t s, void *work, size_t work_bytes) {
  if (a.m <= 0 || a.n <= 0 || a.d <= 0 || a.k <= 0) return hipSuccess;
  if constexpr (sizeof(T) == 4) {
    if (use_f32_mfma(a.m, a.d, a.k, mode)) return launch_knn_f32_mfma(a, s, work, work_bytes, nullptr);
  }
  const Plan p = make_plan(a.m, a.n, a.d, a.k, (int)sizeof(T), mode);
  if (p.qpb == 0) return hipErrorNotSupported;
  a.nsplit = p.nsplit;
  a.cps = p.cps;
  if (p.nsplit > 1) {
    const size_t need = part_bytes(p, a.n, a.k, (int)sizeof(T));
    if (work == nullptr || work_bytes < need) return hipErrorInvalidValue;
    const size_t ents = (size_t)p.nsplit * a.k * a.n;
    a.partD = (T *)work;
    a.partI = (int *)((char *)work + al16(ents * sizeof(T)));
  }
  hipError_t e;
  if (p.kind == 1) {
    if constexpr (sizeof(T) == 8) {
      switch (p.ks) {
        case 4: e = launch_mfma_ks<4>(p, a, s); break;
        case 8: e = launch_mfma_ks<8>(p, a, s); break;
        case 16: e = launch_mfma_ks<16>(p, a, s); break;
        case 32: e = launch_mfma_ks<32>(p, a, s); break;
        case 64: e = launch_mfma_ks<64>(p, a, s); break;
        default: return hipErrorNotSupported;
      }
    } else {
      return hipErrorNotSupported;
    }
  } else {
    e = launch_exact(p, a, s);
  }
  if (e != hipSuccess) return e;
  if (p.nsplit > 1) e = launch_merge(a, s);
  return e;
}

}  // namespace

size_t knn_work_bytes(int m, int n, int d, int k, int elem_bytes, int mode) {
  if (m <= 0 || n <= 0 || d <= 0 || k <= 0) return 0;
  if (elem_bytes == 4 && use_f32_mfma(m, d, k, mode)) return f32_work_bytes(m, n, d, k);
  const Plan p = make_plan(m, n, d, k, elem_bytes, mode);
  return part_bytes(p, n, k, elem_bytes);
}

hipError_t launch_knn_f64(const KnnArgs<double> &a, int mode, hipStream_t s, void *work,
                          size_t work_bytes) {
  return launch_knn<double>(a, mode, s, work, work_bytes);
}

hipError_t launch_knn_f32(const KnnArgs<float> &a, int mode, hipStream_t s, void *work,
                          size_t work_bytes) {
  return launch_knn<float>(a, mode, s, work, work_bytes);
}

hipError_t launch_sq2nrm_f64(int m, int n, int d, const double *XA, const double *XA2,
                             const int *amap, const double *XB, const double *XB2,
                             const int *bmap, double *C, int ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (n > 65535) {
    for (int j0 = 0; j0 < n; j0 += 65535) {
      const int nn = n - j0 < 65535 ? n - j0 : 65535;
      hipError_t e = launch_sq2nrm_f64(m, nn, d, XA, XA2, amap, XB, XB2, bmap + j0,
                                       C + (size_t)j0 * ldc, ldc, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  hipLaunchKernelGGL(sq2nrm_kernel, dim3((m + 255) / 256, n), dim3(256), 0, s, m, n, d, XA, XA2,
                     amap, XB, XB2, bmap, C, ldc);
  return hipGetLastError();
}

hipError_t launch_hdr_to_rows(int n, int k, int ldk, const double *Dh, const int *Ih,
                              double *Dt, int *It, hipStream_t s) {
  if (n <= 0 || k <= 0) return hipSuccess;
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const int nn = n - j0 < 65535 ? n - j0 : 65535;
    hipLaunchKernelGGL(hdr_to_rows_kernel, dim3((k + 255) / 256, nn), dim3(256), 0, s, nn, k, ldk,
                       Dh + (size_t)j0 * ldk, Ih + (size_t)j0 * ldk, Dt + (size_t)j0 * k,
                       It + (size_t)j0 * k);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_rows_to_hdr(int n, int k, int ldk, const double *Dt, const int *It,
                              double *Dh, int *Ih, hipStream_t s) {
  if (n <= 0 || k <= 0) return hipSuccess;
  hipLaunchKernelGGL(rows_to_hdr_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, k, ldk, Dt, It,
                     Dh, Ih);
  return hipGetLastError();
}

hipError_t launch_sqnorm_f64(int N, int d, const double *X, double *X2, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(sqnorm_kernel<double>, dim3((N + 255) / 256), dim3(256), 0, s, N, d, X, X2);
  return hipGetLastError();
}

hipError_t launch_sqnorm_f32(int N, int d, const float *X, float *X2, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(sqnorm_kernel<float>, dim3((N + 255) / 256), dim3(256), 0, s, N, d, X, X2);
  return hipGetLastError();
}

}  // namespace gsknn
