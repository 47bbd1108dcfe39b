// L1-norm (KNN_1NORM) exact kNN on the vector ALU of gfx950.
//
// Reference behaviour (ChenhanYu/rnn): the FP64 kernel tables register L1
// kernels at index KNN_1NORM (kernels/x86_64/sandybridge/gsknn_kernel.h:
// 125-163), rnn_rank_k_abs_int_d8x4.c for the leading dimension blocks and
// rnn_r_1norm_int_d8x4_row.c for the last one; dgsknn_var1 drives them
// (frame/gsknn.c:996-1086).  Per pair, |q_p - c_p| is summed in p order
// within blocks of DKNN_KC = 256 dimensions, each block sum starting from
// zero and the block sums added in order; no norms are involved; the row heap
// select is that of the 2-norm.  Ties go to the lower index (canonical order).
//
// Work layout: one thread per query, candidates walked by the whole
// workgroup in groups of TC.  A candidate's id and coordinates are the same
// for every lane, so they are read with scalar loads into SGPRs: each
// |q_p - c_p| is one v_add_f64 with a negated scalar operand and each sum
// step one v_add_f64 with an |x| source modifier, and candidates cost no LDS
// or vector-memory traffic.  The query row sits in registers for d = 8, 16,
// 32, 64 (D = d) and in LDS, one column per thread, otherwise (D = 0).
// Each thread's heap lives in LDS, as in the exact kernel, fed through a
// per-lane survivor queue; split launches write ascending partial lists that
// the merge kernels fold into the rows.
#include <limits.h>

#include "gsknn_device.h"

namespace gsknn {
namespace {

constexpr int kL1Block = 256;  // DKNN_KC: dimensions per summation block
constexpr int kL1TC = 4;       // candidates per group

// Dimensions per scalar chunk: 64 bytes of each candidate row (one
// s_load_dwordx16 per candidate) in the LDS-query path; the register path
// reads half chunks, two buffers deep.
template <typename T>
constexpr int l1_chunk() {
  return 64 / (int)sizeof(T);
}

template <typename T>
struct L1Layout {
  size_t hd, hi, sq, bytes;
  // dq = 0: query rows in LDS, [d][QPB].
  __host__ __device__ L1Layout(int k, int qpb, int d, int dq) {
    hd = 0;
    hi = al16(hd + sizeof(T) * (size_t)k * qpb);
    sq = al16(hi + sizeof(int) * (size_t)k * qpb);
    bytes = al16(sq + (dq == 0 ? sizeof(T) * (size_t)d * qpb : 0));
  }
};

// Candidate ids of the group starting at c0.  FULL: all TC candidates are
// below cend; otherwise the last one is repeated (its results are unused).
template <int TC, bool FULL>
__device__ __forceinline__ void group_ids(sptr<int> am, int c0, int cend, int (&id)[TC]) {
#pragma unroll
  for (int u = 0; u < TC; ++u) id[u] = am[FULL ? c0 + u : min(c0 + u, cend - 1)];
}

// acc[u] += |q_i - c_{u,i}| over one chunk of P dimensions, in p order.
template <typename T, int TC, int P>
__device__ __forceinline__ void chunk_sum(const T *qv, const T (&cv)[TC][P], T (&acc)[TC]) {
#pragma unroll
  for (int i = 0; i < P; ++i)
#pragma unroll
    for (int u = 0; u < TC; ++u) acc[u] = xadd(acc[u], fabs(xsub(qv[i], cv[u][i])));
}

// d == D (8..64): the query row in registers, one summation block, no
// partial chunks.
template <typename T, int QPB, int D, int TC, bool FULL>
__device__ __forceinline__ void group_reg(sptr<int> am, sptr<T> xs, int c0, int cend,
                                          const T (&qreg)[D], bool valid,
                                          LaneQueue<T, kLaneQueue> &Q, T *hd, int *hi, int q,
                                          int k, T &rd, int &ri) {
  // Half chunks (32 bytes of each candidate row) in two buffers: the loads
  // of chunk c + 2 wait for chunk c's sums (issued all at once they would
  // need TC * D * sizeof(T) / 4 SGPRs) and are in flight while chunk c + 1
  // is summed.
  constexpr int P0 = l1_chunk<T>() < D ? l1_chunk<T>() : D;
  constexpr int P = P0 / 2;
  constexpr int NC = D / P;
  int id[TC];
  group_ids<TC, FULL>(am, c0, cend, id);
  T tot[TC];
#pragma unroll
  for (int u = 0; u < TC; ++u) tot[u] = T(0);
  T cv[2][TC][P];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int u = 0; u < TC; ++u)
#pragma unroll
      for (int i = 0; i < P; ++i) cv[c][u][i] = xs[(size_t)id[u] * D + c * P + i];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    chunk_sum<T, TC, P>(qreg + c * P, cv[c & 1], tot);
    if (c + 2 < NC) {
      const sptr<T> base = after<T, TC>(xs, tot);
#pragma unroll
      for (int u = 0; u < TC; ++u)
#pragma unroll
        for (int i = 0; i < P; ++i) cv[c & 1][u][i] = base[(size_t)id[u] * D + (c + 2) * P + i];
    }
  }
  queue_offer<T, kLaneQueue, TC>(Q, tot, id, FULL ? TC : cend - c0, valid, hd, hi, QPB, q, k,
                                 rd, ri);
}

// Any d: the query row in this thread's LDS column; blocks of kL1Block
// dimensions, the partial chunk at the end of the row loaded element-wise.
template <typename T, int QPB, int TC, bool FULL>
__device__ __forceinline__ void group_lds(sptr<int> am, sptr<T> xs, int c0, int cend, int d,
                                          const T *sq, bool valid, LaneQueue<T, kLaneQueue> &Q,
                                          T *hd, int *hi, int q, int k, T &rd, int &ri) {
  constexpr int P = l1_chunk<T>();
  int id[TC];
  group_ids<TC, FULL>(am, c0, cend, id);
  T tot[TC];
  for (int b0 = 0; b0 < d; b0 += kL1Block) {
    const int be = min(d, b0 + kL1Block);
    T blk[TC];
#pragma unroll
    for (int u = 0; u < TC; ++u) blk[u] = T(0);
    int p0 = b0;
    for (; p0 + P <= be; p0 += P) {
      const sptr<T> base = after<T, TC>(xs, blk);
      T cv[TC][P], qv[P];
#pragma unroll
      for (int u = 0; u < TC; ++u)
#pragma unroll
        for (int i = 0; i < P; ++i) cv[u][i] = base[(size_t)id[u] * d + p0 + i];
#pragma unroll
      for (int i = 0; i < P; ++i) qv[i] = sq[(size_t)(p0 + i) * QPB + q];
      chunk_sum<T, TC, P>(qv, cv, blk);
    }
    for (; p0 < be; ++p0) {
      const T qp = sq[(size_t)p0 * QPB + q];
#pragma unroll
      for (int u = 0; u < TC; ++u) blk[u] = xadd(blk[u], fabs(xsub(qp, xs[(size_t)id[u] * d + p0])));
    }
#pragma unroll
    for (int u = 0; u < TC; ++u) tot[u] = b0 == 0 ? blk[u] : xadd(tot[u], blk[u]);
  }
  queue_offer<T, kLaneQueue, TC>(Q, tot, id, FULL ? TC : cend - c0, valid, hd, hi, QPB, q, k,
                                 rd, ri);
}

// D > 0: d == D with the query row in registers; D == 0: any d.
template <typename T, int QPB, int D>
__global__ __launch_bounds__(QPB) void knn_l1_kernel(KnnArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int TC = kL1TC;
  const int k = a.k, d = a.d;
  const L1Layout<T> L(k, QPB, d, D);
  T *hd = (T *)(smem + L.hd);
  int *hi = (int *)(smem + L.hi);
  T *sq = (T *)(smem + L.sq);

  const int q = threadIdx.x;
  const int j = blockIdx.x * QPB + q;
  const int split = blockIdx.y;
  const bool valid = j < a.n;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);

  // Query row (zeros for threads past n).  No other thread reads a thread's
  // LDS column, so no barrier is needed.
  T qreg[D > 0 ? D : 1];
  {
    const T *xq = a.XB + (size_t)(valid ? a.bmap[j] : 0) * d;
    if constexpr (D > 0) {
#pragma unroll
      for (int p = 0; p < D; ++p) qreg[p] = valid ? xq[p] : T(0);
    } else {
      for (int p = 0; p < d; ++p) sq[(size_t)p * QPB + q] = valid ? xq[p] : T(0);
    }
  }
  T rd = T(0);
  int ri = 0;
  if (valid) {
    heap_load(a, hd, hi, QPB, q, j, a.nsplit > 1);
    rd = hd[q];
    ri = hi[q];
  }

  const sptr<int> am = scalar_view(a.amap);
  const sptr<T> xs = scalar_view(a.XA);
  LaneQueue<T, kLaneQueue> Q;
  queue_init(Q);
  int c0 = cbeg;
  for (; c0 + TC <= cend; c0 += TC) {
    if constexpr (D > 0)
      group_reg<T, QPB, D, TC, true>(am, xs, c0, cend, qreg, valid, Q, hd, hi, q, k, rd, ri);
    else
      group_lds<T, QPB, TC, true>(am, xs, c0, cend, d, sq, valid, Q, hd, hi, q, k, rd, ri);
  }
  if (c0 < cend) {
    if constexpr (D > 0)
      group_reg<T, QPB, D, TC, false>(am, xs, c0, cend, qreg, valid, Q, hd, hi, q, k, rd, ri);
    else
      group_lds<T, QPB, TC, false>(am, xs, c0, cend, d, sq, valid, Q, hd, hi, q, k, rd, ri);
  }
  queue_drain(Q, valid, hd, hi, QPB, q, k, rd, ri);
  if (valid) heap_store(a, hd, hi, QPB, q, j, split);
}

constexpr size_t kLdsMax = 160 * 1024;
constexpr int kCUs = 256;

// Register paths for the common dimensions; every other d takes the LDS path.
int l1_dq(int d) { return (d == 8 || d == 16 || d == 32 || d == 64) ? d : 0; }

struct L1Plan {
  int qpb = 0, dq = 0, gx = 0, nsplit = 1, cps = 0;
  size_t lds = 0;
};

template <typename T>
const void *l1_kernel_ptr(int qpb, int dq) {
#define GSKNN_L1_PTR(QPB_)                                                       \
  if (qpb == QPB_) {                                                             \
    switch (dq) {                                                                \
      case 8: return (const void *)knn_l1_kernel<T, QPB_, 8>;                    \
      case 16: return (const void *)knn_l1_kernel<T, QPB_, 16>;                  \
      case 32: return (const void *)knn_l1_kernel<T, QPB_, 32>;                  \
      case 64: return (const void *)knn_l1_kernel<T, QPB_, 64>;                  \
      default: return (const void *)knn_l1_kernel<T, QPB_, 0>;                   \
    }                                                                            \
  }
  GSKNN_L1_PTR(64)
  GSKNN_L1_PTR(32)
  GSKNN_L1_PTR(16)
#undef GSKNN_L1_PTR
  return nullptr;
}

template <typename T>
L1Plan plan_l1(int m, int n, int d, int k) {
  L1Plan p;
  p.dq = l1_dq(d);
  for (int qpb : {64, 32, 16}) {
    const size_t b = L1Layout<T>(k, qpb, d, p.dq).bytes;
    if (b <= kLdsMax) {
      p.qpb = qpb;
      p.lds = b;
      break;
    }
  }
  if (p.qpb == 0) return p;
  p.gx = (n + p.qpb - 1) / p.qpb;
  // Resident workgroups per CU (registers and LDS), then enough candidate
  // splits for two full rounds of the chip; each split keeps >= 16 groups.
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, l1_kernel_ptr<T>(p.qpb, p.dq), p.qpb,
                                                   p.lds) != hipSuccess ||
      per_cu <= 0)
    per_cu = (int)(kLdsMax / p.lds);
  (void)hipGetLastError();
  const long long target = 2LL * kCUs * per_cu;
  int ns = 1;
  if (p.gx < target && knn_merge_supported(k, (int)sizeof(T))) {
    ns = (int)((target + p.gx - 1) / p.gx);
    const int maxs = m / (16 * kL1TC);
    if (ns > maxs) ns = maxs;
    if (ns < 1) ns = 1;
  }
  if (ns > 1) {
    int cps = (m + ns - 1) / ns;
    cps = (cps + kL1TC - 1) / kL1TC * kL1TC;
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

size_t l1_part_bytes(const L1Plan &p, int n, int k, int elem) {
  if (p.nsplit <= 1) return 0;
  const size_t ents = (size_t)p.nsplit * k * n;
  return al16(ents * elem) + al16(ents * sizeof(int));
}

template <typename T>
hipError_t launch_l1(KnnArgs<T> a, hipStream_t s, void *work, size_t work_bytes) {
  if (a.m <= 0 || a.n <= 0 || a.d <= 0 || a.k <= 0) return hipSuccess;
  const L1Plan p = plan_l1<T>(a.m, a.n, a.d, a.k);
  if (p.qpb == 0) return hipErrorNotSupported;
  a.nsplit = p.nsplit;
  a.cps = p.cps;
  a.part_qmajor = p.nsplit > 1 && knn_merge_query_major(a.n, a.k, p.nsplit);
  if (p.nsplit > 1) {
    const size_t need = l1_part_bytes(p, a.n, a.k, (int)sizeof(T));
    if (work == nullptr || work_bytes < need) return hipErrorInvalidValue;
    const size_t ents = (size_t)p.nsplit * a.k * a.n;
    a.partD = (T *)work;
    a.partI = (int *)((char *)work + al16(ents * sizeof(T)));
  }
  const void *kern = l1_kernel_ptr<T>(p.qpb, p.dq);
  hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
  if (e != hipSuccess) return e;
  void *args[] = {&a};
  e = hipLaunchKernel(kern, dim3(p.gx, p.nsplit), dim3(p.qpb), args, p.lds, s);
  if (e != hipSuccess) return e;
  if (p.nsplit > 1) e = launch_knn_merge(a, s);
  return e;
}

}  // namespace

size_t knn_l1_work_bytes(int m, int n, int d, int k, int elem_bytes) {
  if (m <= 0 || n <= 0 || d <= 0 || k <= 0) return 0;
  const L1Plan p = elem_bytes == 8 ? plan_l1<double>(m, n, d, k) : plan_l1<float>(m, n, d, k);
  return l1_part_bytes(p, n, k, elem_bytes);
}

hipError_t launch_knn_l1_f64(const KnnArgs<double> &a, hipStream_t s, void *work,
                             size_t work_bytes) {
  return launch_l1<double>(a, s, work, work_bytes);
}

hipError_t launch_knn_l1_f32(const KnnArgs<float> &a, hipStream_t s, void *work,
                             size_t work_bytes) {
  return launch_l1<float>(a, s, work, work_bytes);
}

}  // namespace gsknn
