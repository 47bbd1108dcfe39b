// FP64 MFMA kNN, operand-ring form (gfx950).
//
// Reference behaviour (ChenhanYu/rnn):
//   distance  = (-2 * sum_p a_p b_p + XA2[amap[i]]) + XB2[bmap[j]], no clamp
//               (frame/dgsknn_ref_impl.h:63-72, :91-103)
//   selection = per heap row, keep the k smallest in a max-heap whose root is
//               the current k-th distance (frame/gsknn_heap.c:104-162)
// Selection is by the canonical order (distance, index), as in every kernel
// of this library.
//
// One wave per workgroup owns Q = 16 QB queries and streams the candidates
// of its split in tiles of C = 16 CB.  v_mfma_f64_16x16x4_f64 takes
// A = 16 candidates x 4 dims (lane l: candidate l % 16, k-index l / 16) and
// B = 4 dims x 16 queries; lane l of the result holds query l % 16 against
// candidates l / 16 + 4 v, v = 0..3 (measured on gfx950).
//
// Dims are dealt in pairs of k-steps: k-steps 2j and 2j + 1 of lane group g
// (g = l / 16) carry dims 8j + 2g and 8j + 2g + 1, so one 16-byte load per
// lane and row (8-byte loads when the rows are not 16-byte aligned) brings
// in a pair, and the four groups read 64 contiguous bytes of each of the 16
// rows of a candidate block.  An instantiation has a fixed number KP of
// pairs (d <= 8 KP): the query operands of all of them stay in registers for
// the whole launch, zero past d, and loads past d are moved back inside the
// row.  The candidate operands stream through a ring of R pair slots (R = 4,
// or 2 when KP is not a multiple of 4): right after the MFMAs of pair j read
// slot j % R, the slot is refilled with pair j + R, of this tile or (across
// the tile boundary) of the next one, so each load has R pairs of MFMAs to
// arrive and only R pairs of candidate operands are ever in registers.
// KP = 0 is the form for any d (used above d = 256, where the query
// operands no longer fit in registers): the query operands stream through
// the ring beside the candidate operands (the wave's query rows stay in the
// caches), the pair loop runs JR = ceil(ceil(d / 8) / R) R positions, and
// query dims past d are zeroed where they enter the MFMA.  The
// candidates' norms shifted by the MFMA error margin (a.SN, by position)
// are read a tile ahead.
//
// Epilogue: t = fl(-2 dot + shifted XA2) against a per-row threshold with
// the MFMA error margin folded in (mfma_shift_norm, mfma_threshold), so
// every pair whose oracle distance can reach the row's root passes; wave
// masks in scalar registers; passing pairs wait in a per-lane queue of
// candidate ids.  When a queue would overflow, every lane forms the oracle
// distances of its queued pairs at once (oracle_dist_batch: all queued pairs
// interleaved, rows read from global memory) and the four lane groups
// (each covering the wave's queries once) update their LDS heaps in turn.
// Heaps hold only oracle distances: results are bit-identical to the
// oracle's.
#include <limits.h>

#include <atomic>

#include "gsknn_device.h"

namespace gsknn {

namespace {

typedef double rdbl4 __attribute__((ext_vector_type(4)));
typedef double rdbl2 __attribute__((ext_vector_type(2)));

constexpr int kRing = 4;  // ring slots (pairs of k-steps)
constexpr int kRingQ = 4;  // per-lane queue depth

template <int KP, int QB, int CB, int WPE, bool VEC, bool TWO>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void knn_mfma_f64_ring_kernel(
    KnnArgs<double> a) {
  static_assert(QB <= 2, "query block fits one bit");
  constexpr bool QS = KP == 0;  // query operands streamed through the ring
  constexpr int Q = 16 * QB, C = 16 * CB, R = KP % 4 == 0 ? 4 : 2, QR = kRingQ;
  static_assert(KP % R == 0, "ring slots per tile");
  typedef int idv __attribute__((ext_vector_type(CB * 4)));
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k, d = a.d;
  const int n = a.ncount ? *a.ncount : a.n;
  const int qbase = blockIdx.x * Q;
  if (qbase >= n) return;  // fallback launches over a device-side list
  const int kh = TWO ? a.k2 : k;  // heap entries per query
  const MfmaHeapLds hl(Q, kh, TWO);
  double *hd = (double *)(smem + hl.hd);
  int *hi = (int *)(smem + hl.hi);
  double *r0 = (double *)(smem + hl.r0);  // two-stage: roots of the resident rows

  const int lane = threadIdx.x, grp = lane >> 4, cl = lane & 15;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const bool sentinel = a.nsplit > 1;

  if (lane < Q && qbase + lane < n) {
    if (TWO)
      key_heap_init(a, hd, hi, r0, Q, lane, qbase + lane, kh);
    else
      heap_load(a, hd, hi, Q, lane, qbase + lane, sentinel);
  }

  const int JR = QS ? ((d + 7) / 8 + R - 1) / R * R : KP;  // pair positions per tile
  const double cm = mfma_margin_c(8 * JR, d);

  double bq[QB][QS ? 1 : 2 * KP];  // register-resident query operands (KP > 0)
  double qb2[QB];                  // squared norms of the query rows (KP = 0)
  const double *qp[QB];            // query rows
  double qn[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = qb * 16 + cl;
    const int j = qbase + ql[qb];
    qv[qb] = j < n;
    const int row = a.bmap[qv[qb] ? j : n - 1];
    const double *xq = a.XB + (size_t)row * d;
    qp[qb] = xq;
    qn[qb] = qv[qb] ? a.XB2[row] : 0.0;
    if constexpr (QS) {
      // Each lane sums the whole row (once per launch).
      double b2 = 0.0;
      for (int p = 0; p < d; ++p) b2 = __builtin_fma(xq[p], xq[p], b2);
      qb2[qb] = qv[qb] ? b2 : 0.0;
    } else {
#pragma unroll
      for (int s = 0; s < 2 * KP; ++s) {
        const int p = 8 * (s >> 1) + 2 * grp + (s & 1);
        const double v = xq[p < d ? p : d - 1];
        bq[qb][s] = (qv[qb] && p < d) ? v : 0.0;
      }
    }
  }
  // XB row of query block qb.
  auto qrow = [&](int qb) -> int { return a.bmap[qv[qb] ? qbase + ql[qb] : n - 1]; };

  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  struct Ids {
    int v[CB];
  };
  auto cand_id = [&](int t) -> Ids {
    Ids r;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) r.v[cb] = a.amap[min(cbeg + t * C + cb * 16 + cl, cend - 1)];
    return r;
  };
  auto slot_ok = [&](int t, int cb) -> bool { return cbeg + t * C + cb * 16 + cl < cend; };
  // Pair j of a candidate row: dims 8j + 2g, 8j + 2g + 1, moved back inside
  // the row past d (their query operands are zero).
  const int pmax = VEC ? d - 2 : d - 1;
  auto load_pair = [&](const double *row, int j) -> rdbl2 {
    const int p = 8 * j + 2 * grp;
    rdbl2 v;
    if (VEC) {
      v = *(const rdbl2 *)(row + min(p, pmax));
    } else {
      v.x = row[min(p, pmax)];
      v.y = row[min(p + 1, pmax)];
    }
    return v;
  };
  auto row_of = [&](int id) -> const double * { return a.XA + (size_t)id * d; };

  rdbl2 xr[R][CB];  // candidate operand ring
  rdbl2 yr[R][QS ? QB : 1];  // query operand ring (KP = 0)
  double cn[CB];    // shifted norms of this lane's candidates in the current tile (NaN: empty slot)
  auto pos_sn = [&](int t, int cb) -> double { return a.SN[min(cbeg + t * C + cb * 16 + cl, cend - 1)]; };
  int ci[CB];       // their ids (INT_MAX for empty slots)
  const double nan = __builtin_nan("");
  Ids id_a{}, id_b{};
  Ids id_t{};       // ids of the current tile (valid rows; empty slots repeat the last candidate)
  int safe_id = 0;  // a valid XA row, for queue slots without an entry
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    cn[cb] = nan;
    ci[cb] = INT_MAX;
#pragma unroll
    for (int r = 0; r < R; ++r) xr[r][cb] = rdbl2{0.0, 0.0};
  }
  if (ntiles > 0) {
    id_t = cand_id(0);
    safe_id = id_t.v[0];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
#pragma unroll
      for (int r = 0; r < R; ++r) xr[r][cb] = load_pair(row_of(id_t.v[cb]), r);
      const bool ok = slot_ok(0, cb);
      cn[cb] = ok ? pos_sn(0, cb) : nan;
      ci[cb] = ok ? id_t.v[cb] : INT_MAX;
    }
    if constexpr (QS) {
#pragma unroll
      for (int qb = 0; qb < QB; ++qb)
#pragma unroll
        for (int r = 0; r < R; ++r) yr[r][qb] = load_pair(qp[qb], r);
    }
    id_a = ntiles > 1 ? cand_id(1) : id_t;
    id_b = ntiles > 2 ? cand_id(2) : id_a;
  }
  __syncthreads();  // heaps loaded

  // Filter threshold of a query row (mfma_threshold); the query's squared
  // norm comes from its operand registers (the four lane groups hold
  // disjoint dims).  Every lane of the wave calls it.
  auto root_thr = [&](int qb) -> double {
    double b2 = 0.0;
    if constexpr (QS) {
      b2 = qb2[qb];
    } else {
#pragma unroll
      for (int s = 0; s < 2 * KP; ++s) b2 = __builtin_fma(bq[qb][s], bq[qb][s], b2);
      b2 = xadd(b2, __shfl_xor(b2, 16));
      b2 = xadd(b2, __shfl_xor(b2, 32));
    }
    if (!qv[qb]) return -__builtin_inf();
    if (TWO) return key_threshold(r0[ql[qb]], hd[ql[qb]], qn[qb], xmul(xmul(0.5, cm), b2));
    return mfma_threshold(hd[ql[qb]], qn[qb], xmul(xmul(0.5, cm), b2));
  };
  double thr[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) thr[qb] = root_thr(qb);

  // Per-lane queue of passing pairs: candidate id, query block in the top
  // bit (ids are non-negative).
  unsigned pq_i[QR];
  double pq_t[QR];  // two-stage: the pairs' MFMA keys
  int pq_n = 0;
#pragma unroll
  for (int r = 0; r < QR; ++r) {
    pq_i[r] = 0u;
    pq_t[r] = 0.0;
  }
  rdbl4 acc[QB][CB];
  auto flush = [&]() __attribute__((always_inline)) {
    double ed[QR];
    if constexpr (TWO) {
#pragma unroll
      for (int r = 0; r < QR; ++r) ed[r] = pq_t[r];
    } else {
      const double *qp0 = a.XB + (size_t)qrow(0) * d;
      const double *qp1 = a.XB + (size_t)qrow(QB - 1) * d;
      const double *cp[QR], *qp[QR];
      double cnr[QR], qnr[QR];
#pragma unroll
      for (int r = 0; r < QR; ++r) {
        const bool on = r < pq_n;
        const bool hq = pq_i[r] >> 31;
        const int id = on ? (int)(pq_i[r] & 0x7fffffffu) : safe_id;
        cp[r] = a.XA + (size_t)id * d;
        qp[r] = hq ? qp1 : qp0;
        cnr[r] = a.XA2[id];
        qnr[r] = hq ? qn[QB - 1] : qn[0];
      }
      oracle_dist_batch<QR, VEC, (KP <= 24 ? 2 : 1)>(cp, qp, d, cnr, qnr, ed);
    }
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
      if (grp == g) {
#pragma unroll
        for (int r = 0; r < QR; ++r) {
          if (r < pq_n) {
            const int slot = (pq_i[r] >> 31) ? ql[QB - 1] : ql[0];
            const int id = (int)(pq_i[r] & 0x7fffffffu);
            if (cless(ed[r], id, hd[slot], hi[slot])) sift_down(hd, hi, Q, slot, 0, kh, ed[r], id);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    pq_n = 0;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) thr[qb] = root_thr(qb);
  };

  for (int t = 0; t < ntiles; ++t) {
    const bool more = t + 1 < ntiles;
    const Ids id_c = (t + 3 < ntiles) ? cand_id(t + 3) : id_b;
    double nn[CB];  // shifted norms of tile t + 1 (its ids are id_a)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) nn[cb] = pos_sn(more ? t + 1 : t, cb);
    const double *rc[CB], *rn[CB];  // rows of this tile and of the next
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      rc[cb] = row_of(id_t.v[cb]);
      rn[cb] = row_of(id_a.v[cb]);
    }

#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = rdbl4{0.0, 0.0, 0.0, 0.0};

    __builtin_amdgcn_s_setprio(1);
    if constexpr (!QS) {
#pragma unroll
      for (int j = 0; j < KP; ++j) {
        const int slot = j % R;
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) {
            const double av = e ? xr[slot][cb].y : xr[slot][cb].x;
#pragma unroll
            for (int qb = 0; qb < QB; ++qb)
              acc[qb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bq[qb][2 * j + e], acc[qb][cb], 0, 0, 0);
          }
        // Refill the slot with pair j + R of this tile, or across the tile
        // boundary with the next tile's pair j + R - KP (on the last tile the
        // refill reads valid rows and is never used).
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          xr[slot][cb] = j + R < KP ? load_pair(rc[cb], j + R) : load_pair(rn[cb], j + R - KP);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      for (int j0 = 0; j0 < JR; j0 += R) {
#pragma unroll
        for (int slot = 0; slot < R; ++slot) {
          const int j = j0 + slot;
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            // Query dims past d enter as zero.
            const bool in = 8 * j + 2 * grp + e < d;
            double bv[QB];
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) bv[qb] = in ? (e ? yr[slot][qb].y : yr[slot][qb].x) : 0.0;
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
              const double av = e ? xr[slot][cb].y : xr[slot][cb].x;
#pragma unroll
              for (int qb = 0; qb < QB; ++qb)
                acc[qb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[qb], acc[qb][cb], 0, 0, 0);
            }
          }
          // Refill with pair j + R of this tile or the next tile's pair
          // j + R - JR (the query rows stay the same).
          const bool nxt = j + R >= JR;
          const int jn = nxt ? j + R - JR : j + R;
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) xr[slot][cb] = load_pair(nxt ? rn[cb] : rc[cb], jn);
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) yr[slot][qb] = load_pair(qp[qb], jn);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);

    // Epilogue: one FMA and one compare per pair; per-lane bits and ids
    // only when a pair passes.
    double ctn[CB * 4];
#pragma unroll
    for (int b = 0; b < CB * 4; ++b) ctn[b] = __shfl(cn[b >> 2], grp + 4 * (b & 3));
    unsigned long long msk[QB][CB * 4];
    unsigned long long anym = 0;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double tt = __builtin_fma(acc[qb][cb][v], -2.0, ctn[cb * 4 + v]);
          acc[qb][cb][v] = tt;
          msk[qb][cb * 4 + v] = __ballot(tt <= thr[qb]);
          anym |= msk[qb][cb * 4 + v];
        }
    if (anym) {
      unsigned pass[QB];
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        unsigned long long aq = 0ull;
#pragma unroll
        for (int b = 0; b < CB * 4; ++b) aq |= msk[qb][b];
        pass[qb] = 0u;
        if (aq != 0ull) {
#pragma unroll
          for (int b = 0; b < CB * 4; ++b) pass[qb] |= (unsigned)((msk[qb][b] >> lane) & 1ull) << b;
        }
      }
      idv cid;
#pragma unroll
      for (int b = 0; b < CB * 4; ++b) cid[b] = __shfl(ci[b >> 2], grp + 4 * (b & 3));
      // Queue the passing pairs; whenever some lane's queue is full, apply
      // the queues and drop the pairs left that fail the new thresholds.
      while (true) {
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) {
          while (pass[qb] && pq_n < QR) {
            const int b = __builtin_ctz(pass[qb]);
            pass[qb] &= pass[qb] - 1;
            int idx = cid[0];
            double tv = acc[qb][0][0];
#pragma unroll
            for (int bb = 1; bb < CB * 4; ++bb)
              if (b == bb) {
                idx = cid[bb];
                if (TWO) tv = acc[qb][bb >> 2][bb & 3];
              }
#pragma unroll
            for (int r = 0; r < QR; ++r)
              if (pq_n == r) {
                pq_i[r] = (unsigned)idx | ((unsigned)qb << 31);
                if (TWO) pq_t[r] = tv;
              }
            ++pq_n;
          }
        }
        unsigned left = 0u;
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) left |= pass[qb];
        if (!__any(left != 0u)) break;
        flush();
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
#pragma unroll
          for (int b = 0; b < CB * 4; ++b)
            if (!(acc[qb][b >> 2][b & 3] <= thr[qb])) pass[qb] &= ~(1u << b);
      }
    }

#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const bool ok = more && slot_ok(t + 1, cb);
      cn[cb] = ok ? nn[cb] : nan;
      ci[cb] = ok ? id_a.v[cb] : INT_MAX;
    }
    id_t = id_a;
    id_a = id_b;
    id_b = id_c;
  }

  if (__any(pq_n > 0)) flush();
  __syncthreads();
  if (lane < Q && qbase + lane < n) {
    if (TWO)
      key_heap_store(a, hd, hi, Q, lane, qbase + lane, kh);
    else
      heap_store(a, hd, hi, Q, lane, qbase + lane, split);
  }
}

struct RingCfg {
  int kp, qb, cb, wpe;
};

// Shape per dimension class (d <= 8 KP, KP pairs computed whatever d is):
// one wave per SIMD (512 registers: the query operands of all pairs, the
// ring, the accumulators and an exact-distance flush without spilling), two
// query blocks per wave (each candidate operand feeds two MFMAs) and four
// candidate blocks per tile up to d = 112 (the per-tile epilogue amortised
// over 16 MFMAs per pair), two up to d = 192 and one up to d = 256, where
// the query operands take more registers; for d in 257 .. 320 one query
// block per wave (its 40 pairs of operands in the registers two blocks of 32
// take): d = 300 runs 55.2 ms against 81.7 for the streamed form, odd d = 283
// (8-byte loads) 101 against 109 ms.  Measured and left out: 48 pairs with
// one candidate block (d = 352 / 384: 88.9 / 89.2 ms, no faster than the
// streamed form); 56 and more pairs spill.
constexpr int kRingCfg[][4] = {{2, 2, 4, 1},   {4, 2, 4, 1},   {6, 2, 4, 1},   {8, 2, 4, 1},
                               {10, 2, 4, 1},  {12, 2, 4, 1},  {14, 2, 4, 1},  {16, 2, 2, 1},
                               {20, 2, 2, 1},  {24, 2, 2, 1},  {28, 2, 1, 1},  {32, 2, 1, 1},
                               {40, 1, 2, 1}};
constexpr int kNumRingCfg = 13;

// Above d = 256 otherwise: queries streamed (KP = 0), 2 x 2 blocks.
constexpr int kRingStreamCfg[4] = {0, 2, 2, 1};

bool ring_cfg(int d, RingCfg *c) {
  if (d <= 0) return false;
  for (int i = 0; i < kNumRingCfg; ++i)
    if (d <= 8 * kRingCfg[i][0]) {
      *c = {kRingCfg[i][0], kRingCfg[i][1], kRingCfg[i][2], kRingCfg[i][3]};
      return true;
    }
  *c = {kRingStreamCfg[0], kRingStreamCfg[1], kRingStreamCfg[2], kRingStreamCfg[3]};
  return true;
}

template <int KP, int QB, int CB, int WPE, bool VEC>
int ring_reg_blocks_t() {
  // Both forms must run without spilling; the one-stage form re-runs the
  // two-stage form's unsettled queries.
  const void *kerns[2] = {(const void *)knn_mfma_f64_ring_kernel<KP, QB, CB, WPE, VEC, false>,
                          (const void *)knn_mfma_f64_ring_kernel<KP, QB, CB, WPE, VEC, true>};
  int best = 1 << 30;
  for (const void *kern : kerns) {
    hipFuncAttributes at;
    if (hipFuncGetAttributes(&at, kern) == hipSuccess && at.localSizeBytes > 0) return 0;  // spills
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, 64, 0) != hipSuccess || b <= 0)
      b = 4 * WPE;
    best = std::min(best, b);
  }
  return best;
}

#define GSKNN_RING_SHAPES(X) \
  X(2, 2, 4, 1)              \
  X(4, 2, 4, 1)              \
  X(6, 2, 4, 1)              \
  X(8, 2, 4, 1)              \
  X(10, 2, 4, 1)             \
  X(12, 2, 4, 1)             \
  X(14, 2, 4, 1)             \
  X(16, 2, 2, 1)             \
  X(20, 2, 2, 1)             \
  X(24, 2, 2, 1)             \
  X(28, 2, 1, 1)             \
  X(32, 2, 1, 1)             \
  X(40, 1, 2, 1)             \
  X(0, 2, 2, 1)

}  // namespace

int ring_terms(int d) {
  RingCfg c;
  if (!ring_cfg(d, &c)) return 0;
  return c.kp > 0 ? 8 * c.kp : 8 * (((d + 7) / 8 + 3) / 4 * 4);
}

bool ring_supported(int d) {
  RingCfg c;
  return ring_cfg(d, &c);
}

int ring_queries_per_block(int d) {
  RingCfg c;
  return ring_cfg(d, &c) ? 16 * c.qb : 0;
}

int ring_candidates_per_tile(int d) {
  RingCfg c;
  return ring_cfg(d, &c) ? 16 * c.cb : 0;
}

size_t ring_lds_bytes(int d, int k, bool two) {
  const int Q = ring_queries_per_block(d);
  return MfmaHeapLds(Q, two ? two_stage_k2(k) : k, two).bytes;
}

// Workgroups per CU that registers allow (0: the instantiation spills),
// queried once per instantiation.
int ring_reg_blocks(int d, bool vec) {
  RingCfg c;
  if (!ring_cfg(d, &c)) return 0;
  static std::atomic<int> cache[2 * kNumRingCfg + 2];
  int ci = 0;
  while (ci < kNumRingCfg && kRingCfg[ci][0] != c.kp) ++ci;
  const int ki = ci * 2 + (vec ? 1 : 0);
  int b = cache[ki].load(std::memory_order_relaxed);
  if (b > 0) return b - 1;
#define GSKNN_RING_OCC(KP_, QB_, CB_, WPE_)                                   \
  if (c.kp == KP_) b = vec ? ring_reg_blocks_t<KP_, QB_, CB_, WPE_, true>()   \
                           : ring_reg_blocks_t<KP_, QB_, CB_, WPE_, false>();
  GSKNN_RING_SHAPES(GSKNN_RING_OCC)
#undef GSKNN_RING_OCC
  cache[ki].store(b + 1, std::memory_order_relaxed);
  return b;
}

hipError_t launch_knn_ring(const KnnArgs<double> &a, bool vec, int gx, size_t lds, hipStream_t s,
                           bool two) {
  RingCfg c;
  if (!ring_cfg(a.d, &c)) return hipErrorNotSupported;
  const dim3 grid(gx, a.nsplit);
#define GSKNN_RING_LAUNCH(KP_, QB_, CB_, WPE_)                                          \
  if (c.kp == KP_) {                                                                    \
    auto kern = vec ? (two ? knn_mfma_f64_ring_kernel<KP_, QB_, CB_, WPE_, true, true>   \
                           : knn_mfma_f64_ring_kernel<KP_, QB_, CB_, WPE_, true, false>) \
                    : (two ? knn_mfma_f64_ring_kernel<KP_, QB_, CB_, WPE_, false, true>  \
                           : knn_mfma_f64_ring_kernel<KP_, QB_, CB_, WPE_, false, false>); \
    hipError_t e = hipFuncSetAttribute((const void *)kern,                              \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                      \
    hipLaunchKernelGGL(kern, grid, dim3(64), lds, s, a);                                \
    return hipGetLastError();                                                           \
  }
  GSKNN_RING_SHAPES(GSKNN_RING_LAUNCH)
#undef GSKNN_RING_LAUNCH
  return hipErrorNotSupported;
}

}  // namespace gsknn
