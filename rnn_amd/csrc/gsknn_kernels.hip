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
//   knn_mfma_f64_kernel     FP64 MFMA (v_mfma_f64_16x16x4_f64) filter with a
//                           rigorous error margin, oracle-arithmetic
//                           distances for the pairs that pass (bit-identical
//                           results), wave-owned queries, LDS heaps.
//   knn_merge_kernel<T>     folds the partial lists of a candidate-split
//                           launch into the resident heap rows.
//   sqnorm_kernel<T>        X2[i] = sum_p X[i][p]^2 in p order.
//
// Heaps live in LDS as [entry][query] so that lanes touching the same entry
// of different queries hit different banks.  Rows are written back sorted
// in descending order, which is a valid max-heap.
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "gsknn_device.h"

namespace gsknn {

// ---------------------------------------------------------------------------
// FP64 MFMA kernel.
//
// A workgroup of W waves owns Q = W * 16 * QB queries; wave w owns 16 * QB of
// them.  Candidates stream through LDS in tiles of C = 16 * CB, stored in
// MFMA fragment order so that each A-operand read is 64 consecutive doubles,
// with the tile's candidate norms and ids beside it.
// v_mfma_f64_16x16x4_f64 takes A = 16 candidates x 4 dims (lane l: candidate
// l%16, dim l/16) and B = 4 dims x 16 queries (lane l: dim l/16, query l%16);
// lane l of the 16x16 result holds query l%16 against candidates l/16 + 4*v,
// v = 0..3 (measured on gfx950).  The query fragments stay in registers for
// the whole launch (d <= 4*KS).
//
// NB = 2 tile buffers: the next tile is written while the current one is
// read, one barrier per tile; lane l of every wave holds the norm and id of
// tile candidate l % C in registers (loaded a tile ahead) and the epilogue
// fetches them with lane shuffles.  NB = 1: two barriers per tile, norms and
// ids in LDS.
//
// Epilogue per lane: t = fl(-2 dot + shifted XA2[amap[i]]) against a per-row
// threshold with the MFMA error margin folded in (mfma_shift_norm,
// mfma_threshold), so every pair whose oracle distance can reach the row's
// root passes.  A wave-wide ballot skips the insertion step when no lane has
// a survivor, which is the common case once the rows are warm.  Passing
// pairs wait in a per-lane queue; when one would overflow, every lane forms
// the oracle-arithmetic distance of its queued pairs (oracle_dist) and the
// four lane groups (l/16 = 0..3, each covering the 16 queries once) insert
// in turn, so no two lanes ever touch the same heap at once.  Heaps hold
// only oracle distances: results are bit-identical to the oracle's.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

// NB = 1 keeps the tile's candidate norms and ids in LDS beside the tile;
// NB = 2 keeps them in registers (every LDS byte goes to the two tiles and
// the heaps, which keeps two 128-query workgroups per CU at k = 32, d = 64).
struct MfmaLayout {
  size_t tA, tn, ti, hd, hi, bytes;
  __host__ __device__ MfmaLayout(int KS, int CB, int Q, int k, int NB) {
    const size_t tile = (size_t)KS * CB * 64;
    const size_t C = NB == 1 ? 16 * (size_t)CB : 0;
    tA = 0;
    tn = al16(tA + sizeof(double) * tile * NB);
    ti = al16(tn + sizeof(double) * C);
    hd = al16(ti + sizeof(int) * C);
    hi = al16(hd + sizeof(double) * (size_t)k * Q);
    bytes = al16(hi + sizeof(int) * (size_t)k * Q);
  }
};

// Candidate tiles go global -> registers -> LDS.  The amap entries are read
// two tiles ahead and the point rows one tile ahead, so no load waits on
// another inside the loop and the loads overlap the MFMA work of the
// current tile.
// Two waves per SIMD whenever the query fragments (2 * QB * KS registers)
// and the staged tile elements (KS * CB / W doubles) leave room: the
// compiler then keeps the accumulators in VGPRs and fits the survivor queue
// below 256 registers.
constexpr int mfma_waves_per_eu(int KS, int QB, int CB, int W) {
  return (QB * KS <= 32 && KS * CB <= 8 * W) ? 2 : 1;
}

template <int KS, int QB, int CB, int W, int NB>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(mfma_waves_per_eu(KS, QB, CB, W)))) void knn_mfma_f64_kernel(
    KnnArgs<double> a) {
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB;
  static_assert(NB == 1 || NB == 2, "tile buffers");
  static_assert(C <= 64, "one candidate per lane");
  constexpr int TILE = KS * CB * 64;
  // Candidate ids of a lane's CB * 4 result values; a vector type keeps the
  // per-bit selection in registers.
  typedef int idv __attribute__((ext_vector_type(CB * 4)));
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k, d = a.d;
  const MfmaLayout L(KS, CB, Q, k, NB);
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

  const double cm = mfma_margin_c(4 * KS, d);  // margin constant (query side)
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
  // XB row of query block qb.
  auto qrow = [&](int qb) -> int { return a.bmap[qv[qb] ? qbase + ql[qb] : a.n - 1]; };

  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  // Staging: the tile is BLK fragment blocks of 64 doubles (k-step s,
  // candidate block cb); wave w stages blocks w * BPW .. w * BPW + BPW - 1 and
  // lane l writes element l of each, i.e. dim 4 s + l / 16 of candidate
  // cb * 16 + l % 16.  Every LDS store is then 512 contiguous bytes (no bank
  // conflicts) and each lane gathers 8-byte elements, 32 contiguous bytes
  // per candidate row and instruction.  A thread touches the candidates
  // cb * 16 + l % 16 for cb < CB; their ids travel in an Ids.
  constexpr int BLK = KS * CB;
  static_assert(BLK % W == 0, "fragment blocks per wave");
  constexpr int BPW = BLK / W;
  const int si = lane & 15, skk = lane >> 4;
  struct Ids {
    int v[CB];
  };
  double stage[BPW];
  constexpr int NPW = (CB + W - 1) / W;  // norm/id blocks per wave
  double stage_n[NPW];
#pragma unroll
  for (int r = 0; r < NPW; ++r) stage_n[r] = 0.0;
  auto cand_id = [&](int t) -> Ids {
    Ids r;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) r.v[cb] = a.amap[min(cbeg + t * C + cb * 16 + si, cend - 1)];
    return r;
  };
  // Lanes 0..15 of wave w own the norm and id slots of candidate blocks
  // w, w + W, ... (one block per wave when W >= CB).
  const bool nslot = NB == 1 && skk == 0;
  auto own_id = [&](const Ids &ids, int nb) -> int {
    int r = ids.v[0];
#pragma unroll
    for (int cb = 1; cb < CB; ++cb)
      if (nb == cb) r = ids.v[cb];
    return r;
  };
  auto load_x = [&](const Ids &ids, int t) {
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      const int sb = bidx / CB, cb = bidx % CB;
      const int p = 4 * sb + skk;
      stage[e] = a.XA[(size_t)ids.v[cb] * d + (p < d ? p : d - 1)];
    }
    if (nslot) {
#pragma unroll
      for (int r = 0; r < NPW; ++r)
        if (w + r * W < CB) stage_n[r] = a.SN[min(cbeg + t * C + (w + r * W) * 16 + si, cend - 1)];
    }
  };
  auto store_x = [&](int buf, int t, const Ids &ids) {
    double *dst = tA + buf * TILE;
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      const int sb = bidx / CB, cb = bidx % CB;
      const int p = 4 * sb + skk;
      const bool ok = cbeg + t * C + cb * 16 + si < cend;
      dst[bidx * 64 + lane] = (ok && p < d) ? stage[e] : 0.0;
    }
    if (nslot) {
      // An empty slot gets a NaN norm (no comparison with it passes) and
      // index INT_MAX.
#pragma unroll
      for (int r = 0; r < NPW; ++r) {
        const int nb = w + r * W;
        if (nb < CB) {
          const int slot = nb * 16 + si;
          const bool ok = cbeg + t * C + slot < cend;
          tn[slot] = ok ? stage_n[r] : __builtin_nan("");
          ti[slot] = ok ? own_id(ids, nb) : INT_MAX;
        }
      }
    }
  };
  // NB = 2: this lane's tile candidate (lane % C), id two tiles ahead and
  // norm one tile ahead; empty slots as above.
  const int cl = lane & (C - 1);
  auto lane_id = [&](int t) -> int {
    const int c = min(cbeg + t * C + cl, cend - 1);
    return a.amap[c];
  };
  double cur_tn = __builtin_nan("");
  int cur_ti = INT_MAX;
  int lid_a = 0, lid_b = 0;  // lane ids of tiles t + 1 and t + 2
  if (NB == 2 && ntiles > 0) {
    const int id0 = lane_id(0);
    const bool ok = cbeg + cl < cend;
    cur_tn = ok ? a.SN[cbeg + cl] : __builtin_nan("");
    cur_ti = ok ? id0 : INT_MAX;
    lid_a = ntiles > 1 ? lane_id(1) : id0;
    lid_b = ntiles > 2 ? lane_id(2) : lid_a;
  }

  Ids id_a{}, id_b{};  // staging ids of the tiles after the current one
  if (ntiles > 0) {
    const Ids id0 = cand_id(0);
    load_x(id0, 0);
    id_a = ntiles > 1 ? cand_id(1) : id0;
    __syncthreads();  // heaps loaded
    store_x(0, 0, id0);
    if (NB == 2) {
      if (ntiles > 1) load_x(id_a, 1);
      id_b = ntiles > 2 ? cand_id(2) : id_a;
    }
  }
  __syncthreads();
  // Filter threshold of a query row (mfma_threshold): every pair whose
  // oracle distance can reach the root r passes; lanes without a query get
  // -inf.  The query's squared norm comes from its operand registers (the
  // four lane groups hold disjoint dims of the row); every lane of the wave
  // calls it.
  auto root_thr = [&](int qb) -> double {
    double b2 = 0.0;
#pragma unroll
    for (int s = 0; s < KS; ++s) b2 = __builtin_fma(bq[qb][s], bq[qb][s], b2);
    b2 = xadd(b2, __shfl_xor(b2, 16));
    b2 = xadd(b2, __shfl_xor(b2, 32));
    if (!qv[qb]) return -__builtin_inf();
    return mfma_threshold(hd[ql[qb]], qn[qb], xmul(xmul(0.5, cm), b2));
  };
  double thr[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) thr[qb] = root_thr(qb);

  // Per-lane queue of passing pairs: candidate id, with the query block in
  // the top bit (ids are non-negative).
  static_assert(QB <= 2, "query block fits one bit");
  constexpr int QR = 4;
  unsigned pq_i[QR];
  int pq_n = 0;
#pragma unroll
  for (int r = 0; r < QR; ++r) pq_i[r] = 0u;
  // Apply the queues: every lane forms the oracle distances of its queued
  // pairs at once (rows read from global memory), then the four lane groups,
  // which cover the wave's queries once each (within a group every lane owns
  // a distinct heap), update their heaps in turn.
  auto flush = [&]() __attribute__((always_inline)) {
    double ed[QR];
    const int qr0 = qrow(0), qr1 = qrow(QB - 1);
#pragma unroll
    for (int r = 0; r < QR; ++r) {
      ed[r] = 0.0;
      if (r < pq_n) {
        const bool hi_qb = pq_i[r] >> 31;
        const int id = (int)(pq_i[r] & 0x7fffffffu);
        const double *xq = a.XB + (size_t)(hi_qb ? qr1 : qr0) * d;
        ed[r] = oracle_dist(a.XA + (size_t)id * d, xq, d, a.XA2[id], hi_qb ? qn[QB - 1] : qn[0]);
      }
    }
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
      if (grp == g) {
#pragma unroll
        for (int r = 0; r < QR; ++r) {
          if (r < pq_n) {
            const int slot = (pq_i[r] >> 31) ? ql[QB - 1] : ql[0];
            const int id = (int)(pq_i[r] & 0x7fffffffu);
            if (cless(ed[r], id, hd[slot], hi[slot])) sift_down(hd, hi, Q, slot, 0, k, ed[r], id);
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

  dbl4 acc[QB][CB];
  // MFMAs of candidate blocks [LO, HI) of the tile at `cur`.
  auto mfma_blocks = [&](const double *cur, auto lo_c, auto hi_c) __attribute__((always_inline)) {
    constexpr int LO = decltype(lo_c)::value, HI = decltype(hi_c)::value;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = LO; cb < HI; ++cb) acc[qb][cb] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int cb = LO; cb < HI; ++cb) {
        const double av = cur[(s * CB + cb) * 64 + lane];
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bq[qb][s], acc[qb][cb], 0, 0, 0);
      }
    }
  };
  // Epilogue of candidate blocks [LO, HI): one FMA and one compare per pair
  // against the shifted norm (a.SN) and threshold (mfma_threshold), the
  // compare's wave mask kept in scalar registers; per-lane bits and ids are
  // built only when some pair passes.  Norms and ids come from LDS (NB = 1)
  // or from the lanes' registers tnr/tir (NB = 2).
  auto epilogue = [&](auto lo_c, auto hi_c, double tnr, int tir) __attribute__((always_inline)) {
    constexpr int LO = decltype(lo_c)::value, HI = decltype(hi_c)::value;
    double ctn[CB * 4];
#pragma unroll
    for (int b = LO * 4; b < HI * 4; ++b) {
      const int c = (b >> 2) * 16 + grp + 4 * (b & 3);
      ctn[b] = NB == 2 ? __shfl(tnr, c) : tn[c];
    }
    unsigned long long msk[QB][CB * 4];
    unsigned long long anym = 0;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = LO; cb < HI; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // -2 * dot is exact: t = fl(-2 dot + XA2[amap[i]]).
          const double tt = __builtin_fma(acc[qb][cb][v], -2.0, ctn[cb * 4 + v]);
          acc[qb][cb][v] = tt;
          msk[qb][cb * 4 + v] = __ballot(tt <= thr[qb]);
          anym |= msk[qb][cb * 4 + v];
        }
    // Survivors wait in a per-lane queue of QR entries; the heaps (and the
    // thresholds used by the filter) are updated only when some lane's
    // queue would overflow, which keeps the slow path off most tiles.  A
    // stale threshold only lets more pairs through the filter; every queued
    // pair is tested exactly against the live root when it is applied.
    if (anym) {
      unsigned pass[QB];
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        unsigned long long aq = 0ull;  // scalar: any pass in this block
#pragma unroll
        for (int b = LO * 4; b < HI * 4; ++b) aq |= msk[qb][b];
        pass[qb] = 0u;
        if (aq != 0ull) {
#pragma unroll
          for (int b = LO * 4; b < HI * 4; ++b)
            pass[qb] |= (unsigned)((msk[qb][b] >> lane) & 1ull) << b;
        }
      }
      idv cid = {};
      if (NB == 2) {
#pragma unroll
        for (int b = LO * 4; b < HI * 4; ++b)
          cid[b] = __shfl(tir, (b >> 2) * 16 + grp + 4 * (b & 3));
      }
      // Queue the passing pairs; whenever some lane's queue is full, apply
      // the queues and drop the pairs left that fail the new thresholds.
      while (true) {
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) {
          while (pass[qb] && pq_n < QR) {
            const int b = __builtin_ctz(pass[qb]);
            pass[qb] &= pass[qb] - 1;
            int idx;
            if (NB == 2) {
              idx = cid[0];
#pragma unroll
              for (int bb = 1; bb < CB * 4; ++bb)
                if (b == bb) idx = cid[bb];
            } else {
              idx = ti[(b >> 2) * 16 + grp + 4 * (b & 3)];
            }
#pragma unroll
            for (int r = 0; r < QR; ++r)
              if (pq_n == r) pq_i[r] = (unsigned)idx | ((unsigned)qb << 31);
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
          for (int b = LO * 4; b < HI * 4; ++b)
            if (!(acc[qb][b >> 2][b & 3] <= thr[qb])) pass[qb] &= ~(1u << b);
      }
    }
  };
  typedef std::integral_constant<int, 0> I0;
  typedef std::integral_constant<int, CB> ICB;
  for (int t = 0; t < ntiles; ++t) {
    const bool more = t + 1 < ntiles;
    const double *cur = tA + (NB == 2 ? (t & 1) * TILE : 0);
    // The amap entry two tiles ahead is requested before the rows of the
    // next tile, so the row loads never wait behind it.
    Ids id_c{};
    int lid_c = 0;
    double nxt_tn = 0.0;
    if (NB == 2) {
      // The staged tile t + 1 goes to the other buffer, whose last readers
      // (tile t - 1) passed the barrier at the end of the previous step.
      if (more) store_x((t + 1) & 1, t + 1, id_a);
      id_c = (t + 3 < ntiles) ? cand_id(t + 3) : id_b;
      if (t + 2 < ntiles) load_x(id_b, t + 2);
      lid_c = (t + 3 < ntiles) ? lane_id(t + 3) : lid_b;
      nxt_tn = a.SN[min(cbeg + (t + 1) * C + cl, cend - 1)];
    } else {
      id_c = (t + 2 < ntiles) ? cand_id(t + 2) : id_a;
      if (more) load_x(id_a, t + 1);
    }

    mfma_blocks(cur, I0{}, ICB{});
    epilogue(I0{}, ICB{}, cur_tn, cur_ti);

    if (NB == 2) {
      __syncthreads();  // tile t consumed, tile t + 1 visible
      id_a = id_b;
      id_b = id_c;
      const bool nxt_ok = more && cbeg + (t + 1) * C + cl < cend;
      cur_tn = nxt_ok ? nxt_tn : __builtin_nan("");
      cur_ti = nxt_ok ? lid_a : INT_MAX;
      lid_a = lid_b;
      lid_b = lid_c;
    } else {
      __syncthreads();  // tile consumed
      if (more) store_x(0, t + 1, id_a);
      __syncthreads();  // next tile visible
      id_a = id_c;
    }
  }

  if (__any(pq_n > 0)) flush();
  __syncthreads();
  if (tid < Q && qbase + tid < a.n) heap_store(a, hd, hi, Q, tid, qbase + tid, split);
}

// ---------------------------------------------------------------------------
// Direct form of the FP64 MFMA kernel: every wave streams the candidates on
// its own, global -> registers, already in the MFMA operand layout.  There is
// no LDS tile and no workgroup barrier inside the candidate loop, so one
// wave's filter and heap work overlaps the MFMAs of the other wave on its
// SIMD instead of holding up a workgroup at a barrier.
//
// Lane l supplies k-index g = l / 16 of every 16x16x4 product; which dim
// k-step s of group g carries is free, as long as the candidate operand and
// the query operand agree (the dot product is the same sum, taken in another
// order).  Two deals:
//  * LINE (d = 4 KS): k-steps 2j and 2j + 1 of group g are dims 8j + 2g and
//    8j + 2g + 1, so one 16-byte load per lane covers 64 contiguous bytes of
//    each of the tile's 16 candidate rows, and two consecutive step pairs
//    cover whole 128-byte lines.  The refills of tile t + 1 are issued in
//    groups of four k-steps right behind the MFMAs that free their
//    registers (a scheduling barrier keeps them there): every group reads
//    whole lines while they are in L1, and each refill has a full tile of
//    MFMAs before it is needed.
//  * otherwise, contiguous runs: group g owns dims [g KS, min((g + 1) KS, d))
//    and its k-step s is dim st + s with st = min(g KS, d - KS).  A run moved
//    back from g KS overlaps dims of the groups before it; those steps get a
//    zero query coordinate, so every dim enters the sum exactly once and
//    every lane reads KS contiguous in-bounds doubles of each candidate row.
//    The compiler schedules the refills.
// amap entries are read two tiles ahead, shifted norms (a.SN, by position)
// one tile ahead.
// The result layout, filter, survivor queue and heap updates are those of
// knn_mfma_f64_kernel; the candidate norms and ids of a lane's result values
// come from the lanes that loaded them, by lane shuffles.
// ---------------------------------------------------------------------------
template <int KS, int QB, int CB, int W, bool LINE, bool TWO>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2))) void knn_mfma_f64_direct_kernel(
    KnnArgs<double> a) {
  static_assert(KS % 4 == 0, "k-steps in groups of four");
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB;
  static_assert(QB <= 2, "query block fits one bit");
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

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, grp = lane >> 4, cl = lane & 15;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const bool sentinel = a.nsplit > 1;

  if (tid < Q && qbase + tid < n) {
    if (TWO)
      key_heap_init(a, hd, hi, r0, Q, tid, qbase + tid, kh);
    else
      heap_load(a, hd, hi, Q, tid, qbase + tid, sentinel);
  }

  const int own0 = grp * KS, own1 = min(own0 + KS, d);  // dims this lane's group owns
  const int st = LINE ? 2 * grp : min(own0, d - KS);       // first dim it reads
  // Offset of k-step s from st in either deal.
  auto step_dim = [](int s) -> int { return LINE ? 8 * (s >> 1) + (s & 1) : s; };
  const double cm = mfma_margin_c(4 * KS, d);
  double bq[QB][KS];
  double qn[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = w * QW + qb * 16 + cl;
    const int j = qbase + ql[qb];
    qv[qb] = j < n;
    const int jj = qv[qb] ? j : n - 1;
    const double *xq = a.XB + (size_t)a.bmap[jj] * d;
    qn[qb] = qv[qb] ? a.XB2[a.bmap[jj]] : 0.0;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int p = st + step_dim(s);
      bq[qb][s] = (qv[qb] && (LINE || (p >= own0 && p < own1))) ? xq[p] : 0.0;
    }
  }
  // XB row of query block qb (every lane of the wave calls it).
  auto qrow = [&](int qb) -> int { return a.bmap[qv[qb] ? qbase + ql[qb] : n - 1]; };

  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  struct Ids {
    int v[CB];
  };
  // amap entries of this lane's candidates cb * 16 + l % 16 of tile t.
  auto cand_id = [&](int t) -> Ids {
    Ids r;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) r.v[cb] = a.amap[min(cbeg + t * C + cb * 16 + cl, cend - 1)];
    return r;
  };
  auto slot_ok = [&](int t, int cb) -> bool { return cbeg + t * C + cb * 16 + cl < cend; };
  double xb[KS][CB];  // candidate operands of the current tile
  auto row = [&](const Ids &ids, int cb) -> const double * {
    return a.XA + (size_t)ids.v[cb] * d + st;
  };
  auto load_op = [&](const double *r, int s) -> double { return r[step_dim(s)]; };

  double cn[CB];  // shifted norms (a.SN) of this lane's candidates in the current tile
  int ci[CB];     // their ids (INT_MAX for empty slots)
  Ids id_a{}, id_b{};
  auto pos_sn = [&](int t, int cb) -> double { return a.SN[min(cbeg + t * C + cb * 16 + cl, cend - 1)]; };
  // Empty candidate slots carry a NaN norm: no comparison with one passes.
  const double nan = __builtin_nan("");
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    cn[cb] = nan;
    ci[cb] = INT_MAX;
  }
  if (ntiles > 0) {
    const Ids id0 = cand_id(0);
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const double *r = row(id0, cb);
#pragma unroll
      for (int s = 0; s < KS; ++s) xb[s][cb] = load_op(r, s);
      const bool ok = slot_ok(0, cb);
      cn[cb] = ok ? pos_sn(0, cb) : nan;
      ci[cb] = ok ? id0.v[cb] : INT_MAX;
    }
    id_a = ntiles > 1 ? cand_id(1) : id0;
    id_b = ntiles > 2 ? cand_id(2) : id_a;
  }
  __syncthreads();  // heaps loaded

  // Filter threshold of a query row (mfma_threshold): every pair whose
  // oracle distance can reach the root r passes; lanes without a query get
  // -inf.  The query's squared norm comes from its operand registers (the
  // four lane groups hold disjoint dims of the row); every lane of the wave
  // calls it.
  auto root_thr = [&](int qb) -> double {
    double b2 = 0.0;
#pragma unroll
    for (int s = 0; s < KS; ++s) b2 = __builtin_fma(bq[qb][s], bq[qb][s], b2);
    b2 = xadd(b2, __shfl_xor(b2, 16));
    b2 = xadd(b2, __shfl_xor(b2, 32));
    if (!qv[qb]) return -__builtin_inf();
    if (TWO) return key_threshold(r0[ql[qb]], hd[ql[qb]], qn[qb], xmul(xmul(0.5, cm), b2));
    return mfma_threshold(hd[ql[qb]], qn[qb], xmul(xmul(0.5, cm), b2));
  };
  double thr[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) thr[qb] = root_thr(qb);

  // Per-lane queue of passing pairs: candidate id, query block in the top
  // bit (ids are non-negative).
  constexpr int QR = 4;
  unsigned pq_i[QR];
  double pq_t[QR];  // two-stage: the pairs' MFMA keys
  int pq_n = 0;
#pragma unroll
  for (int r = 0; r < QR; ++r) {
    pq_i[r] = 0u;
    pq_t[r] = 0.0;
  }
  dbl4 acc[QB][CB];
  // Apply the queues: every lane forms the oracle distances of its queued
  // pairs at once (rows read from global memory), then the four lane groups,
  // which cover the wave's queries once each, update their heaps in turn.
  auto flush = [&]() __attribute__((always_inline)) {
    double ed[QR];
    const int qr0 = TWO ? 0 : qrow(0), qr1 = TWO ? 0 : qrow(QB - 1);
#pragma unroll
    for (int r = 0; r < QR; ++r) {
      ed[r] = pq_t[r];
      if (!TWO && r < pq_n) {
        const bool hi_qb = pq_i[r] >> 31;
        const int id = (int)(pq_i[r] & 0x7fffffffu);
        const double *xq = a.XB + (size_t)(hi_qb ? qr1 : qr0) * d;
        ed[r] = oracle_dist(a.XA + (size_t)id * d, xq, d, a.XA2[id], hi_qb ? qn[QB - 1] : qn[0]);
      }
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
    const Ids id_c = (t + 3 < ntiles) ? cand_id(t + 3) : id_b;
    double nn[CB];  // shifted norms of tile t + 1 (its ids are id_a)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) nn[cb] = pos_sn(t + 1 < ntiles ? t + 1 : t, cb);
    const double *nr[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) nr[cb] = row(id_a, cb);

#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = dbl4{0.0, 0.0, 0.0, 0.0};
    // Each operand register is refilled with tile t + 1's value once the
    // MFMAs of tile t have read it (on the last
    // tile the refill reads valid rows and is never used).  The wave issues
    // its MFMAs at raised priority, so of the two waves of a SIMD the one in
    // its MFMA phase keeps the matrix pipe while the other runs its epilogue.
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const double av = xb[s][cb];
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bq[qb][s], acc[qb][cb], 0, 0, 0);
        if (!LINE) xb[s][cb] = load_op(nr[cb], s);
      }
      if (LINE) {
        if ((s & 3) == 3) {
#pragma unroll
          for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int u = 3; u >= 0; --u) xb[s - u][cb] = load_op(nr[cb], s - u);
        }
        // (None after the last k-step: the first candidate block's epilogue
        // may start under the second block's MFMAs.)
        if (s + 1 < KS) __builtin_amdgcn_sched_barrier(0);
      }
    }
    __builtin_amdgcn_s_setprio(0);

    // Epilogue: one FMA and one compare per pair against the shifted norm
    // (a.SN) and threshold (mfma_threshold); wave masks in scalar registers;
    // per-lane bits and ids only when a pair passes.
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
      const bool ok = t + 1 < ntiles && slot_ok(t + 1, cb);
      cn[cb] = ok ? nn[cb] : nan;
      ci[cb] = ok ? id_a.v[cb] : INT_MAX;
    }
    id_a = id_b;
    id_b = id_c;
  }

  if (__any(pq_n > 0)) flush();
  __syncthreads();
  if (tid < Q && qbase + tid < n) {
    if (TWO)
      key_heap_store(a, hd, hi, Q, tid, qbase + tid, kh);
    else
      heap_store(a, hd, hi, Q, tid, qbase + tid, split);
  }
}

// ---------------------------------------------------------------------------
// Small candidate sets (m <= 64 P, P <= 32, k <= 64, d <= 64): one wave per
// query, four queries per workgroup, the oracle's arithmetic.  The workgroup
// stages the candidates in chunks of 64 CH rows into LDS, dim-major (row
// loads coalesced along each row; lane l then reads candidate l + 64 u of
// the chunk at dim p without bank conflicts).  Lane l computes the exact
// distances of candidates l, l + 64, ... (dot product in p order, product
// and sum rounded separately, then x(-2), + XA2[amap[i]], + XB2[bmap[j]]) and
// keeps them in registers beside entry l of the resident row.  The wave then
// takes the k smallest entries of (resident row + candidates) under the
// canonical (distance, index) order by k rounds of a wave arg-min and writes
// them descending.  That is the row the heap insertion rule yields (a
// candidate enters only when it is below the current k-th, so the row always
// holds the k smallest entries seen, duplicates included), with no heap, no
// candidate split and no merge.
// ---------------------------------------------------------------------------
constexpr int kSmallMaxK = 64;
constexpr int kSmallMaxP = 32;

// Value of the lane that DPP control CTRL selects within this lane's row
// (every lane of the wave active; CTRL is a full permutation of each row).
template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<CTRL>((int)(b & 0xffffffffll)), hi = dpp_move<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Candidates per lane per staged chunk: LDS holds the [d][64 CH + 4] tile
// (CH d <= 128: at most 68 KiB for FP64), the chunk's norms and the four
// query rows.
constexpr int small_ch(int d) { return d <= 16 ? 8 : d <= 32 ? 4 : 2; }
constexpr int small_stride(int ch) { return 64 * ch + 4; }

// Bitonic sort of one (distance, index) entry per lane across the wave,
// ascending in canonical order (lane r ends with the r-th smallest).
template <typename T>
__device__ __forceinline__ void wave_sort_asc(T &d, int &i, int lane) {
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const T od = __shfl_xor(d, stride);
      const int oi = __shfl_xor(i, stride);
      // Ascending blocks (lane & size == 0) keep the smaller entry in the
      // lower lane of each pair, descending blocks the larger.
      const bool keep_min = ((lane & size) == 0) == ((lane & stride) == 0);
      const bool take = keep_min ? cless(od, oi, d, i) : cless(d, i, od, oi);
      if (take) {
        d = od;
        i = oi;
      }
    }
  }
}

// Offer entry (x, xi) at slot xs to a lane's two smallest entries so far,
// (bd, bi, bs) <= (sd, si, ss) in canonical order; a slot of -1 is empty.
template <typename T>
__device__ __forceinline__ void top2_consider(T x, int xi, int xs, T &bd, int &bi, int &bs, T &sd,
                                              int &si, int &ss) {
  const bool first = bs < 0 || cless(x, xi, bd, bi);
  const bool second = !first && (ss < 0 || cless(x, xi, sd, si));
  sd = first ? bd : (second ? x : sd);
  si = first ? bi : (second ? xi : si);
  ss = first ? bs : (second ? xs : ss);
  bd = first ? x : bd;
  bi = first ? xi : bi;
  bs = first ? xs : bs;
}

template <typename T, int P, int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void knn_small_m_kernel(KnnArgs<T> a) {
  static_assert(P <= 32 && P % CH == 0, "candidate slots per lane");
  constexpr int C = 64 * CH;  // candidates per chunk
  // Row stride of the staged tile: C + 4 puts the 16 dims a staging store
  // writes (one row apart) in different LDS banks.
  constexpr int S = small_stride(CH);
  extern __shared__ __align__(16) unsigned char smem[];
  T *xs = (T *)smem;           // [d][S]
  T *xn = xs + (size_t)S * a.d;  // [C]: the chunk's candidate norms XA2
  T *qs = xn + C;                // [4][64]: the query rows
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int j = blockIdx.x * 4 + w;
  const bool qok = j < a.n;  // waves past n still help stage
  const int k = a.k, d = a.d, m = a.m;
  const int qid = qok ? a.bmap[j] : 0;
  if (lane < d) qs[w * 64 + lane] = a.XB[(size_t)qid * d + lane];

  int id[P];
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const int c = lane + 64 * i;
    id[i] = c < m ? a.amap[c] : -1;
  }
  T v[P];
#pragma unroll
  for (int i = 0; i < P; ++i) v[i] = T(0);
  // Staging: 16 threads per row, thread t reading dims t % 16, t % 16 + 16,
  // ... (NH column groups cover d <= 16 NH) of rows rr, rr + 16, ... (R
  // rows), and the norms of its rows r with r % 16 == t % 16.  The ids of a
  // chunk's rows are read while the previous chunk is staged and summed, so
  // a chunk costs one round trip (rows past m read row 0, never used).
  constexpr int R = C / 16, NH = 8 / CH;
  static_assert(R * NH == 32, "32 staged values per thread and chunk");
  const int p0 = tid & 15, rr = tid >> 4;
  constexpr int NE = (R + 15) / 16;
  int rid[R], nid[NE];  // the thread's rows; the rows whose norms it stages
  auto load_ids = [&](int c0) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = c0 + r * 16 + rr;
      rid[r] = c < m ? a.amap[c] : 0;
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int c = c0 + (p0 + 16 * e) * 16 + rr;
      nid[e] = p0 + 16 * e < R && c < m ? a.amap[c] : 0;
    }
  };
  load_ids(0);
#pragma clang loop unroll(full)
  for (int ch = 0; ch < P / CH; ++ch) {
    const int c0 = ch * C;
    if (c0 < m) {  // uniform
      T x[R][NH];
      T xq[NE];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int h = 0; h < NH; ++h)
          if (p0 + 16 * h < d) x[r][h] = a.XA[(size_t)rid[r] * d + p0 + 16 * h];
#pragma unroll
      for (int e = 0; e < NE; ++e) xq[e] = a.XA2[nid[e]];
      if (c0 + C < m) load_ids(c0 + C);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int h = 0; h < NH; ++h)
          if (p0 + 16 * h < d) xs[(p0 + 16 * h) * S + r * 16 + rr] = x[r][h];
#pragma unroll
      for (int e = 0; e < NE; ++e)
        if (p0 + 16 * e < R) xn[(p0 + 16 * e) * 16 + rr] = xq[e];
      __syncthreads();
      if (qok) {
#pragma unroll 4
        for (int p = 0; p < d; ++p) {
          const T qp = qs[w * 64 + p];
#pragma unroll
          for (int u = 0; u < CH; ++u)
            v[ch * CH + u] = xadd(v[ch * CH + u], xmul(xs[p * S + lane + 64 * u], qp));
        }
        // The oracle's epilogue: x(-2), + XA2[amap[i]] (the query norm later).
#pragma unroll
        for (int u = 0; u < CH; ++u)
          v[ch * CH + u] = xadd(xmul(v[ch * CH + u], T(-2)), xn[lane + 64 * u]);
      }
      __syncthreads();
    }
  }
  if (!qok) return;
  const T qn = a.XB2[qid];
  // Taken entries and empty slots: bit i of tk for candidate slot i, rtk for
  // the resident entry.
  unsigned tk = 0u;
  bool rtk = false;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    if (lane + 64 * i < m) {
      v[i] = xadd(v[i], qn);
    } else {
      tk |= 1u << i;
      v[i] = T(__builtin_inf());  // empty slots sort after every entry
      id[i] = INT_MAX;
    }
  }
  const size_t rbase = (size_t)heap_row(a, j) * a.ldk;
  T rv = T(0);
  int ri = -1;
  if (lane < k) {
    rv = a.D[rbase + lane];
    ri = a.I[rbase + lane];
  } else {
    rtk = true;
  }

  // Selection by sorting.  T = the k-th smallest of the 64 lane minima (k
  // distinct entries, so T is at or above the row's final k-th entry); the
  // entries at or below T in canonical order (few more than k while k is
  // well below 64) go to LDS, one per lane after a wave prefix sum, and one
  // bitonic sort over the wave orders them.  More than 64 such entries (k
  // near 64, or many ties): the rounds below.
  {
    T md = __builtin_inf();
    int mi = INT_MAX;
    if (!rtk) {
      md = rv;
      mi = ri;
    }
#pragma unroll
    for (int i = 0; i < P; ++i)
      if (cless(v[i], id[i], md, mi)) {
        md = v[i];
        mi = id[i];
      }
    wave_sort_asc(md, mi, lane);
    const T Td = __shfl(md, k - 1);
    const int Ti = __shfl(mi, k - 1);
    // Entries at or below T: bit i for candidate slot i, rsel for the
    // resident entry.
    const bool rsel = !rtk && !cless(Td, Ti, rv, ri);
    unsigned sel = 0u;
#pragma unroll
    for (int i = 0; i < P; ++i) sel |= (!cless(Td, Ti, v[i], id[i]) ? 1u : 0u) << i;
    sel &= ~tk;
    const int cnt = __popc(sel) + (rsel ? 1 : 0);
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    const int tot = __shfl(inc, 63);
    if (tot <= 64) {
      // Scratch after the chunk loop's last barrier (no wave reads the tile
      // or the query rows any more): 64 entries per wave.
      T *sd_ = (T *)smem + w * 64;
      int *si_ = (int *)(smem + 256 * sizeof(T)) + w * 64;
      int pos = inc - cnt;
      if (rsel) {
        sd_[pos] = rv;
        si_[pos] = ri;
        ++pos;
      }
#pragma unroll
      for (int i = 0; i < P; ++i)
        if ((sel >> i) & 1u) {
          sd_[pos] = v[i];
          si_[pos] = id[i];
          ++pos;
        }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      T ed = lane < tot ? sd_[lane] : T(__builtin_inf());
      int ei = lane < tot ? si_[lane] : INT_MAX;
      wave_sort_asc(ed, ei, lane);
      if (lane < k) {  // rows are written descending
        a.D[rbase + k - 1 - lane] = ed;
        a.I[rbase + k - 1 - lane] = ei;
      }
      return;
    }
  }

  // This lane's smallest entry not yet taken: (bd, bi) at slot bs (P for the
  // resident entry, -1 when none is left), and the next one (sd, si) at
  // slot ss (-1 when none, -2 when not known).  A lane wins rarely (k
  // rounds over 64 lanes), so a win promotes the next entry and only a
  // second win rescans the lane's entries.
  T bd = T(0), sd = T(0);
  int bi = 0, bs = -1, si = 0, ss = -1;
  auto rescan = [&]() __attribute__((always_inline)) {
    bs = -1;
    ss = -1;
    if (!rtk) top2_consider(rv, ri, P, bd, bi, bs, sd, si, ss);
#pragma unroll
    for (int i = 0; i < P; ++i)
      if (!((tk >> i) & 1u)) top2_consider(v[i], id[i], i, bd, bi, bs, sd, si, ss);
  };
  rescan();
  T outd = T(0);
  int outi = -1;
  for (int r = 0; r < k; ++r) {
    // Wave arg-min of (bd, bi, lane); lanes with nothing left never win.
    // Within each row of 16 lanes by DPP moves (partner lane l ^ 1, l ^ 2,
    // then the mirror in each half row and in the row: each step pairs two
    // groups that already hold their own minimum), across rows by lane
    // shuffles (l ^ 16, l ^ 32).
    T md = bd;
    int mi = bi, ml = bs < 0 ? 64 : lane;
    auto combine = [&](T od, int oi, int ol) {
      const bool take = ol < 64 && (ml == 64 || cless(od, oi, md, mi) ||
                                    (od == md && oi == mi && ol < ml));
      if (take) {
        md = od;
        mi = oi;
        ml = ol;
      }
    };
    combine(dpp_move<0xB1>(md), dpp_move<0xB1>(mi), dpp_move<0xB1>(ml));     // quad_perm 1,0,3,2
    combine(dpp_move<0x4E>(md), dpp_move<0x4E>(mi), dpp_move<0x4E>(ml));     // quad_perm 2,3,0,1
    combine(dpp_move<0x141>(md), dpp_move<0x141>(mi), dpp_move<0x141>(ml));  // row_half_mirror
    combine(dpp_move<0x140>(md), dpp_move<0x140>(mi), dpp_move<0x140>(ml));  // row_mirror
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      const T od = __shfl_xor(md, o);
      const int oi = __shfl_xor(mi, o);
      const int ol = __shfl_xor(ml, o);
      combine(od, oi, ol);
    }
    if (lane == k - 1 - r) {  // rows are written descending
      outd = md;
      outi = mi;
    }
    if (lane == ml) {
      if (bs == P)
        rtk = true;
      else
        tk |= 1u << bs;
      if (ss != -2) {  // the next entry is known (or there is none)
        bd = sd;
        bi = si;
        bs = ss;
        ss = -2;
      } else {
        rescan();
      }
    }
  }
  if (lane < k) {
    a.D[rbase + lane] = outd;
    a.I[rbase + lane] = outi;
  }
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
  // The head entries of up to kHeads lists are loaded together (independent
  // loads in flight at once); a list is read further only while it beats the
  // root, which few lists do.
  constexpr int kHeads = 8;
  for (int s0 = 0; s0 < a.nsplit; s0 += kHeads) {
    T h_d[kHeads];
    int h_i[kHeads];
#pragma unroll
    for (int u = 0; u < kHeads; ++u) {
      const size_t o = ((size_t)min(s0 + u, a.nsplit - 1) * k) * a.n + j;
      h_d[u] = a.partD[o];
      h_i[u] = a.partI[o];
    }
#pragma unroll
    for (int u = 0; u < kHeads; ++u) {
      if (s0 + u >= a.nsplit) break;
      T xd = h_d[u];
      int xi = h_i[u];
      for (int t = 0; cless(xd, xi, rd, ri);) {
        sift_down(hd, hi, QPB, q, 0, k, xd, xi);
        rd = hd[q];
        ri = hi[q];
        if (++t >= k) break;
        const size_t o = ((size_t)(s0 + u) * k + t) * a.n + j;
        xd = a.partD[o];
        xi = a.partI[o];
      }
    }
  }
  KnnArgs<T> b = a;
  b.nsplit = 1;
  heap_store(b, hd, hi, QPB, q, j, 0);
}

// Merge for calls with few queries: one wave per query over query-major
// partial lists.  The 64 lanes load the resident row and every partial list
// at once (up to 64 * R entries, the lists contiguous in memory), drop
// list entries that do not beat the resident row's maximum (the split
// launches' sentinels among them), and extract the k smallest in canonical
// order by k rounds of a wave-wide arg-min; identical (distance, index) pairs
// are distinct entries, broken by slot.  Same result as knn_merge_kernel.
template <typename T, int R>
__global__ __launch_bounds__(256) void knn_merge_wave_kernel(KnnArgs<T> a) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= a.n) return;  // whole wave
  const int k = a.k;
  const size_t row = (size_t)heap_row(a, j) * a.ldk;
  const int E = k * (1 + a.nsplit);
  T ed[R];
  int ei[R];
  // Entry e < k: resident slot e; e >= k: list (e - k) / k, position (e - k) % k.
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = r * 64 + lane;
    ed[r] = (T)__builtin_inf();
    ei[r] = INT_MAX;
    if (e < k) {
      ed[r] = a.D[row + e];
      ei[r] = a.I[row + e];
    } else if (e < E) {
      // query-major lists: this query's entries are contiguous
      const size_t o = (size_t)j * (E - k) + (e - k);
      ed[r] = a.partD[o];
      ei[r] = a.partI[o];
    }
  }
  // Resident maximum (md, mi): list entries must beat it strictly.
  T md = (T)-__builtin_inf();
  int mi = INT_MIN;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r * 64 + lane < k && cless(md, mi, ed[r], ei[r])) {
      md = ed[r];
      mi = ei[r];
    }
  for (int off = 32; off > 0; off >>= 1) {
    const T od = __shfl_xor(md, off);
    const int oi = __shfl_xor(mi, off);
    if (cless(md, mi, od, oi)) {
      md = od;
      mi = oi;
    }
  }
  bool live[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = r * 64 + lane;
    live[r] = e < k || (e < E && cless(ed[r], ei[r], md, mi));
  }
  // k rounds: each lane offers its smallest live entry; the wave minimum by
  // (distance, index, lane) is written and retired.  Each lane keeps its two
  // smallest live entries (slot -1: none, -2: not known), so a win promotes
  // the second and only a lane's second win rescans its entries; the
  // arg-min steps within a 16-lane row are DPP moves, across rows shuffles.
  T ld = T(0), nd = T(0);
  int li = 0, lr = -1, ni = 0, nr = -1;
  auto rescan = [&]() __attribute__((always_inline)) {
    lr = -1;
    nr = -1;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (live[r]) top2_consider(ed[r], ei[r], r, ld, li, lr, nd, ni, nr);
  };
  rescan();
  for (int t = 0; t < k; ++t) {
    T bd = ld;
    int bi = li, bl = lr >= 0 ? lane : 64;
    auto combine = [&](T od, int oi, int ol) {
      if (ol < 64 && (bl == 64 || cless(od, oi, bd, bi) || (od == bd && oi == bi && ol < bl))) {
        bd = od;
        bi = oi;
        bl = ol;
      }
    };
    combine(dpp_move<0xB1>(bd), dpp_move<0xB1>(bi), dpp_move<0xB1>(bl));     // quad_perm 1,0,3,2
    combine(dpp_move<0x4E>(bd), dpp_move<0x4E>(bi), dpp_move<0x4E>(bl));     // quad_perm 2,3,0,1
    combine(dpp_move<0x141>(bd), dpp_move<0x141>(bi), dpp_move<0x141>(bl));  // row_half_mirror
    combine(dpp_move<0x140>(bd), dpp_move<0x140>(bi), dpp_move<0x140>(bl));  // row_mirror
    for (int off = 16; off < 64; off <<= 1) {
      const T od = __shfl_xor(bd, off);
      const int oi = __shfl_xor(bi, off);
      const int ol = __shfl_xor(bl, off);
      combine(od, oi, ol);
    }
    if (lane == bl) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r == lr) live[r] = false;
      if (nr != -2) {
        ld = nd;
        li = ni;
        lr = nr;
        nr = -2;
      } else {
        rescan();
      }
    }
    // Rows are written descending, as heap_store does.
    if (lane == 0) {
      a.D[row + (k - 1 - t)] = bd;
      a.I[row + (k - 1 - t)] = bi;
    }
  }
}

// Two-stage FP64 MFMA form, stage 2 (the contract is stated with
// key_heap_init): one wave per query.  Lanes form the oracle distances of
// the key heap's entries (one entry per lane, rows read from global memory;
// many waves per SIMD hide the latency), the wave sorts resident row + heap
// entries (P = next power of two >= k + k2, padding (+inf, INT_MAX) last)
// with a bitonic network in LDS and takes the first k.  The row is written
// (descending, as every kernel of this library leaves it) when the heap
// never filled or its root key T exceeds mfma_threshold(E) for the k-th
// distance E; otherwise the row stays as it was and the query goes to the
// fallback list (its XB row and heap row), for the one-stage kernel.
__global__ __launch_bounds__(256) void knn_two_stage_finalize_kernel(KnnArgs<double> a, int terms,
                                                                     int P, int *fail_n,
                                                                     int *fail_b, int *fail_h) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  double *sd = (double *)smem + (size_t)w * P;
  int *si = (int *)(smem + al16(sizeof(double) * 4 * (size_t)P)) + (size_t)w * P;
  if (j >= a.n) return;  // whole wave; only wave barriers below
  const int k = a.k, kh = a.k2, d = a.d;
  const int hrow = a.hrow ? a.hrow[j] : j;
  const size_t row = (size_t)hrow * a.ldk;
  const int qr = a.bmap[j];
  const double *xq = a.XB + (size_t)qr * d;
  const double qn = a.XB2[qr];
  const double *hk = a.hk + (size_t)j * kh;
  const int *hid = a.hid + (size_t)j * kh;
  for (int e = lane; e < kh; e += 64) {
    const int id = hid[e];
    double x = __builtin_inf();
    if (id != INT_MAX) x = oracle_dist(a.XA + (size_t)id * d, xq, d, a.XA2[id], qn);
    sd[k + e] = x;
    si[k + e] = id;
  }
  for (int e = lane; e < k; e += 64) {
    sd[e] = a.D[row + e];
    si[e] = a.I[row + e];
  }
  for (int e = k + kh + lane; e < P; e += 64) {
    sd[e] = __builtin_inf();
    si[e] = INT_MAX;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int half = P / 2;
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = lane; q < half; q += 64) {
        const int lo = 2 * stride * (q / stride) + q % stride;
        const int hi = lo + stride;
        const double x = sd[lo], y = sd[hi];
        const int ix = si[lo], iy = si[hi];
        const bool up = (lo & size) == 0;
        if (up ? cless(y, iy, x, ix) : cless(x, ix, y, iy)) {
          sd[lo] = y;
          sd[hi] = x;
          si[lo] = iy;
          si[hi] = ix;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  bool ok = true;
  if (hid[0] != INT_MAX) {
    // The query's squared norm for the margin (any rounding order: the
    // margin constant carries the slack).
    double b2 = 0.0;
    for (int p = lane; p < d; p += 64) b2 = __builtin_fma(xq[p], xq[p], b2);
    for (int off = 32; off >= 1; off >>= 1) b2 = xadd(b2, __shfl_xor(b2, off));
    const double cm = mfma_margin_c(terms, d);
    ok = hk[0] > mfma_threshold(sd[k - 1], qn, xmul(xmul(0.5, cm), b2));
  }
  if (ok) {
    for (int t = lane; t < k; t += 64) {
      a.D[row + t] = sd[k - 1 - t];
      a.I[row + t] = si[k - 1 - t];
    }
  } else if (lane == 0) {
    const int f = atomicAdd(fail_n, 1);
    fail_b[f] = qr;
    fail_h[f] = hrow;
  }
}

// Rows to ascending (distance, index) order, in place: each 64-lane
// workgroup sorts R rows of P = next power of two >= k entries with a
// bitonic network in LDS (padding (+inf, INT_MAX) sorts last).
template <typename T>
__global__ __launch_bounds__(64) void rows_ascending_kernel(int n, int k, T *D, int *I, int ldk,
                                                            int P, int R) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *sd = (T *)smem;
  int *si = (int *)(smem + al16(sizeof(T) * (size_t)P * R));
  const int r0 = blockIdx.x * R;
  for (int t = threadIdx.x; t < P * R; t += 64) {
    const int r = r0 + t / P, e = t % P;
    const bool v = r < n && e < k;
    sd[t] = v ? D[(size_t)r * ldk + e] : (T)__builtin_inf();
    si[t] = v ? I[(size_t)r * ldk + e] : INT_MAX;
  }
  __syncthreads();
  const int half = P / 2;
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < half * R; t += 64) {
        const int base = (t / half) * P, q = t % half;
        const int lo = 2 * stride * (q / stride) + q % stride;
        const int hi = lo + stride;
        const T a = sd[base + lo], b = sd[base + hi];
        const int ia = si[base + lo], ib = si[base + hi];
        const bool up = (lo & size) == 0;
        if (up ? cless(b, ib, a, ia) : cless(a, ia, b, ib)) {
          sd[base + lo] = b;
          sd[base + hi] = a;
          si[base + lo] = ib;
          si[base + hi] = ia;
        }
      }
      __syncthreads();
    }
  for (int t = threadIdx.x; t < P * R; t += 64) {
    const int r = r0 + t / P, e = t % P;
    if (r < n && e < k) {
      D[(size_t)r * ldk + e] = sd[t];
      I[(size_t)r * ldk + e] = si[t];
    }
  }
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

// Shifted candidate norms by position for the FP64 MFMA filter:
// SN[i] = mfma_shift_norm(XA2[amap[i]], |XA[amap[i]]|^2, c) with c the
// margin constant of `terms` MFMA terms.  One wave per row (coalesced row
// loads, a wave sum; any summation order, the margin has slack for it).
__global__ __launch_bounds__(256) void cand_shift_kernel(int m, int d, int terms, const double *XA,
                                                         const double *XA2, const int *amap,
                                                         double *SN, int *zero) {
  const int lane = threadIdx.x & 63;
  // The two-stage form's fallback counter, cleared by a kernel (a launch
  // sequence captured into a HIP graph replays it).
  if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
  const int nw = gridDim.x * 4;
  const double cm = mfma_margin_c(terms, d);
  for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < m; i += nw) {
    const int id = amap[i];
    const double *x = XA + (size_t)id * d;
    double s = 0.0;
    for (int p = lane; p < d; p += 64) s = __builtin_fma(x[p], x[p], s);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) SN[i] = mfma_shift_norm(XA2[id], s, cm);
  }
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
  int ct = 0;    // exact: candidates per LDS tile (0: scalar-candidate form)
  int ks = 0, qb = 0, w = 0, nb = 1;  // mfma
  bool direct = false;                // mfma: a per-wave form (no LDS tile)
  bool ring = false;                  // mfma: the operand-ring form (gsknn_ring.hip)
  int per_cu = 0;                     // direct form: resident workgroups per CU
  int small = 0;                      // exact, one wave per query: candidate slots per lane
  size_t lds = 0;
  int gx = 0, nsplit = 1, cps = 0;
  // FP64 MFMA per-wave forms, unsplit: the two-stage form's heap size and
  // LDS bytes (k2 = 0: not available).  The launch takes it when the
  // workspace holds its heap lists.
  int k2 = 0;
  size_t lds2 = 0;
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

// Direct form: query blocks and candidate blocks per wave; the query
// fragments (QB KS doubles) and the operand registers (CB KS doubles) must
// fit two waves per SIMD.
constexpr int direct_qb(int ks) { return ks <= 16 ? 2 : 1; }
constexpr int direct_cb(int ks) { return ks <= 16 ? 2 : 1; }

// LDS of the direct form: the Q heap rows only.
size_t mfma_direct_lds(int Q, int k) { return al16(sizeof(double) * (size_t)k * Q) + sizeof(int) * (size_t)k * Q; }

size_t merge_lds(int k, int elem, int qpb) {
  return al16((size_t)elem * k * qpb) + sizeof(int) * (size_t)k * qpb;
}

int merge_qpb(int k, int elem) {
  for (int qpb : {64, 32, 16})
    if (merge_lds(k, elem, qpb) <= kLdsMax) return qpb;
  return 0;
}

bool verbose() {
  static const bool v = [] {
    const char *e = getenv("GSKNN_VERBOSE");
    return e && *e && *e != '0';
  }();
  return v;
}

// MFMA launch shapes, in order of preference at equal occupancy: more
// queries per wave (each LDS fragment feeds QB MFMAs), then two tile buffers.
constexpr int kMfmaCfg[7][3] = {{2, 4, 1}, {2, 4, 2}, {2, 8, 2}, {1, 4, 2}, {1, 4, 1}, {1, 2, 1}, {1, 1, 1}};
constexpr int kNumMfmaCfg = 7;

// Workgroups per CU that registers allow; 0 for an instantiation that
// spills to scratch (never chosen).
template <int KS, int QB, int W, int NB>
int mfma_reg_blocks() {
  const void *kern = (const void *)knn_mfma_f64_kernel<KS, QB, mfma_cb(KS), W, NB>;
  hipFuncAttributes at;
  if (hipFuncGetAttributes(&at, kern) == hipSuccess && at.localSizeBytes > 0) return 0;
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, 64 * W, 0) != hipSuccess || b <= 0)
    b = 8 / W;  // two waves per SIMD
  return b;
}

template <int KS, int C>
int mfma_reg_blocks_cfg() {
  return mfma_reg_blocks<KS, kMfmaCfg[C][0], kMfmaCfg[C][1], kMfmaCfg[C][2]>();
}

// Workgroups per CU that registers and wave slots allow (LDS aside), per
// kernel instantiation, queried once.
int mfma_blocks_per_cu(int ks, int cfg) {
  static std::atomic<int> cache[5][kNumMfmaCfg];
  const int ki = ks == 4 ? 0 : ks == 8 ? 1 : ks == 16 ? 2 : ks == 32 ? 3 : 4;
  int b = cache[ki][cfg].load(std::memory_order_relaxed);  // blocks + 1
  if (b > 0) return b - 1;
#define GSKNN_OCC(KS_)                                                     \
  case KS_:                                                                \
    switch (cfg) {                                                         \
      case 0: b = mfma_reg_blocks_cfg<KS_, 0>(); break;                    \
      case 1: b = mfma_reg_blocks_cfg<KS_, 1>(); break;                    \
      case 2: b = mfma_reg_blocks_cfg<KS_, 2>(); break;                    \
      case 3: b = mfma_reg_blocks_cfg<KS_, 3>(); break;                    \
      case 4: b = mfma_reg_blocks_cfg<KS_, 4>(); break;                    \
      case 5: b = mfma_reg_blocks_cfg<KS_, 5>(); break;                    \
      default: b = mfma_reg_blocks_cfg<KS_, 6>(); break;                   \
    }                                                                      \
    break;
  switch (ks) {
    GSKNN_OCC(4)
    GSKNN_OCC(8)
    GSKNN_OCC(16)
    GSKNN_OCC(32)
    GSKNN_OCC(64)
  }
#undef GSKNN_OCC
  cache[ki][cfg].store(b + 1, std::memory_order_relaxed);
  if (verbose()) fprintf(stderr, "gsknn: ks=%d cfg=%d: %d workgroups/CU by registers\n", ks, cfg, b);
  return b;
}

// Direct-form workgroups (one wave each) per CU that registers allow; 0 for
// an instantiation that spills.
template <int KS, bool LINE>
int mfma_direct_reg_blocks() {
  const void *kern = (const void *)knn_mfma_f64_direct_kernel<KS, direct_qb(KS), direct_cb(KS), 1, LINE, false>;
  hipFuncAttributes at;
  if (hipFuncGetAttributes(&at, kern) == hipSuccess && at.localSizeBytes > 0) return 0;
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, 64, 0) != hipSuccess || b <= 0) b = 8;
  return b;
}

int mfma_direct_blocks_per_cu(int ks, bool line) {
  static std::atomic<int> cache[8];
  const int ki = (ks == 4 ? 0 : ks == 8 ? 1 : ks == 16 ? 2 : 3) + (line ? 4 : 0);
  int b = cache[ki].load(std::memory_order_relaxed);  // blocks + 1
  if (b > 0) return b - 1;
  if (line)
    b = ks == 4 ? mfma_direct_reg_blocks<4, true>() : ks == 8 ? mfma_direct_reg_blocks<8, true>()
        : ks == 16 ? mfma_direct_reg_blocks<16, true>() : mfma_direct_reg_blocks<32, true>();
  else
    b = ks == 4 ? mfma_direct_reg_blocks<4, false>() : ks == 8 ? mfma_direct_reg_blocks<8, false>()
        : ks == 16 ? mfma_direct_reg_blocks<16, false>() : mfma_direct_reg_blocks<32, false>();
  cache[ki].store(b + 1, std::memory_order_relaxed);
  if (verbose()) fprintf(stderr, "gsknn: direct ks=%d line=%d: %d workgroups/CU by registers\n", ks, (int)line, b);
  return b;
}

// AUTO's choice of the FP64 MFMA forms (every form returns the oracle's
// rows; this is speed only).  Measured with empty rows: MFMA for d >= 32;
// below, exact calls small enough for the one-wave-per-query kernel stay
// exact (m = n = 2048, d = 16: 0.10 against 0.35 ms); the scalar-candidate
// dims 8 and 16 stay exact up to 2^26 pairs (m = n = 4097, d = 16: 0.20
// against 0.34 ms; 16384^2: 2.31 against 1.49 ms), the other dims from 4 up
// take MFMA from 2^22 pairs (4097^2, d = 24: 0.51 against 0.35 ms; 65536^2,
// d = 4 / 8: 12.7 / 13.1 against 5.9 / 6.1 ms).
bool auto_mfma(int m, int n, int d, int k) {
  if (d >= 32) return true;
  if (d < 4) return false;
  if (m <= 64 * kSmallMaxP && k <= kSmallMaxK) return false;
  const double pairs = (double)m * (double)n;
  return pairs >= ((d == 8 || d == 16) ? 67108864.0 : 4194304.0);
}

Plan make_plan(int m, int n, int d, int k, int elem, int mode) {
  Plan p;
  const bool f64 = elem == 8;
  const int ks = f64 ? mfma_ks(d) : 0;
  // Above d = 256 only the operand-ring kernel has an MFMA form.
  const bool mfma = f64 && (ks > 0 || (ring_supported(d) && mode != MODE_MFMA_TILE)) &&
                    (mode == MODE_MFMA || mode == MODE_MFMA_TILE ||
                     (mode == MODE_AUTO && auto_mfma(m, n, d, k)));
  int tile_c = 0;
  if (mfma) {
    p.kind = 1;
    p.ks = ks;
    // Score = resident waves per CU (LDS and registers) x QB (MFMAs fed by
    // one LDS fragment read); ties go to the earlier entry of kMfmaCfg.
    // GSKNN_MFMA_CFG=<index> pins one shape (tuning).
    static const int pinned = [] {
      const char *e = getenv("GSKNN_MFMA_CFG");
      return e ? atoi(e) : -1;
    }();
    int best = 0, best_waves = 0;
    for (int c = 0; c < kNumMfmaCfg && ks > 0; ++c) {
      if (pinned >= 0 && pinned < kNumMfmaCfg && c != pinned) continue;
      const int qb = kMfmaCfg[c][0], w = kMfmaCfg[c][1], nb = kMfmaCfg[c][2];
      const int Q = w * 16 * qb;
      const size_t b = MfmaLayout(ks, mfma_cb(ks), Q, k, nb).bytes;
      if (b > kLdsMax) continue;
      const int by_lds = (int)(kLdsMax / b);
      const int blocks = std::min(by_lds, mfma_blocks_per_cu(ks, c));
      const int score = blocks * w * qb;
      if (score > best) {
        best = score;
        best_waves = blocks * w;
        p.qb = qb;
        p.w = w;
        p.nb = nb;
        p.qpb = Q;
        p.lds = b;
      }
    }
    tile_c = ks > 0 ? 16 * mfma_cb(ks) : 0;
    // The direct form (d <= 128): one wave of 16 QB queries per workgroup
    // (QB = 2 up to d = 64, 1 above), only the heap rows in LDS.  Taken by
    // MODE_MFMA wherever it fits; by MODE_AUTO when it keeps at least as many
    // waves resident per CU as the LDS-tile kernel's shape (measured faster
    // at equal residency for d = 64, 96 and 128).
    // GSKNN_MFMA_DIRECT=0 turns it off, =2 takes it wherever it fits
    // (tuning).
    static const int direct_env = [] {
      const char *e = getenv("GSKNN_MFMA_DIRECT");
      return e ? atoi(e) : 1;
    }();
    const bool direct_on = direct_env != 0;
    // The operand-ring form for 64 < d <= 256 (measured faster than the
    // direct and LDS-tile forms there; slower at d <= 64, where the direct
    // form keeps two waves per SIMD) wherever its heaps fit and it does not
    // spill.  GSKNN_MFMA_RING=0 turns it off, =2 takes it for every d it
    // supports (tuning).
    static const int ring_env = [] {
      const char *e = getenv("GSKNN_MFMA_RING");
      return e ? atoi(e) : 1;
    }();
    if (direct_on && ring_env != 0 && (d > 64 || ring_env == 2) && pinned < 0 &&
        mode != MODE_MFMA_TILE && ring_supported(d)) {
      const size_t b = ring_lds_bytes(d, k, false);
      if (b <= kLdsMax) {
        const int blocks = std::min((int)(kLdsMax / b), ring_reg_blocks(d, d % 2 == 0));
        if (blocks > 0) {
          p.direct = true;
          p.ring = true;
          p.qb = ring_queries_per_block(d) / 16;
          p.w = 1;
          p.nb = 0;
          p.qpb = 16 * p.qb;
          p.lds = b;
          p.per_cu = blocks;
          tile_c = ring_candidates_per_tile(d);
        }
      }
    }
    if (!p.ring && direct_on && pinned < 0 && mode != MODE_MFMA_TILE && ks <= 32 && d >= ks) {
      const int qb = direct_qb(ks), Q = 16 * qb;
      const size_t b = mfma_direct_lds(Q, k);
      if (b <= kLdsMax) {
        const int blocks = std::min((int)(kLdsMax / b), mfma_direct_blocks_per_cu(ks, d == 4 * ks));
        if (blocks > 0 && (blocks >= best_waves || direct_env == 2 || mode == MODE_MFMA)) {
          best = qb * blocks;
          p.direct = true;
          p.qb = qb;
          p.w = 1;
          p.nb = 0;
          p.qpb = Q;
          p.lds = b;
          p.per_cu = blocks;
          tile_c = 16 * direct_cb(ks);
        }
      }
    }
    if (p.qb == 0) p.kind = 0;  // heap too large for LDS: exact kernel
    if (verbose())
      fprintf(stderr, "gsknn: mfma plan d=%d k=%d: ks=%d qb=%d w=%d nb=%d direct=%d ring=%d lds=%zu score=%d\n",
              d, k, ks, p.qb, p.w, p.nb, (int)p.direct, (int)p.ring, p.lds, best);
  }
  if (p.kind == 0) {
    // d in {8, 16, 32, 64}: the scalar-candidate form (heaps only in LDS).
    const bool scalar = scalar_dim(d);
    p.ct = scalar ? 0 : exact_ct(d, elem);
    p.qpb = 0;
    for (int qpb : {64, 32, 16}) {
      const bool qreg = d <= kExactDQ;
      size_t b;
      if (scalar)
        b = f64 ? ExactScalarLayout<double>(k, qpb).bytes : ExactScalarLayout<float>(k, qpb).bytes;
      else
        b = f64 ? ExactLayout<double>(k, qpb, d, p.ct, qreg).bytes
                : ExactLayout<float>(k, qpb, d, p.ct, qreg).bytes;
      if (b <= kLdsMax) {
        p.qpb = qpb;
        p.lds = b;
        break;
      }
    }
    // Scalar form: a split keeps at least 8 x 8 = 64 candidates.
    tile_c = scalar ? 2 * kScalarTC : p.ct;
  }
  if (p.qpb == 0) return p;
  p.gx = (n + p.qpb - 1) / p.qpb;
  // Split the candidates when the query blocks alone cannot fill the chip;
  // each split keeps at least 8 tiles of work.  The scalar-candidate form
  // waits on scalar loads and wants every resident wave LDS allows, over
  // two rounds of the chip.
  int target = kTargetBlocks;
  if (p.direct) target = 256 * p.per_cu;
  if (p.kind == 0 && p.ct == 0) {
    const int per_cu = (int)std::min<size_t>(16, kLdsMax / p.lds);
    target = 2 * 256 * per_cu;
  }
  int ns = 1;
  if (p.gx < target && merge_qpb(k, elem) > 0) {
    ns = (target + p.gx - 1) / p.gx;
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
  // The two-stage form (key heaps of k2 entries, oracle distances formed
  // afterwards for those entries only) for unsplit operand-ring launches,
  // when its heaps keep the same residency: measured 25 .. 30 % faster than
  // one stage at d = 128 and 256 with empty rows, 7 .. 8 % with carried-over
  // rows.  The direct form (d <= 64) stays in one stage: with the config-2
  // carried-over rows the finalize pass costs more than it saves (9.71
  // against 9.53 ms).  GSKNN_MFMA_TWO=0 turns the form off, =2 takes it for
  // the direct form too (tuning).
  static const int two_env = [] {
    const char *e = getenv("GSKNN_MFMA_TWO");
    return e ? atoi(e) : 1;
  }();
  if (two_env != 0 && p.kind == 1 && p.direct && (p.ring || two_env == 2) && p.nsplit == 1) {
    const int k2 = two_stage_k2(k);
    const size_t b = p.ring ? ring_lds_bytes(d, k, true) : MfmaHeapLds(p.qpb, k2, true).bytes;
    if (b <= kLdsMax && (int)(kLdsMax / b) >= p.per_cu && two_stage_sort_size(k, k2) <= kTwoMaxP) {
      p.k2 = k2;
      p.lds2 = b;
    }
  }
  // Exact calls too small to fill the chip with one thread per query
  // (they would split the candidates and merge) and with m <= 2048, k <= 64,
  // d <= 64:
  // one wave per query instead (knn_small_m_kernel), no split.
  // GSKNN_SMALL_M=0 turns it off (tuning).
  static const bool small_on = [] {
    const char *e = getenv("GSKNN_SMALL_M");
    return !(e && *e == '0');
  }();
  if (small_on && p.kind == 0 && p.nsplit > 1 && m <= 64 * kSmallMaxP && k <= kSmallMaxK &&
      d <= 64) {
    p.small = m <= 512 ? 8 : m <= 1024 ? 16 : 32;
    p.qpb = 4;
    p.gx = (n + 3) / 4;
    p.lds = 0;
    p.nsplit = 1;
    p.cps = m;
  }
  return p;
}

template <typename K>
hipError_t set_lds(K kernel, size_t bytes);

template <int KS>
hipError_t launch_mfma_direct_ks(const Plan &p, const KnnArgs<double> &a, hipStream_t s, bool two) {
  const dim3 grid(p.gx, p.nsplit);
  hipError_t e = hipSuccess;
  constexpr int QB = direct_qb(KS), CB = direct_cb(KS);
#define GSKNN_DIRECT_CASE(W_, LINE_, TWO_)                                    \
  if (p.w == W_ && (a.d == 4 * KS) == LINE_ && two == TWO_) {                 \
    auto kern = knn_mfma_f64_direct_kernel<KS, QB, CB, W_, LINE_, TWO_>;      \
    if ((e = set_lds(kern, p.lds)) != hipSuccess) return e;                   \
    hipLaunchKernelGGL(kern, grid, dim3(64 * W_), p.lds, s, a);               \
    return hipGetLastError();                                                 \
  }
  GSKNN_DIRECT_CASE(1, true, false)
  GSKNN_DIRECT_CASE(1, false, false)
  GSKNN_DIRECT_CASE(1, true, true)
  GSKNN_DIRECT_CASE(1, false, true)
#undef GSKNN_DIRECT_CASE
  return hipErrorNotSupported;
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
#define GSKNN_MFMA_CASE(QB_, W_, NB_)                                         \
  if (p.qb == QB_ && p.w == W_ && p.nb == NB_) {                              \
    auto kern = knn_mfma_f64_kernel<KS, QB_, mfma_cb(KS), W_, NB_>;           \
    if ((e = set_lds(kern, p.lds)) != hipSuccess) return e;                   \
    hipLaunchKernelGGL(kern, grid, dim3(64 * W_), p.lds, s, a);               \
    return hipGetLastError();                                                 \
  }
  GSKNN_MFMA_CASE(2, 8, 2)
  GSKNN_MFMA_CASE(2, 4, 2)
  GSKNN_MFMA_CASE(2, 4, 1)
  GSKNN_MFMA_CASE(1, 4, 2)
  GSKNN_MFMA_CASE(1, 4, 1)
  GSKNN_MFMA_CASE(1, 2, 1)
  GSKNN_MFMA_CASE(1, 1, 1)
#undef GSKNN_MFMA_CASE
  return hipErrorNotSupported;
}

template <typename T>
hipError_t launch_small_m(const Plan &p, const KnnArgs<T> &a, hipStream_t s) {
  const dim3 grid(p.gx);
  void (*kern)(KnnArgs<T>) = nullptr;
  const int ch = small_ch(a.d);
#define GSKNN_SMALL_CASE(P_, CH_) \
  if (p.small == P_ && ch == CH_) kern = knn_small_m_kernel<T, P_, CH_>;
  GSKNN_SMALL_CASE(8, 8)
  GSKNN_SMALL_CASE(8, 4)
  GSKNN_SMALL_CASE(8, 2)
  GSKNN_SMALL_CASE(16, 8)
  GSKNN_SMALL_CASE(16, 4)
  GSKNN_SMALL_CASE(16, 2)
  GSKNN_SMALL_CASE(32, 8)
  GSKNN_SMALL_CASE(32, 4)
  GSKNN_SMALL_CASE(32, 2)
#undef GSKNN_SMALL_CASE
  if (kern == nullptr) return hipErrorNotSupported;
  // [d][S] tile, C norms, 4 query rows
  const size_t lds = ((size_t)small_stride(ch) * a.d + 64 * ch + 256) * sizeof(T);
  hipError_t e = set_lds(kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_exact(const Plan &p, const KnnArgs<T> &a, hipStream_t s) {
  const dim3 grid(p.gx, p.nsplit);
  hipError_t e = hipSuccess;
#define GSKNN_EXACT_S_CASE(QPB_, D_)                                          \
  if (p.qpb == QPB_ && a.d == D_) {                                           \
    auto kern = knn_exact_s_kernel<T, QPB_, D_>;                              \
    if ((e = set_lds(kern, p.lds)) != hipSuccess) return e;                   \
    hipLaunchKernelGGL(kern, grid, dim3(QPB_), p.lds, s, a);                  \
    return hipGetLastError();                                                 \
  }
#define GSKNN_EXACT_CASE(QPB_)                                                \
  if (p.ct == 0) {                                                            \
    GSKNN_EXACT_S_CASE(QPB_, 8)                                               \
    GSKNN_EXACT_S_CASE(QPB_, 16)                                              \
    GSKNN_EXACT_S_CASE(QPB_, 32)                                              \
    GSKNN_EXACT_S_CASE(QPB_, 64)                                              \
  } else if (p.qpb == QPB_) {                                                 \
    auto kern = a.d <= kExactDQ ? knn_exact_kernel<T, QPB_, kExactDQ>         \
                                : knn_exact_kernel<T, QPB_, 0>;               \
    if ((e = set_lds(kern, p.lds)) != hipSuccess) return e;                   \
    hipLaunchKernelGGL(kern, grid, dim3(QPB_), p.lds, s, a, p.ct);            \
    return hipGetLastError();                                                 \
  }
  GSKNN_EXACT_CASE(64)
  GSKNN_EXACT_CASE(32)
  GSKNN_EXACT_CASE(16)
#undef GSKNN_EXACT_CASE
#undef GSKNN_EXACT_S_CASE
  return hipErrorNotSupported;
}

// Calls with fewer queries than this merge with one wave per query.
constexpr int kWaveMergeMaxN = 16384;

bool merge_query_major(int n, int k, int nsplit) {
  return n <= kWaveMergeMaxN && k * (1 + nsplit) <= 64 * 16;
}

template <typename T>
hipError_t launch_merge(const KnnArgs<T> &a, hipStream_t s) {
  const int E = a.k * (1 + a.nsplit);
  if (a.part_qmajor) {
    const dim3 grid((a.n + 3) / 4);
#define GSKNN_WMERGE_CASE(R_)                                                 \
    if (E <= 64 * R_) {                                                       \
      hipLaunchKernelGGL((knn_merge_wave_kernel<T, R_>), grid, dim3(256), 0, s, a); \
      return hipGetLastError();                                               \
    }
    GSKNN_WMERGE_CASE(2)
    GSKNN_WMERGE_CASE(4)
    GSKNN_WMERGE_CASE(8)
    GSKNN_WMERGE_CASE(16)
#undef GSKNN_WMERGE_CASE
    return hipErrorInvalidValue;  // query-major lists always fit (merge_query_major)
  }
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

// Two-stage workspace after the shifted norms: key heaps [n][k2] (keys, then
// ids), the fallback count and the fallback lists (XB rows, heap rows).
struct TwoLayout {
  size_t hk, hid, cnt, fb, fh, bytes;
  TwoLayout(size_t base, int n, int k2) {
    hk = base;
    hid = hk + al16(sizeof(double) * (size_t)n * k2);
    cnt = hid + al16(sizeof(int) * (size_t)n * k2);
    fb = cnt + 16;
    fh = fb + al16(sizeof(int) * (size_t)n);
    bytes = fh + al16(sizeof(int) * (size_t)n) - base;
  }
};

size_t part_bytes(const Plan &p, int n, int k, int elem) {
  if (p.nsplit <= 1) return 0;
  const size_t ents = (size_t)p.nsplit * k * n;
  return al16(ents * elem) + al16(ents * sizeof(int));
}

// Workspace of an FP64 MFMA launch: the shifted candidate norms (m doubles),
// then the split partial lists.
size_t a2_bytes(const Plan &p, int m) { return p.kind == 1 ? al16(sizeof(double) * (size_t)m) : 0; }

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
  a.part_qmajor = p.nsplit > 1 && merge_query_major(a.n, a.k, p.nsplit);
  const size_t a2b = a2_bytes(p, a.m);
  const size_t need = a2b + part_bytes(p, a.n, a.k, (int)sizeof(T));
  if (need > 0 && (work == nullptr || work_bytes < need)) return hipErrorInvalidValue;
  const bool two = p.k2 > 0 && work_bytes >= a2b + TwoLayout(a2b, a.n, p.k2).bytes;
  if (p.nsplit > 1) {
    const size_t ents = (size_t)p.nsplit * a.k * a.n;
    a.partD = (T *)((char *)work + a2b);
    a.partI = (int *)((char *)work + a2b + al16(ents * sizeof(T)));
  }
  hipError_t e;
  const int terms = p.ring ? ring_terms(a.d) : 4 * p.ks;
  int *fail_n = nullptr;
  if (two) {
    const TwoLayout L(a2b, a.n, p.k2);
    a.k2 = p.k2;
    a.hk = (T *)((char *)work + L.hk);
    a.hid = (int *)((char *)work + L.hid);
    fail_n = (int *)((char *)work + L.cnt);
  }
  if (a2b > 0) {
    if constexpr (sizeof(T) == 8) {
      a.SN = (const T *)work;
      const int blocks = std::min((a.m + 3) / 4, 2048);
      hipLaunchKernelGGL(cand_shift_kernel, dim3(blocks), dim3(256), 0, s, a.m, a.d, terms, a.XA,
                         a.XA2, a.amap, (double *)work, fail_n);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  if (p.kind == 1 && p.direct) {
    if constexpr (sizeof(T) == 8) {
      // 16-byte row loads (ring form) need d even and both tables 16-byte
      // aligned.
      const bool vec = a.d % 2 == 0 && ((uintptr_t)a.XA % 16) == 0 && ((uintptr_t)a.XB % 16) == 0;
      auto per_wave = [&](const KnnArgs<T> &x, bool two_stage) -> hipError_t {
        if (p.ring)
          return launch_knn_ring(x, vec, p.gx, two_stage ? p.lds2 : p.lds, s, two_stage);
        Plan q = p;
        q.lds = two_stage ? p.lds2 : p.lds;
        switch (p.ks) {
          case 4: return launch_mfma_direct_ks<4>(q, x, s, two_stage);
          case 8: return launch_mfma_direct_ks<8>(q, x, s, two_stage);
          case 16: return launch_mfma_direct_ks<16>(q, x, s, two_stage);
          case 32: return launch_mfma_direct_ks<32>(q, x, s, two_stage);
          default: return hipErrorNotSupported;
        }
      };
      e = per_wave(a, two);
      if (e == hipSuccess && two) {
        const TwoLayout L(a2b, a.n, p.k2);
        int *fail_b = (int *)((char *)work + L.fb);
        int *fail_h = (int *)((char *)work + L.fh);
        const int P = two_stage_sort_size(a.k, p.k2);
        const size_t lds = al16(sizeof(double) * 4 * (size_t)P) + sizeof(int) * 4 * (size_t)P;
        if ((e = set_lds(knn_two_stage_finalize_kernel, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(knn_two_stage_finalize_kernel, dim3((a.n + 3) / 4), dim3(256), lds, s, a,
                           terms, P, fail_n, fail_b, fail_h);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // Queries whose rows the finalize kernel could not settle: the
        // one-stage kernel over the fallback list (workgroups past the
        // list's length return at once).
        KnnArgs<T> f = a;
        f.bmap = fail_b;
        f.hrow = fail_h;
        f.ncount = fail_n;
        f.k2 = 0;
        f.hk = nullptr;
        f.hid = nullptr;
        e = per_wave(f, false);
      }
    } else {
      return hipErrorNotSupported;
    }
  } else if (p.kind == 1) {
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
    e = p.small ? launch_small_m(p, a, s) : launch_exact(p, a, s);
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
  const size_t two = p.k2 > 0 ? TwoLayout(0, n, p.k2).bytes : 0;
  return a2_bytes(p, m) + part_bytes(p, n, k, elem_bytes) + two;
}

size_t knn_two_stage_count_offset(int m, int n, int d, int k, int mode) {
  if (m <= 0 || n <= 0 || d <= 0 || k <= 0) return (size_t)-1;
  const Plan p = make_plan(m, n, d, k, 8, mode);
  if (p.k2 == 0) return (size_t)-1;
  const size_t a2b = a2_bytes(p, m);
  return TwoLayout(a2b, n, p.k2).cnt;
}

size_t knn_work_bytes_min(int m, int n, int d, int k, int elem_bytes, int mode) {
  if (m <= 0 || n <= 0 || d <= 0 || k <= 0) return 0;
  if (elem_bytes == 4 && use_f32_mfma(m, d, k, mode)) return f32_work_bytes(m, n, d, k);
  const Plan p = make_plan(m, n, d, k, elem_bytes, mode);
  return a2_bytes(p, m) + part_bytes(p, n, k, elem_bytes);
}

bool knn_merge_supported(int k, int elem_bytes) { return merge_qpb(k, elem_bytes) > 0; }

bool knn_f32_takes_mfma(int m, int d, int k, int mode) { return use_f32_mfma(m, d, k, mode); }

bool knn_merge_query_major(int n, int k, int nsplit) { return merge_query_major(n, k, nsplit); }

hipError_t launch_knn_merge(const KnnArgs<double> &a, hipStream_t s) { return launch_merge(a, s); }

hipError_t launch_knn_merge(const KnnArgs<float> &a, hipStream_t s) { return launch_merge(a, s); }

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

template <typename T>
hipError_t launch_rows_ascending_t(int n, int k, T *D, int *I, int ldk, hipStream_t s) {
  if (n <= 0 || k <= 0) return hipSuccess;
  if (k > kRowsAscendingMaxK || ldk < k) return hipErrorInvalidValue;
  int P = 1;
  while (P < k) P <<= 1;
  const int R = P >= 64 ? 1 : 64 / P;  // rows per workgroup
  const size_t lds = al16(sizeof(T) * (size_t)P * R) + sizeof(int) * (size_t)P * R;
  hipError_t e = hipFuncSetAttribute((const void *)rows_ascending_kernel<T>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rows_ascending_kernel<T>, dim3((n + R - 1) / R), dim3(64), lds, s, n, k, D, I,
                     ldk, P, R);
  return hipGetLastError();
}

hipError_t launch_rows_ascending_f64(int n, int k, double *D, int *I, int ldk, hipStream_t s) {
  return launch_rows_ascending_t<double>(n, k, D, I, ldk, s);
}

hipError_t launch_rows_ascending_f32(int n, int k, float *D, int *I, int ldk, hipStream_t s) {
  return launch_rows_ascending_t<float>(n, k, D, I, ldk, s);
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
