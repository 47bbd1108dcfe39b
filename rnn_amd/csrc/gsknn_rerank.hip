// BF16 candidate pass + FP64 re-rank: exact kNN from a BF16 MFMA contraction.
//
// No counterpart in the reference (whose kernels are FP64/FP32 only,
// kernels/x86_64/sandybridge/); the result is the exact k-set of the FP64
// distances, merged into dgsknn-style heap rows (frame/gsknn_heap.c:104-162
// semantics, canonical (distance, index) ties).
//
// Inputs are FP32 point tables.  Preparation (once per table) stores, per row:
//   xb   the row rounded to BF16 (round to nearest even), zero padded to DP,
//   n64  sum_p x_p^2 in FP64,
//   xn   |bf16(x)| and en = |x - bf16(x)|, both rounded up to FP32.
//
// Error bound.  For candidate a and query b with BF16 images a~, b~ and
// residuals e_a = a - a~, e_b = b - b~ (exact in FP32):
//   |a.b - a~.b~| <= |e_a||b~| + |a~||e_b| + |e_a||e_b|            (inputs)
//   BF16 products are exact in FP32; an FP32 sum of d terms in any order,
//   each addition rounding to within 2u (u = 2^-24), is off by at most
//   g |a~||b~| with g = (d+1) 2u / (1 - (d+1) 2u)                    (sum)
//   t = fma(-2, dot, nc) and t + nq round twice; the FP32 norms round once
//   each: at most 4u (nq + nc + 2 |a~||b~|)                           (epilogue)
// so with dist~ = (nq + nc) - 2 dot~ evaluated in FP32 and the true
// distance D = |a - b|^2 = nq + nc - 2 a.b,
//   |dist~ - D| <= delta = 2(en_a xn_b + xn_a en_b + en_a en_b)
//                          + (2g + 8u) xn_a xn_b + 4u (nq + nc)
// (times 1 + 2^-10 for the rounding of delta itself).  lower = dist~ - delta
// and upper = dist~ + delta bracket D.
//
// Pass 1 (knn_bf16_filter2_kernel for k <= 64, knn_bf16_filter_kernel
// otherwise): per query, U = k-th smallest of {upper bounds of all
// candidates} U {resident heap distances} (or a valid bound above it); U
// bounds the final k-th distance from above.  Every candidate whose lower
// bound is <= U at the time it is seen is appended to the query's list (id,
// lower, upper).
// Pass 2 (knn_rerank_select_kernel for k <= 256): one wave per query takes T,
// the k-th smallest of the listed upper bounds and the resident row, computes
// FP64 direct differences sum_p (a_p - b_p)^2 (wave_row_dists: a fixed
// summation order per pair) for the entries whose lower bound is <= T, and
// keeps the k smallest of (resident row + those) in canonical order.  A query
// whose list overflowed, and every query for larger k, goes to the heap form
// (knn_rerank_kernel), which re-scans an overflowed split in FP64.
#include <limits.h>

#include "gsknn_common.h"

namespace gsknn {

namespace {

__host__ __device__ constexpr size_t al16r(size_t x) { return (x + 15) & ~size_t(15); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kU32 = 5.9604644775390625e-08f;  // 2^-24

__device__ __forceinline__ bool cless_d(double da, int ia, double db, int ib) {
  return da < db || (da == db && ia < ib);
}

// x moved outward by 2^-22 |x|, so that a bound computed with round-to-nearest
// additions (each off by at most 2^-24 of its result) still holds.
__device__ __forceinline__ float widen_up(float x) { return __builtin_fmaf(0x1p-22f, fabsf(x), x); }
__device__ __forceinline__ float widen_down(float x) { return __builtin_fmaf(-0x1p-22f, fabsf(x), x); }

}  // namespace

// BF16 row length: d rounded up to 32 KS dims, KS = 1, 2, 4 or 8 (the
// k-steps the filter kernels are built for), zero padded.
__host__ __device__ constexpr int bf16_row_dims(int d) {
  return d <= 32 ? 32 : d <= 64 ? 64 : d <= 128 ? 128 : 256;
}

// Prepared-table layout: bf16 rows [N][DP] | n64 [N] | xn [N] | en [N] |
// maxima {xn, en, n32} as 3 floats.
struct PrepLayout {
  size_t xb, n64, xn, en, mx, bytes;
  int DP;
  __host__ __device__ PrepLayout(int N, int d) {
    DP = bf16_row_dims(d);
    xb = 0;
    n64 = al16r(xb + (size_t)N * DP * 2);
    xn = al16r(n64 + (size_t)N * 8);
    en = al16r(xn + (size_t)N * 4);
    mx = al16r(en + (size_t)N * 4);
    bytes = al16r(mx + 16);
  }
};

size_t bf16_prep_bytes(int N, int d) { return PrepLayout(N, d).bytes; }

__global__ void prep_zero_kernel(float *mx) {
  if (threadIdx.x < 4) mx[threadIdx.x] = 0.0f;
}

// Counters are cleared by a kernel, not hipMemsetAsync, so that a launch
// sequence captured into a HIP graph replays correctly.
__global__ void zero_int_kernel(int *p) {
  if (threadIdx.x == 0) *p = 0;
}

__global__ void prep_bf16_kernel(int N, int d, const float *X, unsigned char *prep) {
  const PrepLayout L(N, d);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  __bf16 *xb = (__bf16 *)(prep + L.xb) + (size_t)i * L.DP;
  const float *x = X + (size_t)i * d;
  double s = 0.0, sb = 0.0, se = 0.0;
  for (int p = 0; p < d; ++p) {
    const float v = x[p];
    const __bf16 hb = (__bf16)v;  // round to nearest even
    const float vb = (float)hb;
    const float e = v - vb;       // exact
    xb[p] = hb;
    s += (double)v * (double)v;
    sb += (double)vb * (double)vb;
    se += (double)e * (double)e;
  }
  for (int p = d; p < L.DP; ++p) xb[p] = (__bf16)0.0f;
  ((double *)(prep + L.n64))[i] = s;
  // sqrt in FP64 then round up to FP32; the 1 + 2^-40 covers the FP64 sum.
  const float xn = __double2float_ru(sqrt(sb) * (1.0 + 0x1p-40));
  const float en = __double2float_ru(sqrt(se) * (1.0 + 0x1p-40));
  const float n32 = __double2float_ru(s);
  ((float *)(prep + L.xn))[i] = xn;
  ((float *)(prep + L.en))[i] = en;
  float *mx = (float *)(prep + L.mx);
  // Non-negative floats order like their bit patterns.
  atomicMax((int *)&mx[0], __float_as_int(xn));
  atomicMax((int *)&mx[1], __float_as_int(en));
  atomicMax((int *)&mx[2], __float_as_int(n32));
}

hipError_t launch_bf16_prep(int N, int d, const float *X, void *prep, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  const PrepLayout L(N, d);
  hipLaunchKernelGGL(prep_zero_kernel, dim3(1), dim3(64), 0, s, (float *)((char *)prep + L.mx));
  hipLaunchKernelGGL(prep_bf16_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, d, X,
                     (unsigned char *)prep);
  return hipGetLastError();
}

struct Bf16Args {
  int m, n, d, k, ldk, cap;
  int nsplit, cps;             // candidate splits (grid.y) and candidates per split
  int NA, NB;                  // rows of the candidate / query tables
  const float *XA, *XB;        // FP32 tables (re-rank)
  const unsigned char *prepA, *prepB;
  const int *amap, *bmap;
  double *D;
  int *I;
  const int *hrow;
  const unsigned char *panel;  // register form: candidates gathered in position order
  int G;        // lists per (split, query): the cap entries split into G segments
  int2 *lst;    // [nsplit][n][cap] (candidate id, lower bound bits)
  float *lstu;  // [nsplit][n][cap] upper bound of the same entry
  int *cnt;     // [nsplit][G][n] entries appended to each list
  float *U;     // [nsplit][n]
  int *nover;   // lists that overflowed
  int warm;     // register form: tiles of each split in the warm-up pass
};

// List gg of split sp of query j: its count and its first entry.
__device__ __forceinline__ int list_count(const Bf16Args &a, int sp, int gg, int j) {
  return a.cnt[((size_t)sp * a.G + gg) * a.n + j];
}
__device__ __forceinline__ size_t list_base(const Bf16Args &a, int sp, int gg, int j) {
  return ((size_t)sp * a.n + j) * a.cap + (size_t)gg * (a.cap / a.G);
}

// Lane l of v_mfma_f32_16x16x32_bf16 holds 8 consecutive k of row l%16 of A
// (and of column l%16 of B) starting at k = 8*(l/16); result entry v of lane l
// is row kRowOf(l, v), column l%16.
__device__ __forceinline__ int out_row(int lane, int v) { return 4 * (lane >> 4) + v; }


struct FilterLayout {
  size_t tA, nc, xn, en, id, heap, cnt, bytes;
  __host__ __device__ FilterLayout(int KS, int CB, int Q, int k) {
    const size_t tile = (size_t)KS * CB * 64 * 8;  // bf16 elements
    const size_t C = 16 * (size_t)CB;
    tA = 0;
    nc = al16r(tA + tile * 2);
    xn = al16r(nc + C * 4);
    en = al16r(xn + C * 4);
    id = al16r(en + C * 4);
    heap = al16r(id + C * 4);
    cnt = al16r(heap + (size_t)4 * Q * ((k + 3) / 4) * 4);
    bytes = al16r(cnt + (size_t)Q * 4);
  }
};


// Sift x down from slot s of the max-heap of k floats at h[q], h[q + QS], ...
__device__ void sift_vals(float *h, int QS, int q, int k, float x, int s = 0) {
  while (true) {
    int c = 2 * s + 1;
    if (c >= k) break;
    float cv = h[c * QS + q];
    if (c + 1 < k) {
      const float rv = h[(c + 1) * QS + q];
      if (cv < rv) {
        ++c;
        cv = rv;
      }
    }
    if (!(x < cv)) break;
    h[s * QS + q] = cv;
    s = c;
  }
  h[s * QS + q] = x;
}
// Candidate tiles go global -> registers -> LDS (one buffer, two barriers per
// tile) in whole fragment blocks per wave; the amap entries of a tile are read
// one tile ahead.  Lanes hold their query fragments in registers for the
// whole launch.
//
// Hot path per pair: tt = fma(-2, dot~, nc) and a running minimum; a query
// block whose minimum reaches the threshold thr = U - nq + delta_max
// (delta_max bounds delta over all candidates) compares its pairs.
// Survivors get their exact lower/upper bounds and wait in a per-lane queue
// of QR entries; the queue is applied (list append, upper-bound heap update)
// when some lane's queue would overflow, and U is then refreshed.  A stale U
// only admits more candidates.
// Two waves per SIMD when the query fragments (4 * QB * KS registers) leave
// room for the rest.
constexpr int bf16_waves_per_eu(int KS, int QB) { return QB * KS <= 16 ? 2 : 1; }

template <int KS, int QB, int CB, int W, bool PL>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(bf16_waves_per_eu(KS, QB)))) void knn_bf16_filter_kernel(Bf16Args a) {
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB;
  constexpr int DP = 32 * KS;
  constexpr int TILE = KS * CB * 64 * 8;  // bf16 elements per tile
  constexpr int QR = 4;
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k;
  const FilterLayout L(KS, CB, Q, k);
  __bf16 *tA = (__bf16 *)(smem + L.tA);
  float *tnc = (float *)(smem + L.nc);
  float *txn = (float *)(smem + L.xn);
  float *ten = (float *)(smem + L.en);
  int *tcid = (int *)(smem + L.id);
  float *hv = (float *)(smem + L.heap);
  int *hc = (int *)(smem + L.cnt);

  const PrepLayout PA(a.NA, a.d), PB(a.NB, a.d);
  const __bf16 *xbA = (const __bf16 *)(a.prepA + PA.xb);
  const __bf16 *xbB = (const __bf16 *)(a.prepB + PB.xb);
  const double *n64A = (const double *)(a.prepA + PA.n64);
  const double *n64B = (const double *)(a.prepB + PB.n64);
  const float *xnA = (const float *)(a.prepA + PA.xn);
  const float *xnB = (const float *)(a.prepB + PB.xn);
  const float *enA = (const float *)(a.prepA + PA.en);
  const float *enB = (const float *)(a.prepB + PB.en);
  const float *mxA = (const float *)(a.prepA + PA.mx);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, grp = lane >> 4;
  const int qbase = blockIdx.x * Q;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const size_t sbase = (size_t)split * a.n;  // this split's rows of lst/cnt/U

  // Upper-bound heaps (U, the k-th smallest upper bound seen, bounds the
  // k-th distance from above), seeded with the resident row's distances
  // rounded up to FP32.
  //  * PL (per lane): lane group g (lanes 16 g .. 16 g + 15) keeps, for each
  //    of its queries, a max-heap of its own of kg = (k + 3 - g) / 4 upper
  //    bounds: the resident row's entries g, g + 4, ... and the candidates
  //    that group's lanes see.  The four heaps of a query hold k values in
  //    all, each at or below the largest of the four roots, so that maximum
  //    is a valid U (somewhat above the exact one).  Each lane updates only
  //    its own heaps, so every lane applies its survivors at once.
  //  * otherwise one heap of k per query, the exact U; the four lane groups
  //    holding a query's candidates take turns to update it.
  constexpr int NS = 4 * Q;  // per-lane heap slots: (lane, query block) of the workgroup
  const int kg = PL ? (k + 3 - grp) / 4 : k;
  auto hslot = [&](int qb) -> int { return (w * 64 + lane) * QB + qb; };
  if constexpr (PL) {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      const int j = qbase + w * QW + qb * 16 + (lane & 15);
      const int hs = hslot(qb);
      if (j < a.n) {
        const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
        for (int e = 0; e < kg; ++e) hv[e * NS + hs] = __double2float_ru(a.D[row + grp + 4 * e]);
        for (int e = kg / 2 - 1; e >= 0; --e) sift_vals(hv, NS, hs, kg, hv[e * NS + hs], e);
      }
    }
  } else if (tid < Q && qbase + tid < a.n) {
    const size_t row = (size_t)(a.hrow ? a.hrow[qbase + tid] : qbase + tid) * a.ldk;
    for (int e = 0; e < k; ++e) hv[e * Q + tid] = __double2float_ru(a.D[row + e]);
    for (int e = k / 2 - 1; e >= 0; --e) sift_vals(hv, Q, tid, k, hv[e * Q + tid], e);
  }
  if (tid < Q) hc[tid] = 0;

  const float u = kU32;
  const float g = (float)((a.d + 1) * 2.0 * 0x1p-24 / (1.0 - (a.d + 1) * 2.0 * 0x1p-24));
  const float cxx = 2.0f * g + 8.0f * u;
  bf16x8 bq[QB][KS];
  float nq[QB], al[QB], be[QB], cq[QB], dmx[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = w * QW + qb * 16 + (lane & 15);
    const int j = qbase + ql[qb];
    qv[qb] = j < a.n;
    const int row = a.bmap[qv[qb] ? j : a.n - 1];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const bf16x8 v = *(const bf16x8 *)(xbB + (size_t)row * DP + s * 32 + 8 * grp);
      bq[qb][s] = qv[qb] ? v : bf16x8{};
    }
    const float xq = xnB[row];
    const float eq = enB[row];
    nq[qb] = (float)n64B[row];
    al[qb] = 2.0f * eq + cxx * xq;
    be[qb] = 2.0f * xq + 2.0f * eq;
    cq[qb] = 4.0f * u * nq[qb];
    dmx[qb] = (mxA[0] * al[qb] + mxA[1] * be[qb] + 4.0f * u * mxA[2] + cq[qb]) * 1.002f + 1e-30f;
  }

  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  // Staging: the tile is BLK fragment blocks of 64 x 8 bf16 (k-step s,
  // candidate block cb); wave w stages blocks w * BPW .. w * BPW + BPW - 1 and
  // lane l writes 16-byte element l of each (dims 32 s + 8 (l / 16) .. + 8 of
  // candidate cb * 16 + l % 16): every LDS store is 1 KB contiguous (no bank
  // conflicts).  Lanes 0..15 of wave w own the norm slots of candidate blocks
  // w, w + W, ... (one block per wave when W >= CB).
  constexpr int BLK = KS * CB;
  static_assert(BLK % W == 0, "fragment blocks per wave");
  constexpr int BPW = BLK / W;
  const int si = lane & 15, skk = lane >> 4;
  constexpr int NPW = (CB + W - 1) / W;  // norm blocks per wave
  uint4 stage[BPW];
  float st_n[NPW], st_x[NPW], st_e[NPW];
  int st_i[NPW];
#pragma unroll
  for (int r = 0; r < NPW; ++r) {
    st_n[r] = st_x[r] = st_e[r] = 0.f;
    st_i[r] = -1;
  }
  // Ids of this thread's candidates: of the tile being staged (sid) and of
  // the tile after it (nid, requested a tile ahead).
  typedef int idvec __attribute__((ext_vector_type(CB)));  // keeps selections in registers
  idvec sid, nid;
  auto cand_pos = [&](int t, int cb) { return min(cbeg + t * C + cb * 16 + si, cend - 1); };
  // Id of the candidate staged as element e (block w * BPW + e); constant
  // register index when BPW is a multiple of CB.
  auto stage_id = [&](int e) __attribute__((always_inline)) -> int {
    if (BPW % CB == 0) return sid[e % CB];
    const int cb = (w * BPW + e) % CB;
    int r = sid[0];
#pragma unroll
    for (int c = 1; c < CB; ++c)
      if (cb == c) r = sid[c];
    return r;
  };
  auto load_ids = [&](int t, idvec &ids) __attribute__((always_inline)) {
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) ids[cb] = a.amap[cand_pos(t, cb)];
  };
  auto load_x = [&]() {
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      const int ch = 4 * (bidx / CB) + skk;  // 16-byte chunk of the padded row
      stage[e] = *(const uint4 *)(xbA + (size_t)stage_id(e) * DP + 8 * ch);
    }
    if (skk == 0) {
#pragma unroll
      for (int r = 0; r < NPW; ++r) {
        const int nb = w + r * W;
        if (nb < CB) {
          int id = sid[0];
#pragma unroll
          for (int cb = 1; cb < CB; ++cb)
            if (nb == cb) id = sid[cb];
          st_n[r] = (float)n64A[id];
          st_i[r] = id;
          st_x[r] = xnA[id];
          st_e[r] = enA[id];
        }
      }
    }
  };
  auto store_x = [&](int t) {
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      const bool ok = cbeg + t * C + (bidx % CB) * 16 + si < cend;
      *(uint4 *)(tA + ((size_t)bidx * 64 + lane) * 8) = ok ? stage[e] : make_uint4(0, 0, 0, 0);
    }
    if (skk == 0) {
#pragma unroll
      for (int r = 0; r < NPW; ++r) {
        const int nb = w + r * W;
        if (nb < CB) {
          const int slot = nb * 16 + si;
          const bool ok = cbeg + t * C + slot < cend;
          tnc[slot] = ok ? st_n[r] : __builtin_inff();
          tcid[slot] = st_i[r];
          txn[slot] = ok ? st_x[r] : 0.0f;
          ten[slot] = ok ? st_e[r] : 0.0f;
        }
      }
    }
  };

  if (ntiles > 0) {
    load_ids(0, sid);
    load_ids(ntiles > 1 ? 1 : 0, nid);
    load_x();
    __syncthreads();  // heaps loaded
    store_x(0);
  }
  __syncthreads();
  // U of each query of this lane (every lane of the wave calls it).
  auto group_u = [&](int qb) -> float {
    if constexpr (!PL) return hv[ql[qb]];
    float r = kg > 0 ? hv[hslot(qb)] : -__builtin_inff();
    r = fmaxf(r, __shfl_xor(r, 16));
    r = fmaxf(r, __shfl_xor(r, 32));
    return r;
  };
  float uq[QB], thr[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    uq[qb] = group_u(qb);
    thr[qb] = qv[qb] ? f32_pass_thr(uq[qb], nq[qb], dmx[qb]) : -__builtin_inff();
  }
  auto refresh = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      uq[qb] = group_u(qb);
      if (qv[qb]) thr[qb] = f32_pass_thr(uq[qb], nq[qb], dmx[qb]);
    }
  };

  // Per-lane queue: lower, upper, candidate id, query block.
  float pq_lo[QR], pq_up[QR];
  int pq_ps[QR], pq_q[QR];
  int pq_n = 0;
#pragma unroll
  for (int r = 0; r < QR; ++r) {
    pq_lo[r] = pq_up[r] = 0.f;
    pq_ps[r] = pq_q[r] = 0;
  }
  // Apply one candidate of query block qb: append it if it can still reach
  // the k-th distance, tighten the heap with its upper bound.  PL: against U
  // as of the last refresh (a stale U only admits more), appended through
  // an LDS counter, this lane's heap; otherwise against the live root of the
  // query's heap (the caller's lane group holds it alone).
  auto apply = [&](int qb, float lower, float upper, int ps) __attribute__((always_inline)) {
    const float Uq = PL ? uq[qb] : hv[ql[qb]];
    if (lower <= Uq) {
      const int at = PL ? atomicAdd(&hc[ql[qb]], 1) : hc[ql[qb]]++;
      if (at < a.cap) {
        const size_t li = (sbase + qbase + ql[qb]) * a.cap + at;
        a.lst[li] = make_int2(ps, __float_as_int(lower));
        a.lstu[li] = upper;
      }
    }
    if constexpr (PL) {
      const int hs = hslot(qb);
      if (kg > 0 && upper < hv[hs]) sift_vals(hv, NS, hs, kg, upper);
    } else {
      if (upper < Uq) sift_vals(hv, Q, ql[qb], k, upper);
    }
  };
  // Lane groups one after another (shared heaps), or all at once (PL).
  auto by_groups = [&](auto &&body) __attribute__((always_inline)) {
    if constexpr (PL) {
      body();
    } else {
#pragma unroll 1
      for (int gsel = 0; gsel < 4; ++gsel) {
        if (grp == gsel) body();
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      }
    }
  };
  auto flush = [&]() {
    by_groups([&]() {
      for (int r = 0; r < QR; ++r) {
        if (r < pq_n) {
#pragma unroll
          for (int qb = 0; qb < QB; ++qb)
            if (pq_q[r] == qb) apply(qb, pq_lo[r], pq_up[r], pq_ps[r]);
        }
      }
    });
    pq_n = 0;
    refresh();
  };

  for (int t = 0; t < ntiles; ++t) {
    const bool more = t + 1 < ntiles;
    sid = nid;
    if (t + 2 < ntiles) load_ids(t + 2, nid);
    if (more) load_x();
    f32x4 acc[QB][CB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const bf16x8 av = *(const bf16x8 *)(tA + ((size_t)(s * CB + cb) * 64 + lane) * 8);
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bq[qb][s], acc[qb][cb], 0, 0, 0);
      }
    }
    // Hot path: one FMA and one min per pair; a query block's pass bits are
    // built only when its smallest t reaches the threshold.
    float tmin[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      tmin[qb] = __builtin_inff();
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = cb * 16 + out_row(lane, v);
          const float tt = __builtin_fmaf(-2.0f, acc[qb][cb][v], tnc[c]);
          acc[qb][cb][v] = tt;
          tmin[qb] = fminf(tmin[qb], tt);
        }
    }
    bool hit = false;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) hit |= tmin[qb] <= thr[qb];
    if (__any(hit)) {
      unsigned pass[QB];
      int npass = 0;
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        pass[qb] = 0u;
        if (tmin[qb] <= thr[qb]) {
#pragma unroll
          for (int b = 0; b < CB * 4; ++b)
            pass[qb] |= (acc[qb][b >> 2][b & 3] <= thr[qb] ? 1u : 0u) << b;
        }
        npass += __builtin_popcount(pass[qb]);
      }
      if (__any(pq_n + npass > QR)) flush();
      // After a flush the queues are empty; if some lane still cannot queue
      // all of its survivors, every lane applies this tile's survivors now.
      const bool direct = pq_n + npass > QR;
      if (__any(direct)) {
        by_groups([&]() {
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) {
            unsigned bits = pass[qb];
            while (bits) {
              const int b = __builtin_ctz(bits);
              bits &= bits - 1;
              float tt = acc[qb][0][0];
#pragma unroll
              for (int bb = 1; bb < CB * 4; ++bb)
                if (b == bb) tt = acc[qb][bb >> 2][bb & 3];
              const int c = (b >> 2) * 16 + out_row(lane, b & 3);
              if (cbeg + t * C + c >= cend) continue;
              const float dist = tt + nq[qb];
              const float delta = (txn[c] * al[qb] + ten[c] * be[qb] + 4.0f * u * tnc[c] + cq[qb]) *
                                      1.0009765625f + 1e-30f;
              apply(qb, widen_down(dist - delta), widen_up(dist + delta), tcid[c]);
            }
          }
        });
        refresh();
      } else {
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) {
          unsigned bits = pass[qb];
          while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            float tt = acc[qb][0][0];
#pragma unroll
            for (int bb = 1; bb < CB * 4; ++bb)
              if (b == bb) tt = acc[qb][bb >> 2][bb & 3];
            const int c = (b >> 2) * 16 + out_row(lane, b & 3);
            if (cbeg + t * C + c >= cend) continue;
            const float dist = tt + nq[qb];
            const float delta = (txn[c] * al[qb] + ten[c] * be[qb] + 4.0f * u * tnc[c] + cq[qb]) *
                                    1.0009765625f + 1e-30f;
#pragma unroll
            for (int r = 0; r < QR; ++r)
              if (pq_n == r) {
                pq_lo[r] = widen_down(dist - delta);
                pq_up[r] = widen_up(dist + delta);
                pq_ps[r] = tcid[c];
                pq_q[r] = qb;
              }
            ++pq_n;
          }
        }
      }
    }
    __syncthreads();  // tile consumed
    if (more) store_x(t + 1);
    __syncthreads();  // next tile visible
  }
  if (__any(pq_n > 0)) flush();
  float uf[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) uf[qb] = group_u(qb);
  __syncthreads();  // every append counted

  if (grp == 0) {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      const int j = qbase + ql[qb];
      if (j < a.n) {
        a.U[sbase + j] = uf[qb];
        a.cnt[sbase + j] = hc[ql[qb]];  // G = 1: [split][n]
        if (hc[ql[qb]] > a.cap) atomicAdd(a.nover, 1);
      }
    }
  }
}

// Register form of pass 1 (k from 48 to 4 RK): the same contraction, bounds
// and lists, with
//  * upper-bound sets in registers: lane group g keeps, for each of its
//    queries, its kg = (k + 3 - g) / 4 smallest upper bounds (the resident
//    row's entries g, g + 4, ... and the candidates its lanes see) as a
//    sorted array of RK floats, padded below with -inf; an insertion is a
//    branch-free min/max shift (a value at or above the largest changes
//    nothing), and U = the largest of the four groups' largest (the four
//    sets hold k values, all at or below it);
//  * one list segment per (query, lane group): the group appends to its own
//    segment of cap / 4 entries with a counter in a register (no atomics);
//  * the candidates gathered beforehand into a panel in candidate order
//    (bf16_gather_kernel: BF16 rows and a record {nc, xn, en, id} per
//    position), so every load of the tile loop has an affine address;
//  * candidate tiles loaded straight into LDS (global_load_lds_dwordx4,
//    lane l's 16 bytes at the block's base + 16 l, the fragment layout the
//    MFMA reads), two tile buffers, one barrier per tile: tile t + 1 loads
//    while tile t is computed;
//  * survivors applied in the tile they pass; lane l of every wave holds the
//    bound terms and id of the tile's candidate l in registers and survivors
//    fetch them by lane shuffles (no LDS reads after the next tile's loads
//    are issued, which the compiler would make wait for those loads).
// Candidate panel of the register form: BF16 rows [m + kPanelPad][DP] in
// candidate position order, then a record {nc, xn, en, id} per position
// (nc = the FP64 squared norm rounded to FP32, id = amap[position]).  The
// kPanelPad zero rows and records after the last position let a tile that
// runs past the end load without clamping (its extra candidates are masked).
constexpr int kPanelPad = 64;
struct PanelLayout {
  size_t rows, rec, bytes;
  __host__ __device__ PanelLayout(int m, int DP) {
    rows = 0;
    rec = al16r((size_t)(m + kPanelPad) * DP * 2);
    bytes = al16r(rec + (size_t)(m + kPanelPad) * 16);
  }
};

// One wave per candidate position (the pad positions written as zeros):
// 16 bytes of the row per lane.
__global__ __launch_bounds__(256) void bf16_gather_kernel(int m, int NA, int d, const int *amap,
                                                          const unsigned char *prepA,
                                                          unsigned char *panel) {
  const PrepLayout PA(NA, d);
  const int DP = PA.DP;
  const PanelLayout P(m, DP);
  const int pos = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pos >= m + kPanelPad) return;
  if (pos >= m) {
    uint4 *dst = (uint4 *)(panel + P.rows + (size_t)pos * DP * 2);
    for (int c = lane; c < DP / 8; c += 64) dst[c] = make_uint4(0u, 0u, 0u, 0u);
    if (lane == 0) ((float4 *)(panel + P.rec))[pos] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const int id = amap[pos];
  const uint4 *src = (const uint4 *)(prepA + PA.xb + (size_t)id * DP * 2);
  uint4 *dst = (uint4 *)(panel + P.rows + (size_t)pos * DP * 2);
  for (int c = lane; c < DP / 8; c += 64) dst[c] = src[c];
  if (lane == 0) {
    const float nc = (float)((const double *)(prepA + PA.n64))[id];
    const float xn = ((const float *)(prepA + PA.xn))[id];
    const float en = ((const float *)(prepA + PA.en))[id];
    ((float4 *)(panel + P.rec))[pos] = make_float4(nc, xn, en, __int_as_float(id));
  }
}

// Per buffer: the tile and the candidates' FP32 norms (read by every pair).
struct Filter2Layout {
  size_t tile, nc, buf, bytes;
  __host__ __device__ Filter2Layout(int KS, int CB) {
    const size_t C = 16 * (size_t)CB;
    tile = 0;
    nc = al16r(tile + (size_t)KS * CB * 64 * 16);
    buf = al16r(nc + C * 4);  // one buffer; two are used
    bytes = 2 * buf;
  }
};

template <int KS, int QB, int CB, int W, int RK>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(2))) void knn_bf16_filter2_kernel(Bf16Args a) {
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB;
  constexpr int DP = 32 * KS;
  constexpr int BLK = KS * CB;  // 64 x 8 bf16 fragment blocks per tile
  static_assert(BLK % W == 0, "fragment blocks per wave");
  static_assert(C == 64, "one candidate per lane");
  constexpr int BPW = BLK / W;
  extern __shared__ __align__(16) unsigned char smem[];
  const Filter2Layout L(KS, CB);
  const int k = a.k;

  const PrepLayout PA(a.NA, a.d), PB(a.NB, a.d);
  const __bf16 *xbB = (const __bf16 *)(a.prepB + PB.xb);
  const double *n64B = (const double *)(a.prepB + PB.n64);
  const float *xnB = (const float *)(a.prepB + PB.xn);
  const float *enB = (const float *)(a.prepB + PB.en);
  const float *mxA = (const float *)(a.prepA + PA.mx);
  const PanelLayout PL(a.m, DP);
  const __bf16 *xg = (const __bf16 *)(a.panel + PL.rows);
  const float4 *grec = (const float4 *)(a.panel + PL.rec);

  const int tid = threadIdx.x, lane = tid & 63, grp = lane >> 4, si = lane & 15;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int qbase = blockIdx.x * Q;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const int capg = a.cap / 4;

  const float u = kU32;
  const float g = (float)((a.d + 1) * 2.0 * 0x1p-24 / (1.0 - (a.d + 1) * 2.0 * 0x1p-24));
  const float cxx = 2.0f * g + 8.0f * u;
  bf16x8 bq[QB][KS];
  float nq[QB], al[QB], be[QB], cq[QB], dmx[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = w * QW + qb * 16 + si;
    const int j = qbase + ql[qb];
    qv[qb] = j < a.n;
    const int row = a.bmap[qv[qb] ? j : a.n - 1];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const bf16x8 v = *(const bf16x8 *)(xbB + (size_t)row * DP + s * 32 + 8 * grp);
      bq[qb][s] = qv[qb] ? v : bf16x8{};
    }
    const float xq = xnB[row];
    const float eq = enB[row];
    nq[qb] = (float)n64B[row];
    al[qb] = 2.0f * eq + cxx * xq;
    be[qb] = 2.0f * xq + 2.0f * eq;
    cq[qb] = 4.0f * u * nq[qb];
    // At least every candidate's delta.
    dmx[qb] = (mxA[0] * al[qb] + mxA[1] * be[qb] + 4.0f * u * mxA[2] + cq[qb]) * 1.002f + 1e-30f;
  }

  // Upper-bound sets, seeded with the resident row rounded up to FP32.  The
  // sets hold FP32 bit patterns ordered as integers: upper bounds of
  // distances are not negative (a negative one sorts below every
  // non-negative value, which keeps the largest a valid U), -inf pads below
  // and +inf marks an empty place.
  const int kg = (k + 3 - grp) / 4;
  int hu[QB][RK];
  auto insert_bits = [&](int qb, int x) __attribute__((always_inline)) {
#pragma unroll
    for (int i = RK - 1; i >= 1; --i) hu[qb][i] = max(hu[qb][i - 1], min(hu[qb][i], x));
    hu[qb][0] = min(hu[qb][0], x);
  };
  auto insert = [&](int qb, float xf) __attribute__((always_inline)) {
    insert_bits(qb, __float_as_int(xf));
  };
  constexpr int kNegInf = (int)0xFF800000u, kPosInf = 0x7F800000;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
#pragma unroll
    for (int i = 0; i < RK; ++i) hu[qb][i] = i < RK - kg ? kNegInf : kPosInf;
    if (qv[qb]) {
      const int j = qbase + ql[qb];
      const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
      for (int e = 0; e < kg; ++e) insert(qb, __double2float_ru(a.D[row + grp + 4 * e]));
    }
  }
  auto group_u = [&](int qb) -> float {
    float r = __int_as_float(hu[qb][RK - 1]);
    r = fmaxf(r, __shfl_xor(r, 16));
    r = fmaxf(r, __shfl_xor(r, 32));
    return r;
  };
  float uq[QB], thr[QB];
  auto refresh = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      uq[qb] = group_u(qb);
      thr[qb] = qv[qb] ? f32_pass_thr(uq[qb], nq[qb], dmx[qb]) : -__builtin_inff();
    }
  };
  refresh();
  int lc[QB];  // entries appended to this group's segment of each query's list
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) lc[qb] = 0;

  // Tile staging.  Wave w loads fragment blocks w BPW .. w BPW + BPW - 1 of
  // each tile (k-step bidx / CB, candidate block bidx % CB): lane l reads
  // dims 32 s + 8 (l / 16) .. + 8 of candidate (bidx % CB) 16 + l % 16.
  // Lane l of every wave loads the bound terms of the tile's candidate l
  // (block l / 16); wave w writes the FP32 norms of block w to LDS.
  // Tile t covers positions cbeg + t C .. + C - 1 (the panel's pad rows
  // cover a last tile that runs past the end): a uniform base per tile and
  // fixed per-lane offsets.
  auto tile_rows = [&](int t) { return (const char *)xg + (size_t)(cbeg + t * C) * DP * 2; };
  auto tile_recs = [&](int t) { return grec + cbeg + t * C; };
  unsigned doff[BPW];  // byte offset of this lane's 16 bytes of block w BPW + e
#pragma unroll
  for (int e = 0; e < BPW; ++e) {
    const int bidx = w * BPW + e;
    doff[e] = (unsigned)(((bidx % CB) * 16 + si) * DP * 2 + ((bidx / CB) * 32 + 8 * grp) * 2);
  }
  // The launch walks a step sequence: steps 0 .. T0 - 1 are the warm-up pass
  // over tiles 0 .. T0 - 1, steps T0 .. T0 + ntiles - 1 the main pass over
  // every tile.  Step v uses LDS buffer v & 1.
  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  const int T0 = min(a.warm, ntiles);
  const int nsteps = ntiles > 0 ? T0 + ntiles : 0;
  auto tile_of = [&](int v) { return v < T0 ? v : v - T0; };
  auto bufp = [&](int v) -> unsigned char * { return smem + (size_t)(v & 1) * L.buf; };
  // Record of candidate `lane` of the next tile, held as loaded until the
  // tile's loads have landed, and the bound terms of the current tile.
  float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
  float rc_x = 0.f, rc_e = 0.f, rc_n = 0.f;
  int rc_i = 0;
  auto issue_tile = [&](int v) __attribute__((always_inline)) {
    unsigned char *b = bufp(v);
    const int t = tile_of(v);
    const char *rows = tile_rows(t);
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      __builtin_amdgcn_global_load_lds((const void *)(rows + doff[e]),
                                       (__attribute__((address_space(3))) void *)(b + L.tile + (size_t)bidx * 1024),
                                       16, 0, 0);
    }
    st = tile_recs(t)[lane];
  };
  // After the wave's loads of tile t have landed: the bound terms become
  // current, and wave w writes the norms of block w.
  auto finish_tile = [&](int v) __attribute__((always_inline)) {
    unsigned char *b = bufp(v);
    const bool ok = cbeg + tile_of(v) * C + lane < cend;
    rc_n = ok ? st.x : __builtin_inff();
    rc_x = ok ? st.y : 0.0f;
    rc_e = ok ? st.z : 0.0f;
    rc_i = __float_as_int(st.w);
    if (grp == w) ((float *)(b + L.nc))[lane] = rc_n;
  };
  // vmcnt(0), other counters untouched: every load of this wave (the tile's
  // LDS loads included) has completed.
  auto wait_loads = []() __attribute__((always_inline)) { __builtin_amdgcn_s_waitcnt(0x0F70); };

  if (nsteps > 0) {
    issue_tile(0);
    wait_loads();
    finish_tile(0);
  }

  for (int step = 0; step < nsteps; ++step) {
    __syncthreads();  // this step's tile in LDS; every wave done with the last step
    const int t = tile_of(step);
    const bool more = step + 1 < nsteps;
    if (more) issue_tile(step + 1);
    const unsigned char *b = bufp(step);
    const __bf16 *tA = (const __bf16 *)(b + L.tile);
    const float *tnc = (const float *)(b + L.nc);
    f32x4 acc[QB][CB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const bf16x8 av = *(const bf16x8 *)(tA + ((size_t)(s * CB + cb) * 64 + lane) * 8);
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bq[qb][s], acc[qb][cb], 0, 0, 0);
      }
    }
    // Hot path: one FMA and one min per pair (the values as scalars, slot
    // cb 4 + v).
    float tq[QB][CB * 4];
    float tmin[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      tmin[qb] = __builtin_inff();
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = cb * 16 + out_row(lane, v);
          const float tt = __builtin_fmaf(-2.0f, acc[qb][cb][v], tnc[c]);
          tq[qb][cb * 4 + v] = tt;
          tmin[qb] = fminf(tmin[qb], tt);
        }
    }
    if (step < T0) {
      // Warm-up step: each lane's smallest tt of the tile (its 16 slots, a
      // chunk of candidates no other lane or step sees) gives an upper bound
      // of that one candidate's distance, (tmin + nq) + delta_max rounded
      // up; the sets take these bounds.  Every valid lane inserts at once:
      // no survivor loop, no list entries.
      int xb[QB];
      bool need = false;
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        const float s = (tmin[qb] + nq[qb]) + dmx[qb];
        xb[qb] = qv[qb] ? __float_as_int(widen_up(s)) : kPosInf;
        need |= xb[qb] < hu[qb][RK - 1];
      }
      if (__any(need)) {
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) insert_bits(qb, xb[qb]);
        refresh();
      }
    }
    // Main pass.  Over tiles the warm-up pass covered, survivors are listed
    // but their upper bounds stay out of the sets (each set entry stands for
    // a distinct candidate).
    const bool ins = step >= 2 * T0;
    bool hit = false;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) hit |= step >= T0 && tmin[qb] <= thr[qb];
    if (__any(hit)) {
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        unsigned bits = 0u;
        if (tmin[qb] <= thr[qb]) {
#pragma unroll
          for (int bb = 0; bb < CB * 4; ++bb) bits |= (tq[qb][bb] <= thr[qb] ? 1u : 0u) << bb;
        }
        const int j = qbase + ql[qb];
        const size_t lb = ((size_t)split * a.n + j) * a.cap + (size_t)grp * capg;
        // Wave-uniform loop: the record shuffles read other lanes, which
        // must be active (a shuffle from an inactive lane returns 0).  The
        // body is straight-line (a lane without a survivor inserts +inf, which
        // leaves its set unchanged), so the sets update in place.
        if (__any(bits != 0u)) do {
          const bool act = bits != 0u;
          const int bb = act ? __builtin_ctz(bits) : 0;
          bits &= bits - 1;
          const float tt = select_tree<0, CB * 4, CB * 4>(tq[qb], bb);
          const int c = (bb >> 2) * 16 + out_row(lane, bb & 3);
          // This tile's bound terms (finish_tile of the next tile, after
          // the survivors, replaces them).
          const float rx = __shfl(rc_x, c), re = __shfl(rc_e, c), rn = __shfl(rc_n, c);
          const int rid = __shfl(rc_i, c);
          const bool valid = act && cbeg + t * C + c < cend;
          const float dist = tt + nq[qb];
          const float delta = (rx * al[qb] + re * be[qb] + 4.0f * u * rn + cq[qb]) *
                                  1.0009765625f + 1e-30f;
          const float lower = widen_down(dist - delta), upper = widen_up(dist + delta);
          const bool app = valid && lower <= uq[qb];
          const int at = lc[qb];
          lc[qb] = at + (app ? 1 : 0);
          if (app && at < capg) {
            a.lst[lb + at] = make_int2(rid, __float_as_int(lower));
            a.lstu[lb + at] = upper;
          }
          // max with +inf's pattern for a lane without a valid survivor (an
          // integer max of constants, which the compiler keeps branch-free).
          const int ux = max(__float_as_int(upper), valid && ins ? (int)0x80000000 : kPosInf);
          if (__any(ux < hu[qb][RK - 1])) insert_bits(qb, ux);
        } while (__any(bits != 0u));
      }
      refresh();
    }
    if (more) {
      wait_loads();
      finish_tile(step + 1);
    }
  }

  float uf[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) uf[qb] = group_u(qb);
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    const int j = qbase + ql[qb];
    if (j < a.n) {
      a.cnt[((size_t)split * 4 + grp) * a.n + j] = lc[qb];
      if (grp == 0) a.U[(size_t)split * a.n + j] = uf[qb];
      if (lc[qb] > capg) atomicAdd(a.nover, 1);
    }
  }
}

// Pass 2: one wave per query.  LDS heap per wave: k doubles + k ints.
__device__ void heap_sift_d(double *hd, int *hi, int s, int k, double xd, int xi) {
  while (true) {
    int c = 2 * s + 1;
    if (c >= k) break;
    double cd = hd[c];
    int ci = hi[c];
    if (c + 1 < k && cless_d(cd, ci, hd[c + 1], hi[c + 1])) {
      ++c;
      cd = hd[c];
      ci = hi[c];
    }
    if (!cless_d(xd, xi, cd, ci)) break;
    hd[s] = cd;
    hi[s] = ci;
    s = c;
  }
  hd[s] = xd;
  hi[s] = xi;
}

// FP64 direct differences sum_p (a_p - b_p)^2 of up to kRowBatch candidate
// rows for one query, computed by the whole wave: lane l holds query dims
// 4 l .. 4 l + 3 (qv, zero beyond d) and reads the same dims of each row with
// one 16-byte load (scalar loads when d is not a multiple of 4), sums its four
// terms in order, and a fixed butterfly over the 64 lanes adds the partial
// sums (the same tree for every row).  A pair's distance therefore does not
// depend on the list, the batch or the call it comes in.  Row ids are in
// lanes 0 .. nb - 1 of myid; the distance of row b is returned in lanes
// kRowLanes b .. kRowLanes (b + 1) - 1.  d <= 256.
constexpr int kRowBatch = 4;
constexpr int kRowLanes = 64 / kRowBatch;  // lanes holding each row's distance

__device__ double wave_row_dists(const float *XA, int d, int myid, int nb, const double qv[4]) {
  const int lane = threadIdx.x & 63;
  const bool vec4 = (d & 3) == 0;  // rows 16-byte aligned
  double v[kRowBatch];
#pragma unroll
  for (int b = 0; b < kRowBatch; ++b) {
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    const int id = __builtin_amdgcn_readlane(myid, b);
    const float *xa = XA + (size_t)id * d + 4 * lane;
    if (b < nb) {
      if (vec4) {
        if (4 * lane < d) x = *(const float4 *)xa;
      } else {
        if (4 * lane < d) x.x = xa[0];
        if (4 * lane + 1 < d) x.y = xa[1];
        if (4 * lane + 2 < d) x.z = xa[2];
        if (4 * lane + 3 < d) x.w = xa[3];
      }
    }
    double df = (double)x.x - qv[0];
    double sacc = df * df;
    df = (double)x.y - qv[1];
    sacc = __builtin_fma(df, df, sacc);
    df = (double)x.z - qv[2];
    sacc = __builtin_fma(df, df, sacc);
    df = (double)x.w - qv[3];
    v[b] = __builtin_fma(df, df, sacc);
  }
  // Lanes differing in bit 5, then 4 (then 3, for 8 rows) exchange half of
  // their values; the remaining lane bits add.  Row b ends in the lanes
  // whose top log2(kRowBatch) bits are b.
  const bool b5 = lane & 32, b4 = lane & 16;
  double w1;
  if constexpr (kRowBatch == 8) {
    const bool b3 = lane & 8;
    double w4[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double send = b5 ? v[c] : v[c + 4];
      w4[c] = (b5 ? v[c + 4] : v[c]) + __shfl_xor(send, 32);
    }
    double w2[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double send = b4 ? w4[c] : w4[c + 2];
      w2[c] = (b4 ? w4[c + 2] : w4[c]) + __shfl_xor(send, 16);
    }
    w1 = (b3 ? w2[1] : w2[0]) + __shfl_xor(b3 ? w2[0] : w2[1], 8);
  } else {
    static_assert(kRowBatch == 4, "4 or 8 rows per batch");
    double w2[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double send = b5 ? v[c] : v[c + 2];
      w2[c] = (b5 ? v[c + 2] : v[c]) + __shfl_xor(send, 32);
    }
    w1 = (b4 ? w2[1] : w2[0]) + __shfl_xor(b4 ? w2[0] : w2[1], 16);
    w1 += __shfl_xor(w1, 8);
  }
  w1 += __shfl_xor(w1, 4);
  w1 += __shfl_xor(w1, 2);
  w1 += __shfl_xor(w1, 1);
  return w1;
}

// This lane's query dims 4 lane .. 4 lane + 3 for wave_row_dists.
__device__ __forceinline__ void load_query4(const float *xq, int d, double qv[4]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = 4 * lane + i;
    qv[i] = p < d ? (double)xq[p] : 0.0;
  }
}

// Heap form of pass 2 for query j: the row as a max-heap in LDS (hd, hi: k
// entries), list entries in list order, dropped when their lower bound
// exceeds the filter's U or the row's current k-th distance; an overflowed
// list re-scans every candidate of its split.  Used for overflowed lists and
// for k above the selection form's limit.
__device__ void rerank_heap_path(const Bf16Args &a, int j, double *hd, int *hi) {
  const int k = a.k, d = a.d;
  const int lane = threadIdx.x & 63;
  const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
  for (int t = lane; t < k; t += 64) {
    hd[t] = a.D[row + t];
    hi[t] = a.I[row + t];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane == 0)
    for (int s = k / 2 - 1; s >= 0; --s) heap_sift_d(hd, hi, s, k, hd[s], hi[s]);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  double qv[4];
  load_query4(a.XB + (size_t)a.bmap[j] * d, d, qv);
  // Every split's U bounds the final k-th distance; prune with the smallest.
  float Uf = a.U[j];
  for (int sp = 1; sp < a.nsplit; ++sp) Uf = fminf(Uf, a.U[(size_t)sp * a.n + j]);
  const float Ulim = Uf * (1.0f + 0x1p-20f) + 1e-30f;
  double rd = hd[0];
  int ri = hi[0];
  const int capg = a.cap / a.G;
  for (int sp = 0; sp < a.nsplit; ++sp) {
    bool over = false;
    for (int gg = 0; gg < a.G; ++gg) over |= list_count(a, sp, gg, j) > capg;
    const int cbeg = sp * a.cps;
    // An overflowed split is re-scanned once; otherwise each of its lists.
  for (int gg = 0; gg < (over ? 1 : a.G); ++gg) {
    const int nscan = over ? min(a.m, cbeg + a.cps) - cbeg : list_count(a, sp, gg, j);
    const size_t lb = list_base(a, sp, gg, j);
    for (int base = 0; base < nscan; base += 64) {
      // One list entry (or, for an overflowed list, one candidate) per lane.
      const int e = base + lane;
      int cid = 0;
      float lw = -__builtin_inff();  // the entry's lower bound (none when re-scanning)
      bool keep = e < nscan;
      if (keep) {
        if (over) {
          cid = a.amap[cbeg + e];
        } else {
          const int2 en = a.lst[lb + e];
          cid = en.x;
          lw = __int_as_float(en.y);
          keep = lw <= Ulim && (double)lw <= rd;
        }
      }
      unsigned long long live = __ballot(keep);
      while (live) {
        // Up to kRowBatch kept entries: lane t < nb takes the t-th lowest set bit.
        int nb = 0, mysrc = 0;
#pragma unroll
        for (int t = 0; t < kRowBatch; ++t) {
          if (live) {
            if (lane == t) mysrc = __builtin_ctzll(live);
            live &= live - 1;
            ++nb;
          }
        }
        const int sid = __shfl(cid, mysrc);
        const int myid = lane < nb ? sid : 0;
        const double v = wave_row_dists(a.XA, d, myid, nb, qv);
        // Apply the results in list order (canonical selection does not
        // depend on it).
        for (int t = 0; t < nb; ++t) {
          const double dist = __shfl(v, kRowLanes * t);
          const int tid = __shfl(myid, t);
          if (cless_d(dist, tid, rd, ri)) {
            if (lane == 0) heap_sift_d(hd, hi, 0, k, dist, tid);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            rd = hd[0];
            ri = hi[0];
          }
        }
        // Entries whose lower bound exceeds the row's k-th distance can no
        // longer enter it.
        live &= __ballot((double)lw <= rd);
      }
    }
  }
  }
  if (lane == 0) {
    for (int e2 = k - 1; e2 > 0; --e2) {
      const double xd = hd[e2];
      const int xi = hi[e2];
      hd[e2] = hd[0];
      hi[e2] = hi[0];
      heap_sift_d(hd, hi, 0, e2, xd, xi);
    }
    for (int t = 0; t < k; ++t) {
      a.D[row + t] = hd[k - 1 - t];
      a.I[row + t] = hi[k - 1 - t];
    }
  }
}

// only_over: take only the queries with an overflowed list (the selection
// form leaves them).
__global__ __launch_bounds__(256) void knn_rerank_kernel(Bf16Args a, int only_over) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k;
  const int w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  double *hd = (double *)smem + (size_t)w * k;
  int *hi = (int *)(smem + al16r(sizeof(double) * 4 * (size_t)k)) + (size_t)w * k;
  if (j >= a.n) return;  // whole wave leaves together; no block barrier below
  if (only_over) {
    bool over = false;
    for (int sp = 0; sp < a.nsplit; ++sp)
      for (int gg = 0; gg < a.G; ++gg) over |= list_count(a, sp, gg, j) > a.cap / a.G;
    if (!over) return;
  }
  rerank_heap_path(a, j, hd, hi);
}

// Selection form of pass 2 (k <= kSelMaxK), one wave per query; a query
// with an overflowed list is left to knn_rerank_kernel:
//  1. T = the k-th smallest of {upper bounds of the listed entries} U {the
//     resident row's distances rounded up to FP32}, bounded from above by a
//     radix select over the top 24 bits of the keys (three 8-bit digits,
//     LDS histogram; the keys are staged in LDS by the first digit's pass):
//     T is the largest value with the k-th key's 24-bit prefix, at most
//     2^-15 relative above it.  k items have distances at or below T, so the
//     final k-th distance is at most T: an entry whose lower bound exceeds T
//     (with a relative slack of 2^-20 for the FP64 rounding of the re-ranked
//     sums) cannot enter the row.  Every candidate with one of the k
//     smallest upper bounds was listed, so T is (up to the 2^-15) the exact U
//     of the whole candidate stream.
//  2. The entries at or below T join the resident row in an LDS pool of
//     64 NS entries, then get their FP64 direct-difference distances
//     kRowBatch at a time (wave_row_dists).
//  3. The pool's k smallest entries in (distance, index) order are found by
//     ranking every entry against the pool (equal pairs, duplicates of one
//     candidate, break by pool slot); a full pool is cut back to k the same
//     way.  The row is written in descending order (a valid max-heap).
// List loads go eight per lane at a time, so one memory latency covers 512
// entries.
constexpr int kSelMaxK = 256;
constexpr int kSelKeys = 1024;                        // keys staged in LDS per wave
constexpr int kSelUnroll = 8;                         // list loads per lane in flight

// Pool entries per lane: room for k plus at least two windows of 64.
constexpr int sel_slots(int k) { return (k + 128 + 63) / 64; }

struct SelLayout {
  size_t hist, keys, pd, pi, bytes;
  int P;
  __host__ __device__ SelLayout(int k) {
    P = 64 * sel_slots(k);
    hist = 0;
    keys = al16r(256 * 4);
    pd = al16r(keys + (size_t)kSelKeys * 4);
    pi = al16r(pd + (size_t)P * 8);
    bytes = al16r(pi + (size_t)P * 4);
  }
};

// Order-preserving map of FP32 bit patterns to unsigned keys and back.
__device__ __forceinline__ unsigned fkey(float f) {
  const unsigned b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float keyf(unsigned key) {
  return __uint_as_float((key & 0x80000000u) ? (key & 0x7fffffffu) : ~key);
}

template <int NS>
__global__ __launch_bounds__(256) void knn_rerank_select_kernel(Bf16Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k, d = a.d;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  const SelLayout SL(k);
  unsigned char *wbase = smem + (size_t)w * SL.bytes;
  unsigned *hist = (unsigned *)(wbase + SL.hist);
  unsigned *keys = (unsigned *)(wbase + SL.keys);
  double *pd = (double *)(wbase + SL.pd);
  int *pi = (int *)(wbase + SL.pi);
  if (j >= a.n) return;  // whole wave leaves together; no block barrier below
  bool over = false;
  const int nl = a.nsplit * a.G;  // lists of this query: list L is (L / G, L % G)
  int ntot = k;  // keys: the resident row, then each list's entries
  for (int L = 0; L < nl; ++L) {
    const int c = list_count(a, L / a.G, L % a.G, j);
    over |= c > a.cap / a.G;
    ntot += c;
  }
  if (over) return;
  auto wsync = []() __attribute__((always_inline)) {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  };
  const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
  double qv[4];
  load_query4(a.XB + (size_t)a.bmap[j] * d, d, qv);
  for (int t = lane; t < k; t += 64) {
    pd[t] = a.D[row + t];
    pi[t] = a.I[row + t];
  }
  for (int b = lane; b < 256; b += 64) hist[b] = 0u;
  wsync();

  // 1. Radix select of the k-th smallest key.  Digit 0 reads the keys from
  // the row and the lists and stages the first kSelKeys in LDS.
  for (int t = lane; t < k; t += 64) {
    const unsigned key = fkey(__double2float_ru(pd[t]));
    if (t < kSelKeys) keys[t] = key;
    atomicAdd(&hist[key >> 24], 1u);
  }
  {
    int kbase = k;  // key index of this list's first entry
    for (int L = 0; L < nl; ++L) {
      const int cnt = list_count(a, L / a.G, L % a.G, j);
      const float *up = a.lstu + list_base(a, L / a.G, L % a.G, j);
      for (int base = 0; base < cnt; base += 64 * kSelUnroll) {
        float v[kSelUnroll];
#pragma unroll
        for (int u = 0; u < kSelUnroll; ++u) {
          const int e = base + lane + 64 * u;
          v[u] = e < cnt ? up[e] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kSelUnroll; ++u) {
          const int e = base + lane + 64 * u;
          if (e < cnt) {
            const unsigned key = fkey(v[u]);
            if (kbase + e < kSelKeys) keys[kbase + e] = key;
            atomicAdd(&hist[key >> 24], 1u);
          }
        }
      }
      kbase += cnt;
    }
  }
  unsigned prefix = 0;
  int rank = k;  // 1-based rank still to find below the current prefix
#pragma unroll 1
  for (int pass = 0; pass < 3; ++pass) {
    const int sh = 24 - 8 * pass;
    if (pass > 0) {
      for (int b = lane; b < 256; b += 64) hist[b] = 0u;
      wsync();
      auto count = [&](unsigned key) __attribute__((always_inline)) {
        if (((key >> sh) >> 8) == prefix) atomicAdd(&hist[(key >> sh) & 255u], 1u);
      };
      const int ns = ntot < kSelKeys ? ntot : kSelKeys;
      for (int e = lane; e < ns; e += 64) count(keys[e]);
      if (ntot > kSelKeys) {
        // Keys beyond the LDS stage are read again from the lists.
        int kbase = k;
        for (int L = 0; L < nl; ++L) {
          const int cnt = list_count(a, L / a.G, L % a.G, j);
          const float *up = a.lstu + list_base(a, L / a.G, L % a.G, j);
          for (int e = lane; e < cnt; e += 64)
            if (kbase + e >= kSelKeys) count(fkey(up[e]));
          kbase += cnt;
        }
      }
    }
    wsync();
    // Lane l holds bins 4 l .. 4 l + 3; inclusive scan over the lanes.
    const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                   h3 = hist[4 * lane + 3];
    const int local = (int)(h0 + h1 + h2 + h3);
    int incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    const int excl = incl - local;
    const unsigned long long hitm = __ballot(excl < rank && rank <= incl);
    const int src = (int)__builtin_ctzll(hitm);
    int bin = 0, before = excl;
    {
      int c = excl;
      const unsigned hv4[4] = {h0, h1, h2, h3};
      bool found = false;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!found && c + (int)hv4[t] >= rank) {
          bin = 4 * lane + t;
          before = c;
          found = true;
        }
        c += (int)hv4[t];
      }
    }
    bin = __shfl(bin, src);
    before = __shfl(before, src);
    prefix = (prefix << 8) | (unsigned)bin;
    rank -= before;
    wsync();  // histogram read before the next pass clears it
  }
  // The largest key with the 24-bit prefix (+inf when the prefix is
  // +inf's: keys above it are NaNs).
  const unsigned tkey = min((prefix << 8) | 0xFFu, 0xFF800000u);
  const float T = keyf(tkey);
  const float Tlim = T * (1.0f + 0x1p-20f) + 1e-30f;

  int np = k;    // pool entries; slots np0 .. np - 1 still lack distances
  int np0 = k;
  // 2b. Distances of the pending pool slots, kRowBatch rows at a time.
  auto distances = [&]() {
    wsync();
    for (int s0 = np0; s0 < np; s0 += kRowBatch) {
      const int nb = min(kRowBatch, np - s0);
      const int myid = lane < nb ? pi[s0 + lane] : 0;
      const double v = wave_row_dists(a.XA, d, myid, nb, qv);
      if (lane % kRowLanes == 0 && lane / kRowLanes < nb) pd[s0 + lane / kRowLanes] = v;
    }
    np0 = np;
  };
  // 3. Keep the k smallest pool entries; the last round writes the row.
  auto select = [&](bool last) {
    distances();
    wsync();
    double od[NS];
    int oi[NS], orank[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int e = lane + 64 * s;
      od[s] = 0.0;
      oi[s] = 0;
      orank[s] = 0;
      if (e < np) {
        od[s] = pd[e];
        oi[s] = pi[e];
      }
    }
    const int ns = (np + 63) >> 6;  // occupied slots per lane
    for (int jj = 0; jj < np; ++jj) {
      const double dj = pd[jj];
      const int ij = pi[jj];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (s < ns) {
          const int e = lane + 64 * s;
          const bool lt = dj < od[s] || (dj == od[s] && (ij < oi[s] || (ij == oi[s] && jj < e)));
          orank[s] += lt ? 1 : 0;
        }
      }
    }
    wsync();  // every rank computed before the pool is rewritten
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int e = lane + 64 * s;
      if (e < np && orank[s] < k) {
        if (last) {
          a.D[row + k - 1 - orank[s]] = od[s];
          a.I[row + k - 1 - orank[s]] = oi[s];
        } else {
          pd[orank[s]] = od[s];
          pi[orank[s]] = oi[s];
        }
      }
    }
    np = np0 = k;
    wsync();
  };

  // 2a. Compact the entries at or below T into the pool.
  // A batch whose survivors would overflow the pool is taken one window at a
  // time, cutting the pool back to k before a window that does not fit.
  for (int L = 0; L < nl; ++L) {
    const int cnt = list_count(a, L / a.G, L % a.G, j);
    const int2 *ls = a.lst + list_base(a, L / a.G, L % a.G, j);
    for (int base = 0; base < cnt; base += 64 * kSelUnroll) {
      int2 en[kSelUnroll];
#pragma unroll
      for (int u = 0; u < kSelUnroll; ++u) {
        const int e = base + lane + 64 * u;
        en[u] = e < cnt ? ls[e] : make_int2(0, __float_as_int(__builtin_inff()));
      }
      bool keep[kSelUnroll];
      int id[kSelUnroll];
      int nkeep = 0;
#pragma unroll
      for (int u = 0; u < kSelUnroll; ++u) {
        keep[u] = base + lane + 64 * u < cnt && __int_as_float(en[u].y) <= Tlim;
        id[u] = en[u].x;
        nkeep += __popcll(__ballot(keep[u]));
      }
      if (np + nkeep <= SL.P) {
#pragma unroll
        for (int u = 0; u < kSelUnroll; ++u) {
          const unsigned long long live = __ballot(keep[u]);
          if (keep[u]) {
            const int at = np + __builtin_amdgcn_mbcnt_hi((unsigned)(live >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)live, 0u));
            pi[at] = id[u];
          }
          np += __popcll(live);
        }
      } else {
#pragma unroll 1
        for (int u = 0; u < kSelUnroll; ++u) {
          const int e = base + lane + 64 * u;
          bool kp = false;
          int cid = 0;
          if (e < cnt) {
            const int2 en1 = ls[e];
            kp = __int_as_float(en1.y) <= Tlim;
            cid = en1.x;
          }
          const unsigned long long live = __ballot(kp);
          if (live == 0ull) continue;
          if (np + __popcll(live) > SL.P) select(false);
          if (kp) {
            const int at = np + __builtin_amdgcn_mbcnt_hi((unsigned)(live >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)live, 0u));
            pi[at] = cid;
          }
          np += __popcll(live);
        }
      }
    }
  }
  select(true);
}

namespace {

constexpr size_t kLds = 160 * 1024;
constexpr int kCBb = 4;

// Register form: upper-bound set size per lane, two query blocks per wave
// for k <= 4 kRegRK, one query block (fewer query registers, larger sets)
// for 4 kRegRK < k <= 4 kRegRK1.
constexpr int kRegRK = 16;
constexpr int kRegRK1 = 32;

struct Bf16Plan {
  int ks = 0, qb = 0, w = 0, q = 0, cap = 0, nsplit = 1, cps = 0;
  bool pl = false;   // per-lane upper-bound heaps
  bool reg = false;  // register form of the filter (knn_bf16_filter2_kernel)
  int G = 1;         // list segments per (split, query)
  size_t lds = 0;
};

// Split the candidates when the query blocks alone give fewer than two
// workgroups per CU; each split keeps at least 16 tiles.
void plan_splits(Bf16Plan &p, int m, int n) {
  const int gx = (n + p.q - 1) / p.q;
  const int C = 16 * kCBb;
  int ns = 1;
  if (gx < 512) {
    ns = (512 + gx - 1) / gx;
    const int maxs = m / (16 * C);
    if (ns > maxs) ns = maxs;
    if (ns < 1) ns = 1;
  }
  int cps = (m + ns - 1) / ns;
  cps = (cps + C - 1) / C * C;
  p.nsplit = (m + cps - 1) / cps;
  p.cps = cps;
}

Bf16Plan plan_bf16(int d, int k) {
  Bf16Plan p;
  const int need = (d + 31) / 32;
  for (int ks : {1, 2, 4, 8}) {
    if (need <= ks) {
      p.ks = ks;
      break;
    }
  }
  if (p.ks == 0) return p;
  // Shapes (QB, W) in order of preference; GSKNN_BF16_CFG=<index> pins one
  // (tuning).
  constexpr int kOpts = 4;
  const int opts[kOpts][2] = {{2, 4}, {1, 4}, {1, 2}, {1, 1}};
  static const int pinned = [] {
    const char *e = getenv("GSKNN_BF16_CFG");
    return e ? atoi(e) : -1;
  }();
  for (int c = 0; c < kOpts; ++c) {
    if (pinned >= 0 && pinned < kOpts && c != pinned) continue;
    const int *o = opts[c];
    const int Q = o[1] * 16 * o[0];
    const size_t b = FilterLayout(p.ks, kCBb, Q, k).bytes;
    if (b <= kLds) {
      p.qb = o[0];
      p.w = o[1];
      p.q = Q;
      p.lds = b;
      break;
    }
  }
  // Per-lane upper bounds (the register form for k <= 4 kRegRK, the LDS form's
  // per-lane heaps from k = 48 on) get twice the list capacity: an
  // overflowing list is re-scanned against every candidate.  In the LDS form
  // the largest of four short-heap roots is far above the exact U for small
  // k (one heap per query there); the register form's warm-up pass sets U
  // from each lane's tile minima first, and at config-3 size it runs k = 10,
  // 16, 24, 32 in 12.3, 13.0, 13.3, 14.0 ms against 19.4, 22.2, 25.7, 28.6 ms
  // for the LDS form, with shorter lists.  GSKNN_BF16_REG_MINK=<k> keeps the
  // LDS form below k (tuning).
  static const int reg_mink = [] {
    const char *e = getenv("GSKNN_BF16_REG_MINK");
    return e ? atoi(e) : 1;
  }();
  p.pl = k >= 48 || (k >= reg_mink && k <= 4 * kRegRK1);
  const int cap = (p.pl ? 32 : 16) * k;
  const int lo = p.pl ? 1024 : 512, hi = p.pl ? 8192 : 4096;
  p.cap = cap < lo ? lo : (cap > hi ? hi : cap);
  // Register form for k <= 4 kRegRK (two queries blocks, four waves;
  // GSKNN_BF16_REG=0 keeps the LDS-heap form, for comparison).
  static const bool reg_on = [] {
    const char *e = getenv("GSKNN_BF16_REG");
    return !(e && atoi(e) == 0);
  }();
  static const int reg1_on = [] {
    const char *e = getenv("GSKNN_BF16_REG1");
    return !(e && atoi(e) == 0);
  }();
  if (reg_on && p.pl && (k <= 4 * kRegRK || (reg1_on && k <= 4 * kRegRK1)) &&
      Filter2Layout(p.ks, kCBb).bytes <= kLds) {
    p.reg = true;
    p.qb = k <= 4 * kRegRK ? 2 : 1;
    // At d > 192 (KS = 8) eight waves share each tile: the two 32 KiB tile
    // buffers leave two workgroups per CU, so four waves per SIMD instead of
    // two hide the survivor loop's latency (config 3: 14.6 / 14.8 ms against
    // 16.1 / 16.2 on one box; GSKNN_BF16_W8=0 keeps four waves, for
    // comparison).
    static const bool w8 = [] {
      const char *e = getenv("GSKNN_BF16_W8");
      return !(e && atoi(e) == 0);
    }();
    p.w = w8 && p.ks == 8 ? 8 : 4;
    p.q = p.w * 16 * p.qb;
    p.lds = Filter2Layout(p.ks, kCBb).bytes;
    p.G = 4;
  }
  return p;
}

// panel_bytes: the register form's candidate panel (0 for the other forms).
size_t bf16_work_layout(int n, int cap, int nsplit, int G, size_t panel_bytes, size_t *o_lst,
                        size_t *o_cnt, size_t *o_u, size_t *o_over, size_t *o_lstu = nullptr,
                        size_t *o_panel = nullptr) {
  size_t off = 0;
  const size_t rows = (size_t)n * nsplit;
  *o_lst = off;
  off = al16r(off + rows * cap * sizeof(int2));
  *o_cnt = off;
  off = al16r(off + rows * G * 4);
  *o_u = off;
  off = al16r(off + rows * 4);
  *o_over = off;
  off = al16r(off + 16);
  if (o_lstu) *o_lstu = off;
  off = al16r(off + rows * cap * sizeof(float));
  if (o_panel) *o_panel = off;
  off = al16r(off + panel_bytes);
  return off;
}

size_t panel_bytes(const Bf16Plan &p, int m, int d) {
  return p.reg ? PanelLayout(m, bf16_row_dims(d)).bytes : 0;
}

using SelKernel = void (*)(Bf16Args);
SelKernel sel_kernel(int ns) {
  switch (ns) {
    case 3: return knn_rerank_select_kernel<3>;
    case 4: return knn_rerank_select_kernel<4>;
    case 5: return knn_rerank_select_kernel<5>;
    case 6: return knn_rerank_select_kernel<6>;
    default: return knn_rerank_select_kernel<7>;
  }
}

template <int KS>
hipError_t launch_filter_ks(const Bf16Plan &p, const Bf16Args &a, hipStream_t s) {
  const dim3 grid((a.n + p.q - 1) / p.q, a.nsplit);
  if (p.reg) {
    auto kern = p.qb == 2 ? knn_bf16_filter2_kernel<KS, 2, kCBb, 4, kRegRK>
                          : knn_bf16_filter2_kernel<KS, 1, kCBb, 4, kRegRK1>;
    if constexpr (KS == 8) {
      if (p.w == 8)
        kern = p.qb == 2 ? knn_bf16_filter2_kernel<KS, 2, kCBb, 8, kRegRK>
                         : knn_bf16_filter2_kernel<KS, 1, kCBb, 8, kRegRK1>;
    }
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)p.lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, grid, dim3(64 * p.w), p.lds, s, a);
    return hipGetLastError();
  }
#define GSKNN_BF16_CASE(QB_, W_)                                                       \
  if (p.qb == QB_ && p.w == W_) {                                                      \
    auto kern = p.pl ? knn_bf16_filter_kernel<KS, QB_, kCBb, W_, true>                  \
                     : knn_bf16_filter_kernel<KS, QB_, kCBb, W_, false>;                \
    hipError_t e = hipFuncSetAttribute((const void *)kern,                              \
                                       hipFuncAttributeMaxDynamicSharedMemorySize,      \
                                       (int)p.lds);                                     \
    if (e != hipSuccess) return e;                                                      \
    hipLaunchKernelGGL(kern, grid, dim3(64 * W_), p.lds, s, a);                         \
    return hipGetLastError();                                                           \
  }
  GSKNN_BF16_CASE(2, 4)
  GSKNN_BF16_CASE(1, 4)
  GSKNN_BF16_CASE(1, 2)
  GSKNN_BF16_CASE(1, 1)
#undef GSKNN_BF16_CASE
  return hipErrorNotSupported;
}

}  // namespace

size_t bf16_work_bytes(int m, int n, int d, int k) {
  Bf16Plan p = plan_bf16(d, k);
  if (p.ks == 0 || p.q == 0 || n <= 0 || m <= 0) return 0;
  plan_splits(p, m, n);
  size_t a, b, c, e;
  return bf16_work_layout(n, p.cap, p.nsplit, p.G, panel_bytes(p, m, d), &a, &b, &c, &e);
}

size_t bf16_count_offset(int m, int n, int d, int k, int *nsplit) {
  Bf16Plan p = plan_bf16(d, k);
  if (p.ks == 0 || p.q == 0 || n <= 0 || m <= 0) return 0;
  plan_splits(p, m, n);
  size_t o_lst, o_cnt, o_u, o_over;
  bf16_work_layout(n, p.cap, p.nsplit, p.G, panel_bytes(p, m, d), &o_lst, &o_cnt, &o_u, &o_over);
  if (nsplit) *nsplit = p.nsplit * p.G;  // lists per query
  return o_cnt;
}

hipError_t launch_knn_bf16(int m, int n, int d, int k, int NA, int NB, const float *XA,
                           const void *prepA, const int *amap, const float *XB, const void *prepB,
                           const int *bmap, double *D, int *I, int ldk, const int *hrow,
                           void *work, size_t work_bytes, int *overflow_out, hipStream_t s) {
  if (m <= 0 || n <= 0 || d <= 0 || k <= 0) return hipSuccess;
  Bf16Plan p = plan_bf16(d, k);
  if (p.ks == 0 || p.q == 0) return hipErrorNotSupported;
  plan_splits(p, m, n);
  const size_t rl = al16r(sizeof(double) * 4 * (size_t)k) + sizeof(int) * 4 * (size_t)k;
  if (rl > kLds) return hipErrorNotSupported;
  size_t o_lst, o_cnt, o_u, o_over, o_lstu;
  size_t o_panel;
  const size_t need = bf16_work_layout(n, p.cap, p.nsplit, p.G, panel_bytes(p, m, d), &o_lst, &o_cnt,
                                       &o_u, &o_over, &o_lstu, &o_panel);
  if (work == nullptr || work_bytes < need) return hipErrorInvalidValue;
  Bf16Args a{};
  a.m = m; a.n = n; a.d = d; a.k = k; a.ldk = ldk; a.cap = p.cap;
  a.nsplit = p.nsplit; a.cps = p.cps; a.G = p.G;
  a.NA = NA; a.NB = NB;
  a.XA = XA; a.XB = XB;
  a.prepA = (const unsigned char *)prepA;
  a.prepB = (const unsigned char *)prepB;
  a.amap = amap; a.bmap = bmap;
  a.D = D; a.I = I; a.hrow = hrow;
  a.lst = (int2 *)((char *)work + o_lst);
  a.lstu = (float *)((char *)work + o_lstu);
  a.panel = (const unsigned char *)work + o_panel;
  a.cnt = (int *)((char *)work + o_cnt);
  a.U = (float *)((char *)work + o_u);
  a.nover = (int *)((char *)work + o_over);
  // Warm-up pass of the register form: the first 1/16 of each split's tiles
  // (GSKNN_BF16_WARM=<tiles>, read at each call, sets the count).
  const char *warm_e = getenv("GSKNN_BF16_WARM");
  const int warm_env = warm_e ? atoi(warm_e) : -1;
  a.warm = p.reg ? (warm_env >= 0 ? warm_env : p.cps / (16 * kCBb) / 16) : 0;
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, s, a.nover);
  if (p.reg)
    hipLaunchKernelGGL(bf16_gather_kernel, dim3((m + kPanelPad + 3) / 4), dim3(256), 0, s, m, NA, d, amap,
                       a.prepA, (unsigned char *)work + o_panel);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (p.ks) {
    case 1: e = launch_filter_ks<1>(p, a, s); break;
    case 2: e = launch_filter_ks<2>(p, a, s); break;
    case 4: e = launch_filter_ks<4>(p, a, s); break;
    case 8: e = launch_filter_ks<8>(p, a, s); break;
    default: return hipErrorNotSupported;
  }
  if (e != hipSuccess) return e;
  // Selection form, then the heap form for the queries it leaves
  // (overflowed lists), or the heap form for every query above kSelMaxK.
  const bool sel = k <= kSelMaxK;
  if (sel) {
    const size_t sl = 4 * SelLayout(k).bytes;
    auto kern = sel_kernel(sel_slots(k));
    e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sl);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((n + 3) / 4), dim3(256), sl, s, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  e = hipFuncSetAttribute((const void *)knn_rerank_kernel,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)rl);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(knn_rerank_kernel, dim3((n + 3) / 4), dim3(256), rl, s, a, sel ? 1 : 0);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (overflow_out)
    e = hipMemcpyAsync(overflow_out, a.nover, sizeof(int), hipMemcpyDeviceToDevice, s);
  return e;
}

}  // namespace gsknn
