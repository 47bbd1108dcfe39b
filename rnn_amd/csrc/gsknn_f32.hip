// FP32 exact kNN (sgsknn) from an FP32 MFMA candidate pass and a re-rank in
// the reference oracle's own arithmetic.
//
// Reference: sgsknn_ref (frame/gsknn_ref.c:63-109, frame/sgsknn_ref_impl.h:
// 41-103): distance = ((-2 * sum_p a_p b_p) + XA2[amap[i]]) + XB2[bmap[j]] with
// the dot product summed in p order in FP32, every product and sum rounded
// separately; per heap row the k smallest are kept (frame/gsknn_heap.c:104-130).
//
// Pass 1 (knn_f32_filter_kernel) computes dist~ with v_mfma_f32_16x16x4_f32
// (FP32 products and sums in the matrix core's own order) and brackets the
// oracle's distance d_or by
//   |dist~ - d_or| <= delta = c_d |a| |b| + 4u (|nc| + |nq|),
//   c_d = (2 g_m + 2 g_o + 8.1 u),  g_m = (2d+2) 2u / (1 - (2d+2) 2u),
//   g_o = d u / (1 - d u),  u = 2^-24,
// where |a|, |b| are the true Euclidean norms of the points (computed in FP64,
// rounded up), g_m covers the MFMA sum (each product or addition rounding to
// within 2u, any order), g_o the oracle's sequential sum, and the 4u terms the
// two FP32 epilogue additions on each side.  As in the BF16 form, U = k-th
// smallest upper bound per query, and every candidate whose lower bound is <= U
// when seen is listed.
// Pass 2 recomputes the listed candidates with the oracle's arithmetic (one
// lane per candidate, p in order, no fused multiply-add) and keeps the k
// smallest of the row and the candidates under the canonical (distance,
// index) order: the distances are bit-identical to sgsknn_ref's.  For k <= 256
// a pool and rank selection do the merge (knn_f32_rerank_select_kernel),
// above it an LDS heap (knn_f32_rerank_kernel).
#include <limits.h>

#include "gsknn_common.h"

namespace gsknn {

namespace {

__host__ __device__ constexpr size_t a16(size_t x) { return (x + 15) & ~size_t(15); }

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kU = 5.9604644775390625e-08f;  // 2^-24

// x moved outward by 2^-22 |x|, so that a bound computed with round-to-nearest
// additions (each off by at most 2^-24 of its result) still holds.
__device__ __forceinline__ float widen_up(float x) { return __builtin_fmaf(0x1p-22f, fabsf(x), x); }
__device__ __forceinline__ float widen_dn(float x) { return __builtin_fmaf(-0x1p-22f, fabsf(x), x); }

__device__ __forceinline__ bool cl_f(float da, int ia, float db, int ib) {
  return da < db || (da == db && ia < ib);
}

__device__ void sift_v(float *h, int QS, int q, int k, float x) {
  int s = 0;
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

__device__ void sift_fi(float *hd, int *hi, int s, int k, float xd, int xi) {
  while (true) {
    int c = 2 * s + 1;
    if (c >= k) break;
    float cd = hd[c];
    int ci = hi[c];
    if (c + 1 < k && cl_f(cd, ci, hd[c + 1], hi[c + 1])) {
      ++c;
      cd = hd[c];
      ci = hi[c];
    }
    if (!cl_f(xd, xi, cd, ci)) break;
    hd[s] = cd;
    hi[s] = ci;
    s = c;
  }
  hd[s] = xd;
  hi[s] = xi;
}

}  // namespace

struct F32Args {
  int m, n, d, k, ldk, cap;
  int nsplit, cps;  // candidate splits (grid.y) and candidates per split
  const float *XA, *XA2;
  const int *amap;
  const float *XB, *XB2;
  const int *bmap;
  float *D;
  int *I;
  const int *hrow;
  float *xnA;   // [m] |XA[amap[i]]| rounded up
  float *xnB;   // [n] |XB[bmap[j]]| rounded up
  float *mx;    // {max xnA, max |XA2[amap[i]]|}
  int2 *lst;    // [nsplit][n][cap]
  int *cnt;     // [nsplit][n]
  float *U;     // [nsplit][n]
  int *nover;
  int G;                       // list segments per (split, query): cnt is [nsplit][G][n]
  const unsigned char *panel;  // register form: candidates in position order
  int warm;                    // register form: tiles of each split in the warm-up pass
};

// Norms of the gathered points (FP64 sum, rounded up) and candidate maxima.
__global__ void f32_prep_kernel(F32Args a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool vec = (a.d & 3) == 0 && (((uintptr_t)a.XA | (uintptr_t)a.XB) & 15) == 0;
  // |x| in FP64, p in order, rounded up; rows read 16 bytes at a time when
  // d is a multiple of 4 and the tables are 16-byte aligned.
  auto norm = [&](const float *x) -> float {
    double s = 0.0;
    int p = 0;
    if (vec) {
      for (; p < a.d; p += 4) {
        const f32x4 v = *(const f32x4 *)(x + p);
        s += (double)v.x * (double)v.x;
        s += (double)v.y * (double)v.y;
        s += (double)v.z * (double)v.z;
        s += (double)v.w * (double)v.w;
      }
    }
    for (; p < a.d; ++p) s += (double)x[p] * (double)x[p];
    return __double2float_ru(sqrt(s) * (1.0 + 0x1p-40));
  };
  // Bit patterns of this lane's candidates for the table maxima (the values
  // are >= 0 or NaN, so their bits order as integers).
  int m0 = 0, m1 = 0;
  if (i < a.m) {
    const int id = a.amap[i];
    const float xn = norm(a.XA + (size_t)id * a.d);
    a.xnA[i] = xn;
    m0 = __float_as_int(xn);
    m1 = __float_as_int(fabsf(a.XA2[id]));
  } else if (i < a.m + a.n) {
    const int j = i - a.m;
    a.xnB[j] = norm(a.XB + (size_t)a.bmap[j] * a.d);
  }
  // Wave maxima first, then one atomic per wave instead of one per row.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    m0 = max(m0, __shfl_xor(m0, off));
    m1 = max(m1, __shfl_xor(m1, off));
  }
  if ((threadIdx.x & 63) == 0) {
    if (m0 > 0) atomicMax((int *)&a.mx[0], m0);
    if (m1 > 0) atomicMax((int *)&a.mx[1], m1);
  }
}

__global__ void f32_zero_kernel(float *mx, int *nover) {
  if (threadIdx.x < 2) mx[threadIdx.x] = 0.0f;
  if (threadIdx.x == 0) *nover = 0;
}

struct F32Layout {
  size_t tA, nc, xn, heap, cnt, bytes;
  __host__ __device__ F32Layout(int KS, int CB, int Q, int k) {
    const size_t tile = (size_t)KS * CB * 64;  // floats
    const size_t C = 16 * (size_t)CB;
    tA = 0;
    nc = a16(tile * 4);
    xn = a16(nc + C * 4);
    heap = a16(xn + C * 4);
    cnt = a16(heap + (size_t)k * Q * 4);
    bytes = a16(cnt + (size_t)Q * 4);
  }
};

// Same structure as the BF16 filter: candidate tiles global -> registers ->
// one LDS buffer in MFMA fragment order (lane l of k-step s, block cb holds
// candidate cb*16 + l%16, dim 4s + l/16), query fragments in registers, a
// one-compare hot filter and a per-lane deferred queue for the survivors.
// Two waves per SIMD when the query fragments (QB * KS registers) leave
// room for the rest.
constexpr int f32_waves_per_eu(int KS, int QB) { return QB * KS <= 128 ? 2 : 1; }

template <int KS, int QB, int CB, int W>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(f32_waves_per_eu(KS, QB)))) void knn_f32_filter_kernel(F32Args a) {
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB;
  constexpr int QR = 4;
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k, d = a.d;
  const F32Layout L(KS, CB, Q, k);
  float *tA = (float *)(smem + L.tA);
  float *tnc = (float *)(smem + L.nc);
  float *txn = (float *)(smem + L.xn);
  float *hv = (float *)(smem + L.heap);
  int *hc = (int *)(smem + L.cnt);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, grp = lane >> 4;
  const int qbase = blockIdx.x * Q;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const size_t sbase = (size_t)split * a.n;

  if (tid < Q) {
    const int j = qbase + tid;
    if (j < a.n) {
      const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
      for (int t = 0; t < k; ++t) hv[t * Q + tid] = a.D[row + t];
      for (int s = k / 2 - 1; s >= 0; --s) {
        float x = hv[s * Q + tid];
        int p = s;
        while (true) {
          int c = 2 * p + 1;
          if (c >= k) break;
          float cv = hv[c * Q + tid];
          if (c + 1 < k && cv < hv[(c + 1) * Q + tid]) cv = hv[(++c) * Q + tid];
          if (!(x < cv)) break;
          hv[p * Q + tid] = cv;
          p = c;
        }
        hv[p * Q + tid] = x;
      }
    }
    hc[tid] = 0;
  }

  const float u = kU;
  const double gm = (2.0 * d + 2) * 2.0 * 0x1p-24 / (1.0 - (2.0 * d + 2) * 2.0 * 0x1p-24);
  const double go = d * 0x1p-24 / (1.0 - d * 0x1p-24);
  const float cd = (float)((2.0 * gm + 2.0 * go + 8.1 * 0x1p-24) * 1.01);
  float bq[QB][KS];
  float nq[QB], al[QB], cq[QB], dmx[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = w * QW + qb * 16 + (lane & 15);
    const int j = qbase + ql[qb];
    qv[qb] = j < a.n;
    const int jj = qv[qb] ? j : a.n - 1;
    const float *xq = a.XB + (size_t)a.bmap[jj] * d;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int p = 4 * s + grp;
      const float v = xq[p < d ? p : d - 1];
      bq[qb][s] = (qv[qb] && p < d) ? v : 0.0f;
    }
    nq[qb] = a.XB2[a.bmap[jj]];
    al[qb] = cd * a.xnB[jj];
    cq[qb] = 4.0f * u * fabsf(nq[qb]);
    dmx[qb] = (al[qb] * a.mx[0] + 4.0f * u * a.mx[1] + cq[qb]) * 1.002f + 1e-30f;
  }

  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  // Staging: the tile is BLK fragment blocks of 64 floats (k-step s,
  // candidate block cb); wave w stages blocks w * BPW .. w * BPW + BPW - 1 and
  // lane l writes element l of each, i.e. dim 4 s + l / 16 of candidate
  // cb * 16 + l % 16: every LDS store is 256 contiguous bytes (no bank
  // conflicts) and a load instruction reads 16 contiguous bytes of each of
  // 16 candidate rows.  Lanes 0..15 of wave w own the norm slots of
  // candidate blocks w, w + W, ... (one block per wave when W >= CB).
  constexpr int BLK = KS * CB;
  static_assert(BLK % W == 0, "fragment blocks per wave");
  constexpr int BPW = BLK / W;
  const int si = lane & 15, skk = lane >> 4;
  constexpr int NPW = (CB + W - 1) / W;  // norm blocks per wave
  float stage[BPW];
  float st_n[NPW], st_x[NPW];
#pragma unroll
  for (int r = 0; r < NPW; ++r) st_n[r] = st_x[r] = 0.f;
  // Ids of this thread's candidates: of the tile being staged (sid) and of
  // the tile after it (nid, requested a tile ahead so that no row load
  // waits on its amap entry).
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
  auto load_x = [&](int t) {
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      const int p = 4 * (bidx / CB) + skk;
      stage[e] = a.XA[(size_t)stage_id(e) * d + (p < d ? p : d - 1)];
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
          st_n[r] = a.XA2[id];
          st_x[r] = a.xnA[cand_pos(t, nb)];
        }
      }
    }
  };
  auto store_x = [&](int t) {
#pragma unroll
    for (int e = 0; e < BPW; ++e) {
      const int bidx = w * BPW + e;
      const int cb = bidx % CB;
      const int p = 4 * (bidx / CB) + skk;
      const bool ok = cbeg + t * C + cb * 16 + si < cend;
      tA[(size_t)bidx * 64 + lane] = (ok && p < d) ? stage[e] : 0.0f;
    }
    if (skk == 0) {
#pragma unroll
      for (int r = 0; r < NPW; ++r) {
        const int nb = w + r * W;
        if (nb < CB) {
          const int slot = nb * 16 + si;
          const bool ok = cbeg + t * C + slot < cend;
          tnc[slot] = ok ? st_n[r] : __builtin_inff();
          txn[slot] = ok ? st_x[r] : 0.0f;
        }
      }
    }
  };

  if (ntiles > 0) {
    load_ids(0, sid);
    load_ids(ntiles > 1 ? 1 : 0, nid);
    load_x(0);
    __syncthreads();
    store_x(0);
  }
  __syncthreads();
  float thr[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb)
    thr[qb] = qv[qb] ? f32_pass_thr(hv[ql[qb]], nq[qb], dmx[qb]) : -__builtin_inff();

  float pq_lo[QR], pq_up[QR];
  int pq_ps[QR], pq_q[QR];
  int pq_n = 0;
#pragma unroll
  for (int r = 0; r < QR; ++r) {
    pq_lo[r] = pq_up[r] = 0.f;
    pq_ps[r] = pq_q[r] = 0;
  }
  auto apply = [&](int slot, float lower, float upper, int ps) {
    const float Uq = hv[slot];
    if (lower <= Uq) {
      const int at = hc[slot]++;
      if (at < a.cap) a.lst[(sbase + qbase + slot) * a.cap + at] = make_int2(ps, __float_as_int(lower));
    }
    if (upper < Uq) sift_v(hv, Q, slot, k, upper);
  };
  auto refresh = [&]() {
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
      if (qv[qb]) thr[qb] = f32_pass_thr(hv[ql[qb]], nq[qb], dmx[qb]);
  };
  auto flush = [&]() {
#pragma unroll 1
    for (int gsel = 0; gsel < 4; ++gsel) {
      if (grp == gsel) {
        for (int r = 0; r < QR; ++r) {
          if (r < pq_n) {
            int slot = ql[0];
#pragma unroll
            for (int qb = 1; qb < QB; ++qb)
              if (pq_q[r] == qb) slot = ql[qb];
            apply(slot, pq_lo[r], pq_up[r], pq_ps[r]);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    pq_n = 0;
    refresh();
  };

  for (int t = 0; t < ntiles; ++t) {
    const bool more = t + 1 < ntiles;
    sid = nid;
    if (t + 2 < ntiles) load_ids(t + 2, nid);
    if (more) load_x(t + 1);
    f32x4 acc[QB][CB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const float av = tA[(size_t)(s * CB + cb) * 64 + lane];
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bq[qb][s], acc[qb][cb], 0, 0, 0);
      }
    }
    // Hot path: one FMA and one compare per pair, the compare's wave mask
    // kept in scalar registers; per-lane bits only when some pair passes.
    unsigned long long msk[QB][CB * 4];
    unsigned long long anym = 0ull;
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = cb * 16 + 4 * grp + v;
          const float tt = __builtin_fmaf(-2.0f, acc[qb][cb][v], tnc[c]);
          acc[qb][cb][v] = tt;
          msk[qb][cb * 4 + v] = __ballot(tt <= thr[qb]);
          anym |= msk[qb][cb * 4 + v];
        }
    if (anym) {
      unsigned pass[QB];
      int npass = 0;
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        pass[qb] = 0u;
#pragma unroll
        for (int b = 0; b < CB * 4; ++b) pass[qb] |= (unsigned)((msk[qb][b] >> lane) & 1ull) << b;
        npass += __builtin_popcount(pass[qb]);
      }
      if (__any(pq_n + npass > QR)) flush();
      const bool direct = __any(pq_n + npass > QR);
#pragma unroll 1
      for (int gsel = 0; gsel < (direct ? 4 : 1); ++gsel) {
        if (!direct || grp == gsel) {
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
              const int c = (b >> 2) * 16 + 4 * grp + (b & 3);
              if (cbeg + t * C + c >= cend) continue;
              const float dist = tt + nq[qb];
              const float delta = txn[c] * al[qb] + 4.0f * u * fabsf(tnc[c]) + cq[qb] + 1e-30f;
              if (direct) {
                apply(ql[qb], widen_dn(dist - delta), widen_up(dist + delta), cbeg + t * C + c);
              } else {
#pragma unroll
                for (int r = 0; r < QR; ++r)
                  if (pq_n == r) {
                    pq_lo[r] = widen_dn(dist - delta);
                    pq_up[r] = widen_up(dist + delta);
                    pq_ps[r] = cbeg + t * C + c;
                    pq_q[r] = qb;
                  }
                ++pq_n;
              }
            }
          }
        }
        if (direct) {
          __builtin_amdgcn_wave_barrier();
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
      }
      if (direct) refresh();
    }
    __syncthreads();
    if (more) store_x(t + 1);
    __syncthreads();
  }
  if (__any(pq_n > 0)) flush();
  __syncthreads();

  if (tid < Q) {
    const int j = qbase + tid;
    if (j < a.n) {
      a.U[sbase + j] = hv[tid];
      a.cnt[sbase + j] = hc[tid];
      if (hc[tid] > a.cap) atomicAdd(a.nover, 1);
    }
  }
}

// ---------------------------------------------------------------------------
// Register form of pass 1 (d <= 256, k <= 128), the structure of the BF16
// register-form filter on FP32 operands: candidates gathered into a panel in
// position order (FP32 rows zero padded to DP = 16 NB dims and a record
// {XA2, |a| rounded up} per position), tiles of 64 candidates loaded straight
// into LDS (global_load_lds_dwordx4) into two buffers with one barrier per
// tile, per-lane sets of the k/4 smallest upper bounds in registers, one list
// segment per lane group, a warm-up pass over the first tiles.
//
// Dims are dealt in 16-dim blocks: in block j, k-step i of lane group g
// (g = l / 16) carries dim 16 j + 4 g + i, so one 16-byte load (and one
// ds_read_b128) per lane and candidate row gives the operands of four
// k-steps; the query operands follow the same deal.  Which dims a k-step
// carries is free as long as both operands agree: the MFMA sum is bounded in
// any order.
//
// Upper bounds enter the sets clamped at 0: an FP32 oracle distance can be
// slightly negative (no clamp), and the sets order FP32 bit patterns as
// integers, which orders negative values wrongly; a larger bound stays valid.
// ---------------------------------------------------------------------------
constexpr int kF32Pad = 64;
struct F32Panel {
  size_t rows, rec, bytes;
  __host__ __device__ F32Panel(int m, int DP) {
    rows = 0;
    rec = a16((size_t)(m + kF32Pad) * DP * 4);
    bytes = a16(rec + (size_t)(m + kF32Pad) * 16);
  }
};

// One wave per candidate position (pad positions written as zeros).
__global__ __launch_bounds__(256) void f32_gather_kernel(F32Args a, int DP, unsigned char *panel) {
  const F32Panel P(a.m, DP);
  const int pos = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pos >= a.m + kF32Pad) return;
  float *dst = (float *)(panel + P.rows) + (size_t)pos * DP;
  if (pos >= a.m) {
    for (int c = lane; c < DP; c += 64) dst[c] = 0.0f;
    if (lane == 0) ((float4 *)(panel + P.rec))[pos] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const int id = a.amap[pos];
  const float *src = a.XA + (size_t)id * a.d;
  for (int c = lane; c < DP; c += 64) dst[c] = c < a.d ? src[c] : 0.0f;
  if (lane == 0) ((float4 *)(panel + P.rec))[pos] = make_float4(a.XA2[id], a.xnA[pos], 0.f, 0.f);
}

struct F32RegLayout {
  size_t tile, nc, buf, bytes;
  __host__ __device__ F32RegLayout(int NB, int CB) {
    tile = 0;
    nc = a16(tile + (size_t)NB * CB * 1024);
    buf = a16(nc + 16 * (size_t)CB * 4);
    bytes = 2 * buf;
  }
};

template <int NB, int QB, int CB, int W, int RK>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(2))) void knn_f32_filter2_kernel(F32Args a) {
  constexpr int QW = 16 * QB, Q = W * QW, C = 16 * CB;
  constexpr int DP = 16 * NB;
  constexpr int BLK = NB * CB;  // 64 x 16-byte fragment blocks per tile
  static_assert(BLK % W == 0, "fragment blocks per wave");
  static_assert(C == 64, "one candidate per lane");
  constexpr int BPW = BLK / W;
  extern __shared__ __align__(16) unsigned char smem[];
  const F32RegLayout L(NB, CB);
  const int k = a.k, d = a.d;
  const F32Panel PL(a.m, DP);
  const float *xg = (const float *)(a.panel + PL.rows);
  const float4 *grec = (const float4 *)(a.panel + PL.rec);

  const int tid = threadIdx.x, lane = tid & 63, grp = lane >> 4, si = lane & 15;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int qbase = blockIdx.x * Q;
  const int split = blockIdx.y;
  const int cbeg = split * a.cps;
  const int cend = min(a.m, cbeg + a.cps);
  const int capg = a.cap / 4;

  // Bound of the FP32 MFMA value against the oracle's FP32 distance (the
  // header of this file): delta = cd |a||b| + 4u (|nc| + |nq|).
  const float u = kU;
  const double gm = (2.0 * d + 2) * 2.0 * 0x1p-24 / (1.0 - (2.0 * d + 2) * 2.0 * 0x1p-24);
  const double go = d * 0x1p-24 / (1.0 - d * 0x1p-24);
  const float cd = (float)((2.0 * gm + 2.0 * go + 8.1 * 0x1p-24) * 1.01);
  float4 bq[QB][NB];
  float nq[QB], al[QB], cq[QB], dmx[QB];
  int ql[QB];
  bool qv[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
    ql[qb] = w * QW + qb * 16 + si;
    const int j = qbase + ql[qb];
    qv[qb] = j < a.n;
    const int jj = qv[qb] ? j : a.n - 1;
    const float *xq = a.XB + (size_t)a.bmap[jj] * d;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int p = 16 * b + 4 * grp + i;
        v[i] = (qv[qb] && p < d) ? xq[p] : 0.0f;
      }
      bq[qb][b] = make_float4(v[0], v[1], v[2], v[3]);
    }
    nq[qb] = a.XB2[a.bmap[jj]];
    al[qb] = cd * a.xnB[jj];
    cq[qb] = 4.0f * u * fabsf(nq[qb]);
    dmx[qb] = (al[qb] * a.mx[0] + 4.0f * u * a.mx[1] + cq[qb]) * 1.002f + 1e-30f;
  }

  const int kg = (k + 3 - grp) / 4;
  int hu[QB][RK];
  auto insert_bits = [&](int qb, int x) __attribute__((always_inline)) {
#pragma unroll
    for (int i = RK - 1; i >= 1; --i) hu[qb][i] = max(hu[qb][i - 1], min(hu[qb][i], x));
    hu[qb][0] = min(hu[qb][0], x);
  };
  constexpr int kNegInf = (int)0xFF800000u, kPosInf = 0x7F800000;
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) {
#pragma unroll
    for (int i = 0; i < RK; ++i) hu[qb][i] = i < RK - kg ? kNegInf : kPosInf;
    if (qv[qb]) {
      const int j = qbase + ql[qb];
      const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
      for (int e = 0; e < kg; ++e) insert_bits(qb, max(__float_as_int(a.D[row + grp + 4 * e]), 0));
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
  int lc[QB];
#pragma unroll
  for (int qb = 0; qb < QB; ++qb) lc[qb] = 0;

  // Tile staging: wave w loads fragment blocks w BPW .. w BPW + BPW - 1
  // (dim block bidx / CB, candidate block bidx % CB); lane l reads dims
  // 16 j + 4 (l / 16) .. + 4 of candidate (bidx % CB) 16 + l % 16.
  auto tile_rows = [&](int t) { return (const char *)xg + (size_t)(cbeg + t * C) * DP * 4; };
  auto tile_recs = [&](int t) { return grec + cbeg + t * C; };
  unsigned doff[BPW];
#pragma unroll
  for (int e = 0; e < BPW; ++e) {
    const int bidx = w * BPW + e;
    doff[e] = (unsigned)(((bidx % CB) * 16 + si) * DP * 4 + ((bidx / CB) * 16 + 4 * grp) * 4);
  }
  const int ntiles = cend > cbeg ? (cend - cbeg + C - 1) / C : 0;
  const int T0 = min(a.warm, ntiles);
  const int nsteps = ntiles > 0 ? T0 + ntiles : 0;
  auto tile_of = [&](int v) { return v < T0 ? v : v - T0; };
  auto bufp = [&](int v) -> unsigned char * { return smem + (size_t)(v & 1) * L.buf; };
  float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
  float rc_x = 0.f, rc_n = 0.f;
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
  auto finish_tile = [&](int v) __attribute__((always_inline)) {
    unsigned char *b = bufp(v);
    const bool ok = cbeg + tile_of(v) * C + lane < cend;
    rc_n = ok ? st.x : __builtin_inff();
    rc_x = ok ? st.y : 0.0f;
    if (grp == w) ((float *)(b + L.nc))[lane] = rc_n;
  };
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
    const float4 *tA = (const float4 *)(b + L.tile);
    const float *tnc = (const float *)(b + L.nc);
    f32x4 acc[QB][CB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int jb = 0; jb < NB; ++jb) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const float4 av = tA[(jb * CB + cb) * 64 + lane];
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) {
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bq[qb][jb].x, acc[qb][cb], 0, 0, 0);
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bq[qb][jb].y, acc[qb][cb], 0, 0, 0);
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bq[qb][jb].z, acc[qb][cb], 0, 0, 0);
          acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bq[qb][jb].w, acc[qb][cb], 0, 0, 0);
        }
      }
    }
    float tq[QB][CB * 4];
    float tmin[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      tmin[qb] = __builtin_inff();
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = cb * 16 + 4 * grp + v;
          const float tt = __builtin_fmaf(-2.0f, acc[qb][cb][v], tnc[c]);
          tq[qb][cb * 4 + v] = tt;
          tmin[qb] = fminf(tmin[qb], tt);
        }
    }
    if (step < T0) {
      // Warm-up step: each lane's smallest tt of the tile bounds one
      // candidate's oracle distance from above by (tmin + nq) + dmx.
      int xb[QB];
      bool need = false;
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        const float s = (tmin[qb] + nq[qb]) + dmx[qb];
        xb[qb] = qv[qb] ? max(__float_as_int(widen_up(s)), 0) : kPosInf;
        need |= xb[qb] < hu[qb][RK - 1];
      }
      if (__any(need)) {
#pragma unroll
        for (int qb = 0; qb < QB; ++qb) insert_bits(qb, xb[qb]);
        refresh();
      }
    }
    // Main pass; over the warm-up tiles survivors are listed only.
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
        // Wave-uniform loop (the record shuffles need every lane active);
        // a lane without a survivor inserts +inf, which changes nothing.
        if (__any(bits != 0u)) do {
          const bool act = bits != 0u;
          const int bb = act ? __builtin_ctz(bits) : 0;
          bits &= bits - 1;
          const float tt = select_tree<0, CB * 4, CB * 4>(tq[qb], bb);
          const int c = (bb >> 2) * 16 + 4 * grp + (bb & 3);
          const float rx = __shfl(rc_x, c), rn = __shfl(rc_n, c);
          const int pos = cbeg + t * C + c;
          const bool valid = act && pos < cend;
          const float dist = tt + nq[qb];
          const float delta = rx * al[qb] + 4.0f * u * fabsf(rn) + cq[qb] + 1e-30f;
          const float lower = widen_dn(dist - delta), upper = widen_up(dist + delta);
          const bool app = valid && lower <= uq[qb];
          const int at = lc[qb];
          lc[qb] = at + (app ? 1 : 0);
          if (app && at < capg) a.lst[lb + at] = make_int2(pos, __float_as_int(lower));
          const int ux = max(__float_as_int(upper), valid && ins ? 0 : kPosInf);
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

// The oracle's FP32 distance for candidate position pos and query j.
__device__ __forceinline__ float oracle_dist_f32(const F32Args &a, int pos, int j) {
  const float *xa = a.XA + (size_t)a.amap[pos] * a.d;
  const float *xb = a.XB + (size_t)a.bmap[j] * a.d;
  float s = 0.0f;
  for (int p = 0; p < a.d; ++p) s = __fadd_rn(s, __fmul_rn(xa[p], xb[p]));
  s = __fmul_rn(s, -2.0f);
  s = __fadd_rn(s, a.XA2[a.amap[pos]]);
  return __fadd_rn(s, a.XB2[a.bmap[j]]);
}

// Pass 2: one wave per query.  Lanes take 64 listed candidates at a time,
// compute their oracle distances independently, then lane 0 merges the ones
// that beat the current root into the LDS heap in lane order.
__global__ __launch_bounds__(256) void knn_f32_rerank_kernel(F32Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  float *hd = (float *)smem + (size_t)w * k;
  int *hi = (int *)(smem + a16(sizeof(float) * 4 * (size_t)k)) + (size_t)w * k;
  if (j >= a.n) return;
  const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
  if (lane == 0) {
    for (int t = 0; t < k; ++t) {
      hd[t] = a.D[row + t];
      hi[t] = a.I[row + t];
    }
    for (int s = k / 2 - 1; s >= 0; --s) sift_fi(hd, hi, s, k, hd[s], hi[s]);
  }
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  float Ulim = a.U[j];
  for (int sp = 1; sp < a.nsplit; ++sp) Ulim = fminf(Ulim, a.U[(size_t)sp * a.n + j]);
  const int capg = a.cap / a.G;
  for (int sp = 0; sp < a.nsplit; ++sp) {
    const size_t sj = (size_t)sp * a.n + j;
    bool over = false;
    for (int gg = 0; gg < a.G; ++gg) over |= a.cnt[((size_t)sp * a.G + gg) * a.n + j] > capg;
    const int cbeg = sp * a.cps;
    // An overflowed split is re-scanned once; otherwise each of its lists.
    for (int gg = 0; gg < (over ? 1 : a.G); ++gg) {
    const int nscan = over ? min(a.m, cbeg + a.cps) - cbeg : a.cnt[((size_t)sp * a.G + gg) * a.n + j];
    const size_t lbase = sj * a.cap + (size_t)gg * capg;
    for (int e0 = 0; e0 < nscan; e0 += 64) {
      const int e = e0 + lane;
      bool have = false;
      int pos = 0;
      if (e < nscan) {
        if (over) {
          pos = cbeg + e;
          have = true;
        } else {
          const int2 en = a.lst[lbase + e];
          pos = en.x;
          have = __int_as_float(en.y) <= Ulim;
        }
      }
      float dist = 0.0f;
      int id = 0;
      if (have) {
        dist = oracle_dist_f32(a, pos, j);
        id = a.amap[pos];
      }
      // Merge in lane order; only lanes that may beat the root take a turn.
      unsigned long long todo = __ballot(have && cl_f(dist, id, hd[0], hi[0]));
      while (todo) {
        const int src = __builtin_ctzll(todo);
        todo &= todo - 1;
        const float xd = __shfl(dist, src, 64);
        const int xi = __shfl(id, src, 64);
        if (lane == 0 && cl_f(xd, xi, hd[0], hi[0])) sift_fi(hd, hi, 0, k, xd, xi);
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      }
    }
    }
  }
  if (lane == 0) {
    for (int e2 = k - 1; e2 > 0; --e2) {
      const float xd = hd[e2];
      const int xi = hi[e2];
      hd[e2] = hd[0];
      hi[e2] = hi[0];
      sift_fi(hd, hi, 0, e2, xd, xi);
    }
    for (int t = 0; t < k; ++t) {
      a.D[row + t] = hd[k - 1 - t];
      a.I[row + t] = hi[k - 1 - t];
    }
  }
}

// The oracle's FP32 distance of candidate row xa and query row xb (p in
// order from 0, every product and sum rounded), with the loads of 32 dims of
// both rows issued together when the rows are 16-byte aligned (VEC).
__device__ __forceinline__ float oracle_dist_f32v(const float *xa, const float *xb, int d, float xa2,
                                                  float xb2, bool vec) {
  float s = 0.0f;
  int p = 0;
  if (vec) {
#pragma unroll 2
    for (; p + 16 <= d; p += 16) {
      f32x4 ca[4], cb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ca[u] = *(const f32x4 *)(xa + p + 4 * u);
        cb[u] = *(const f32x4 *)(xb + p + 4 * u);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s = __fadd_rn(s, __fmul_rn(ca[u].x, cb[u].x));
        s = __fadd_rn(s, __fmul_rn(ca[u].y, cb[u].y));
        s = __fadd_rn(s, __fmul_rn(ca[u].z, cb[u].z));
        s = __fadd_rn(s, __fmul_rn(ca[u].w, cb[u].w));
      }
    }
  }
  for (; p < d; ++p) s = __fadd_rn(s, __fmul_rn(xa[p], xb[p]));
  s = __fmul_rn(s, -2.0f);
  s = __fadd_rn(s, xa2);
  return __fadd_rn(s, xb2);
}

// Pass 2, selection form (k <= 64 (NS - 2)): one wave per query.  A pool in
// LDS holds the resident row and every recomputed candidate that beats the
// pool's k-th entry (the resident row's largest at first); when a batch of 64
// would not fit, a rank selection keeps the k smallest under the canonical
// (distance, index) order and gives the new k-th entry.  The last selection
// writes the row in descending order (a valid max-heap, as the heap form
// leaves it).  No lane runs heap operations alone.
template <int NS>
__global__ __launch_bounds__(256) void knn_f32_rerank_select_kernel(F32Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int P = 64 * NS;
  const int k = a.k, d = a.d;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  float *pd = (float *)smem + (size_t)w * P;
  int *pi = (int *)(smem + sizeof(float) * 4 * P) + (size_t)w * P;
  float *rootd = (float *)(smem + sizeof(float) * 8 * P) + w;
  int *rooti = (int *)(smem + sizeof(float) * 8 * P + 16) + w;
  if (j >= a.n) return;  // the whole wave leaves; no block barrier below
  auto wsync = []() __attribute__((always_inline)) {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  };
  const size_t row = (size_t)(a.hrow ? a.hrow[j] : j) * a.ldk;
  float rd = -__builtin_inff();
  int ri = INT_MIN;
  for (int t = lane; t < k; t += 64) {
    const float x = a.D[row + t];
    const int xi = a.I[row + t];
    pd[t] = x;
    pi[t] = xi;
    if (cl_f(rd, ri, x, xi)) {
      rd = x;
      ri = xi;
    }
  }
  // The resident row's largest entry (no candidate at or above it can enter).
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float od = __shfl_xor(rd, off);
    const int oi = __shfl_xor(ri, off);
    if (cl_f(rd, ri, od, oi)) {
      rd = od;
      ri = oi;
    }
  }
  int np = k;
  // Keep the k smallest pool entries; the last call writes the row.
  auto select = [&](bool last) __attribute__((always_inline)) {
    wsync();
    float od[NS];
    int oi[NS], orank[NS];
    const int ns = (np + 63) >> 6;
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const int e = lane + 64 * s2;
      od[s2] = e < np ? pd[e] : 0.0f;
      oi[s2] = e < np ? pi[e] : 0;
      orank[s2] = 0;
    }
    for (int jj = 0; jj < np; ++jj) {
      const float dj = pd[jj];
      const int ij = pi[jj];
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        if (s2 < ns) {
          const int e = lane + 64 * s2;
          const bool lt = dj < od[s2] || (dj == od[s2] && (ij < oi[s2] || (ij == oi[s2] && jj < e)));
          orank[s2] += lt ? 1 : 0;
        }
      }
    }
    wsync();  // every rank taken before the pool is rewritten
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      const int e = lane + 64 * s2;
      if (e < np && orank[s2] < k) {
        if (last) {
          a.D[row + k - 1 - orank[s2]] = od[s2];
          a.I[row + k - 1 - orank[s2]] = oi[s2];
        } else {
          pd[orank[s2]] = od[s2];
          pi[orank[s2]] = oi[s2];
          if (orank[s2] == k - 1) {
            *rootd = od[s2];
            *rooti = oi[s2];
          }
        }
      }
    }
    wsync();
    if (!last) {
      rd = *rootd;
      ri = *rooti;
      np = k;
    }
  };

  const bool vec = (d & 3) == 0 && ((uintptr_t)a.XA & 15) == 0 && ((uintptr_t)a.XB & 15) == 0;
  const int qid = a.bmap[j];
  const float *xb = a.XB + (size_t)qid * d;
  const float xb2 = a.XB2[qid];
  float Ulim = a.U[j];
  for (int sp = 1; sp < a.nsplit; ++sp) Ulim = fminf(Ulim, a.U[(size_t)sp * a.n + j]);
  const int capg = a.cap / a.G;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int sp = 0; sp < a.nsplit; ++sp) {
    const size_t sj = (size_t)sp * a.n + j;
    bool over = false;
    for (int gg = 0; gg < a.G; ++gg) over |= a.cnt[((size_t)sp * a.G + gg) * a.n + j] > capg;
    const int cbeg = sp * a.cps;
    // An overflowed split is re-scanned once; otherwise each of its lists.
    for (int gg = 0; gg < (over ? 1 : a.G); ++gg) {
      const int nscan = over ? min(a.m, cbeg + a.cps) - cbeg : a.cnt[((size_t)sp * a.G + gg) * a.n + j];
      const size_t lbase = sj * a.cap + (size_t)gg * capg;
      for (int e0 = 0; e0 < nscan; e0 += 64) {
        const int e = e0 + lane;
        bool have = false;
        int pos = 0;
        if (e < nscan) {
          if (over) {
            pos = cbeg + e;
            have = true;
          } else {
            const int2 en = a.lst[lbase + e];
            pos = en.x;
            have = __int_as_float(en.y) <= Ulim;
          }
        }
        float dist = 0.0f;
        int id = 0;
        if (have) {
          id = a.amap[pos];
          dist = oracle_dist_f32v(a.XA + (size_t)id * d, xb, d, a.XA2[id], xb2, vec);
        }
        bool keep = have && cl_f(dist, id, rd, ri);
        unsigned long long km = __ballot(keep);
        if (np + __popcll(km) > P) {
          select(false);
          keep = keep && cl_f(dist, id, rd, ri);
          km = __ballot(keep);
        }
        if (keep) {
          const int slot = np + __popcll(km & below);
          pd[slot] = dist;
          pi[slot] = id;
        }
        np += __popcll(km);
      }
    }
  }
  select(true);
}

namespace {

constexpr size_t kLdsF = 160 * 1024;

// GSKNN_F32_SELECT=0 keeps the heap-form re-rank for every k (read at each
// call; tuning only).
bool f32_select_on() {
  const char *e = getenv("GSKNN_F32_SELECT");
  return !(e && atoi(e) == 0);
}
// Register form: upper-bound set size per lane, two query blocks per wave
// for k <= 4 kF32RegRK, one (larger sets) for 4 kF32RegRK < k <= 4 kF32RegRK1.
constexpr int kF32RegRK = 16;
constexpr int kF32RegRK1 = 32;

struct F32Plan {
  int ks = 0, qb = 0, cb = 0, w = 0, q = 0, cap = 0, nsplit = 1, cps = 0;
  size_t lds = 0;
  bool reg = false;  // register form (knn_f32_filter2_kernel)
  int nb = 0;        // register form: 16-dim blocks (DP = 16 nb)
  int G = 1;         // list segments per (split, query)
};

// Split the candidates when the query blocks alone give fewer than two
// workgroups per CU; each split keeps at least 16 tiles.
void plan_f32_splits(F32Plan &p, int m, int n) {
  const int gx = (n + p.q - 1) / p.q;
  const int C = 16 * p.cb;
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

F32Plan plan_f32(int d, int k) {
  F32Plan p;
  const int need = (d + 3) / 4;
  for (int ks : {4, 8, 16, 32, 64}) {
    if (need <= ks) {
      p.ks = ks;
      break;
    }
  }
  if (p.ks == 0) return p;
  p.cb = p.ks >= 32 ? 1 : 2;
  // At d > 128 four waves of QB = 2 exceed 256 registers each (the query
  // fragments alone take 128), so 8-wave workgroups, which stage half as
  // much per thread, come first there (measured at d = 256: 114 ms against
  // 166 ms for 4 waves at one wave per SIMD).
  const int opts[5][2] = {{2, 8}, {2, 4}, {1, 4}, {1, 2}, {1, 1}};
  for (auto &o : opts) {
    if (o[1] == 8 && p.ks < 64) continue;
    const int Q = o[1] * 16 * o[0];
    const size_t b = F32Layout(p.ks, p.cb, Q, k).bytes;
    if (b <= kLdsF) {
      p.qb = o[0];
      p.w = o[1];
      p.q = Q;
      p.lds = b;
      break;
    }
  }
  const int cap = 16 * k;
  p.cap = cap < 512 ? 512 : (cap > 4096 ? 4096 : cap);
  // Register form for d <= 256 and k <= 4 kF32RegRK1 (GSKNN_F32_REG=0 keeps
  // the LDS-tile form, for comparison); its lists get twice the capacity,
  // in four lane-group segments.
  static const bool reg_on = [] {
    const char *e = getenv("GSKNN_F32_REG");
    return !(e && atoi(e) == 0);
  }();
  // Above d = 128 the 64 KiB tiles (two buffers) leave one workgroup per
  // CU, so it takes eight waves (two per SIMD) sharing each tile: at
  // config-3 size (d = 256, m = n = 131072, k = 64) 81 ms against 110 for the
  // LDS-tile form (four waves: 110.5).
  if (reg_on && d <= 256 && k <= 4 * kF32RegRK1) {
    p.reg = true;
    p.nb = d <= 16 ? 1 : d <= 32 ? 2 : d <= 64 ? 4 : d <= 128 ? 8 : 16;
    p.qb = k <= 4 * kF32RegRK ? 2 : 1;
    p.cb = 4;
    p.w = p.nb == 16 ? 8 : 4;
    p.q = p.w * 16 * p.qb;
    p.lds = F32RegLayout(p.nb, 4).bytes;
    p.G = 4;
    const int c2 = 32 * k;
    p.cap = c2 < 1024 ? 1024 : (c2 > 8192 ? 8192 : c2);
  }
  return p;
}

struct F32Work {
  size_t xnA, xnB, mx, lst, cnt, U, nover, panel, bytes;
  F32Work(int m, int n, const F32Plan &p) {
    const size_t rows = (size_t)n * p.nsplit;
    xnA = 0;
    xnB = a16(xnA + (size_t)m * 4);
    mx = a16(xnB + (size_t)n * 4);
    lst = a16(mx + 16);
    cnt = a16(lst + rows * p.cap * 8);
    U = a16(cnt + rows * p.G * 4);
    nover = a16(U + rows * 4);
    panel = a16(nover + 16);
    bytes = a16(panel + (p.reg ? F32Panel(m, 16 * p.nb).bytes : 0));
  }
};

template <int KS>
hipError_t launch_f32_ks(const F32Plan &p, const F32Args &a, hipStream_t s) {
  const dim3 grid((a.n + p.q - 1) / p.q, a.nsplit);
  constexpr int CBk = KS >= 32 ? 1 : 2;
#define GSKNN_F32_CASE(QB_, W_)                                                        \
  if (p.qb == QB_ && p.w == W_) {                                                      \
    auto kern = knn_f32_filter_kernel<KS, QB_, CBk, W_>;                                \
    hipError_t e = hipFuncSetAttribute((const void *)kern,                              \
                                       hipFuncAttributeMaxDynamicSharedMemorySize,      \
                                       (int)p.lds);                                     \
    if (e != hipSuccess) return e;                                                      \
    hipLaunchKernelGGL(kern, grid, dim3(64 * W_), p.lds, s, a);                         \
    return hipGetLastError();                                                           \
  }
  GSKNN_F32_CASE(2, 8)
  GSKNN_F32_CASE(2, 4)
  GSKNN_F32_CASE(1, 4)
  GSKNN_F32_CASE(1, 2)
  GSKNN_F32_CASE(1, 1)
#undef GSKNN_F32_CASE
  return hipErrorNotSupported;
}

}  // namespace

size_t f32_work_bytes(int m, int n, int d, int k) {
  F32Plan p = plan_f32(d, k);
  if (p.q == 0 || m <= 0 || n <= 0) return 0;
  plan_f32_splits(p, m, n);
  return F32Work(m, n, p).bytes;
}

size_t f32_overflow_offset(int m, int n, int d, int k) {
  F32Plan p = plan_f32(d, k);
  if (p.q == 0 || m <= 0 || n <= 0) return (size_t)-1;
  plan_f32_splits(p, m, n);
  return F32Work(m, n, p).nover;
}

bool f32_mfma_supported(int d, int k) {
  const F32Plan p = plan_f32(d, k);
  const size_t rl = a16(sizeof(float) * 4 * (size_t)k) + sizeof(int) * 4 * (size_t)k;
  return p.q > 0 && rl <= kLdsF;
}

hipError_t launch_knn_f32_mfma(const KnnArgs<float> &ka, hipStream_t s, void *work,
                               size_t work_bytes, int *overflow_out) {
  F32Plan p = plan_f32(ka.d, ka.k);
  if (p.q == 0) return hipErrorNotSupported;
  plan_f32_splits(p, ka.m, ka.n);
  const size_t rl = a16(sizeof(float) * 4 * (size_t)ka.k) + sizeof(int) * 4 * (size_t)ka.k;
  if (rl > kLdsF) return hipErrorNotSupported;
  const F32Work W(ka.m, ka.n, p);
  if (work == nullptr || work_bytes < W.bytes) return hipErrorInvalidValue;
  char *wb = (char *)work;
  F32Args a{};
  a.m = ka.m; a.n = ka.n; a.d = ka.d; a.k = ka.k; a.ldk = ka.ldk; a.cap = p.cap;
  a.nsplit = p.nsplit; a.cps = p.cps;
  a.XA = ka.XA; a.XA2 = ka.XA2; a.amap = ka.amap;
  a.XB = ka.XB; a.XB2 = ka.XB2; a.bmap = ka.bmap;
  a.D = ka.D; a.I = ka.I; a.hrow = ka.hrow;
  a.xnA = (float *)(wb + W.xnA);
  a.xnB = (float *)(wb + W.xnB);
  a.mx = (float *)(wb + W.mx);
  a.lst = (int2 *)(wb + W.lst);
  a.cnt = (int *)(wb + W.cnt);
  a.U = (float *)(wb + W.U);
  a.nover = (int *)(wb + W.nover);
  a.G = p.G;
  a.panel = (const unsigned char *)(wb + W.panel);
  // Warm-up pass of the register form: the first 1/16 of each split's tiles,
  // at most 32 (GSKNN_F32_WARM=<tiles>, read at each call, sets the count).
  // Longer passes cost more MFMA work than the list entries they save: at
  // m = n = 65536, d = 128, k = 32, 0 / 32 / 64 / 128 tiles took 10.71 /
  // 10.71 / 10.92 / 11.42 ms, and serving over 2M candidates (about 290 tiles
  // per split at 1/16) ran p50 216 ms with 32 against 271.
  const char *warm_e = getenv("GSKNN_F32_WARM");
  a.warm = p.reg ? (warm_e ? atoi(warm_e) : min(p.cps / 64 / 16, 32)) : 0;
  hipLaunchKernelGGL(f32_zero_kernel, dim3(1), dim3(64), 0, s, a.mx, a.nover);
  hipLaunchKernelGGL(f32_prep_kernel, dim3((a.m + a.n + 255) / 256), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (p.reg) {
    hipLaunchKernelGGL(f32_gather_kernel, dim3((a.m + kF32Pad + 3) / 4), dim3(256), 0, s, a, 16 * p.nb,
                       (unsigned char *)(wb + W.panel));
    const dim3 grid((a.n + p.q - 1) / p.q, a.nsplit);
    void (*kern)(F32Args) = nullptr;
    switch (p.nb * 4 + p.qb) {
      case 1 * 4 + 2: kern = knn_f32_filter2_kernel<1, 2, 4, 4, kF32RegRK>; break;
      case 2 * 4 + 2: kern = knn_f32_filter2_kernel<2, 2, 4, 4, kF32RegRK>; break;
      case 4 * 4 + 2: kern = knn_f32_filter2_kernel<4, 2, 4, 4, kF32RegRK>; break;
      case 8 * 4 + 2: kern = knn_f32_filter2_kernel<8, 2, 4, 4, kF32RegRK>; break;
      case 1 * 4 + 1: kern = knn_f32_filter2_kernel<1, 1, 4, 4, kF32RegRK1>; break;
      case 2 * 4 + 1: kern = knn_f32_filter2_kernel<2, 1, 4, 4, kF32RegRK1>; break;
      case 4 * 4 + 1: kern = knn_f32_filter2_kernel<4, 1, 4, 4, kF32RegRK1>; break;
      case 8 * 4 + 1: kern = knn_f32_filter2_kernel<8, 1, 4, 4, kF32RegRK1>; break;
      case 16 * 4 + 2: kern = knn_f32_filter2_kernel<16, 2, 4, 8, kF32RegRK>; break;
      case 16 * 4 + 1: kern = knn_f32_filter2_kernel<16, 1, 4, 8, kF32RegRK1>; break;
      default: return hipErrorNotSupported;
    }
    e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, grid, dim3(64 * p.w), p.lds, s, a);
    e = hipGetLastError();
  } else switch (p.ks) {
    case 4: e = launch_f32_ks<4>(p, a, s); break;
    case 8: e = launch_f32_ks<8>(p, a, s); break;
    case 16: e = launch_f32_ks<16>(p, a, s); break;
    case 32: e = launch_f32_ks<32>(p, a, s); break;
    case 64: e = launch_f32_ks<64>(p, a, s); break;
    default: return hipErrorNotSupported;
  }
  if (e != hipSuccess) return e;
  const int sel_ns = (a.k + 63) / 64 + 2;
  const void *rk = (const void *)knn_f32_rerank_kernel;
  size_t rlk = rl;
  if (sel_ns <= 6 && f32_select_on()) {
    const void *sk[4] = {(const void *)knn_f32_rerank_select_kernel<3>,
                         (const void *)knn_f32_rerank_select_kernel<4>,
                         (const void *)knn_f32_rerank_select_kernel<5>,
                         (const void *)knn_f32_rerank_select_kernel<6>};
    rk = sk[sel_ns - 3];
    rlk = sizeof(float) * 8 * 64 * (size_t)sel_ns + 32;
  }
  e = hipFuncSetAttribute(rk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rlk);
  if (e != hipSuccess) return e;
  void *kargs[] = {(void *)&a};
  e = hipLaunchKernel(rk, dim3((a.n + 3) / 4), dim3(256), kargs, rlk, s);
  if (e != hipSuccess) return e;
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (overflow_out) e = hipMemcpyAsync(overflow_out, a.nover, sizeof(int), hipMemcpyDeviceToDevice, s);
  return e;
}

}  // namespace gsknn
