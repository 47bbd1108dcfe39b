// Batched all-nearest-neighbour sweep over the leaves of a randomized-
// projection tree: one persistent launch per tree iteration.
//
// No counterpart in the reference; each leaf is the reference's gsknn call
// with amap = bmap = the leaf's point ids and heap rows = the same ids
// (merging into global heap rows, frame/gsknn.c:1138-1184 semantics), so the
// result equals the composition of per-leaf dgsknn_ref/sgsknn_ref calls.
//
// Work list: items[i] = {leaf start in perm, leaf size, first query of the
// block, 0}; the item count is read from device memory so that a launch
// captured once in a hipGraph serves every iteration.  Each workgroup walks
// the list with stride gridDim.x (items sorted by leaf size, largest first,
// balance the persistent grid).  Within an item the exact block routine runs
// one thread per query with the oracle's arithmetic (the scalar-candidate
// form for d in {8, 16, 32, 64}), so distances are bit-identical to the
// oracle's.  Leaves of one iteration are disjoint, so no
// heap row is touched by two items of the same launch.
#include "gsknn_device.h"

namespace gsknn {

// D > 0: d == D in {8, 16, 32, 64}, the scalar-candidate block routine
// (CT unused); D == 0: the LDS-tile routine with DQ.
template <typename T, int QPB, int DQ, int D>
__global__ __launch_bounds__(QPB) void knn_leaf_sweep_kernel(KnnArgs<T> base, const int4 *items,
                                                             const int *nitems, int CT) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int count = *nitems;
  for (int it = blockIdx.x; it < count; it += gridDim.x) {
    const int4 w = items[it];
    KnnArgs<T> a = base;
    a.m = w.y;
    a.n = w.y;
    a.amap = base.amap + w.x;
    a.bmap = base.amap + w.x;
    a.hrow = base.amap + w.x;
    a.nsplit = 1;
    a.cps = w.y;
    if constexpr (D > 0)
      exact_block_s<T, QPB, D>(a, w.z / QPB, 0, smem);
    else
      exact_block<T, QPB, DQ>(a, CT, w.z / QPB, 0, smem);
    __syncthreads();  // LDS is reused by the next item
  }
}

// One level of the randomized-projection tree: for element i of the nodes
// being split (node n holds perm[starts[n], starts[n] + sizes[n]) and its
// elements are numbered from offs[n]), the position in perm and the sort key
// (n, projection order, point id); sorting the keys orders every node by
// (projection, id), ties by id, so the median split is deterministic.  The
// projection is sum_p x_p dir_p in p order, products and sums rounded
// separately; its FP32 bits map to an unsigned order (negative values
// reversed).  Blocks of 256 consecutive elements: one thread finds the node
// of the block's first element by binary search and the others step forward
// from it (a block spans few nodes); rows are read 16 bytes at a time when d
// is a multiple of 4 and the tables are 16-byte aligned.
__global__ __launch_bounds__(256) void tree_level_keys_kernel(
    long long total, int nnodes, int d, int id_bits, const float *X, const long long *perm,
    const long long *starts, const long long *offs, const float *dirs, long long *pos_out,
    long long *key_out) {
  __shared__ int base_node;
  const long long i0 = (long long)blockIdx.x * blockDim.x;
  const long long i = i0 + threadIdx.x;
  if (threadIdx.x == 0) {
    int lo = 0, hi = nnodes - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offs[mid] <= i0) lo = mid;
      else hi = mid - 1;
    }
    base_node = lo;
  }
  __syncthreads();
  if (i >= total) return;
  int lo = base_node;
  while (lo + 1 < nnodes && offs[lo + 1] <= i) ++lo;
  const long long pos = starts[lo] + (i - offs[lo]);
  const long long pt = perm[pos];
  const float *x = X + pt * d;
  const float *dv = dirs + (size_t)lo * d;
  float s = 0.0f;
  if ((d & 3) == 0 && (((uintptr_t)X | (uintptr_t)dirs) & 15) == 0) {
    const float4 *x4 = (const float4 *)x;
    const float4 *d4 = (const float4 *)dv;
    for (int p = 0; p < d / 4; ++p) {
      const float4 a = x4[p], w = d4[p];
      s = __fadd_rn(s, __fmul_rn(a.x, w.x));
      s = __fadd_rn(s, __fmul_rn(a.y, w.y));
      s = __fadd_rn(s, __fmul_rn(a.z, w.z));
      s = __fadd_rn(s, __fmul_rn(a.w, w.w));
    }
  } else {
    for (int p = 0; p < d; ++p) s = __fadd_rn(s, __fmul_rn(x[p], dv[p]));
  }
  int b = __float_as_int(s);
  if (b < 0) b ^= 0x7FFFFFFF;
  const long long ord = (long long)b + (1LL << 31);
  pos_out[i] = pos;
  key_out[i] = ((long long)lo << (32 + id_bits)) | (ord << id_bits) | pt;
}

hipError_t launch_tree_level_keys(long long total, int nnodes, int d, int id_bits,
                                  const float *X, const long long *perm,
                                  const long long *starts, const long long *offs,
                                  const float *dirs, long long *pos, long long *keys,
                                  hipStream_t s) {
  if (total <= 0) return hipSuccess;
  const long long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(tree_level_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, s, total,
                     nnodes, d, id_bits, X, perm, starts, offs, dirs, pos, keys);
  return hipGetLastError();
}

template <typename T>
size_t leaf_sweep_lds(int d, int k, int qpb, int ct) {
  if (scalar_dim(d)) return ExactScalarLayout<T>(k, qpb).bytes;
  return ExactLayout<T>(k, qpb, d, ct, d <= kExactDQ).bytes;
}

template <typename T>
hipError_t launch_leaf_sweep_t(int d, int k, const T *X, const T *X2, const int *perm,
                               const int *items, const int *nitems, T *D, int *I, int ldk,
                               int qpb, int ct, int grid, hipStream_t s) {
  KnnArgs<T> a{};
  a.d = d; a.k = k; a.ldk = ldk;
  a.XA = X; a.XA2 = X2; a.XB = X; a.XB2 = X2;
  a.amap = perm; a.bmap = perm;
  a.D = D; a.I = I;
  a.nsplit = 1;
  const size_t lds = leaf_sweep_lds<T>(d, k, qpb, ct);
  if (lds > 160 * 1024 || ct < 4 || (ct & 3) || grid < 1) return hipErrorNotSupported;
  hipError_t e = hipSuccess;
#define GSKNN_LEAF_CASE(QPB_)                                                             \
  if (qpb == QPB_) {                                                                      \
    auto kern = d == 8    ? knn_leaf_sweep_kernel<T, QPB_, 0, 8>                          \
                : d == 16 ? knn_leaf_sweep_kernel<T, QPB_, 0, 16>                         \
                : d == 32 ? knn_leaf_sweep_kernel<T, QPB_, 0, 32>                         \
                : d == 64 ? knn_leaf_sweep_kernel<T, QPB_, 0, 64>                         \
                : d <= kExactDQ ? knn_leaf_sweep_kernel<T, QPB_, kExactDQ, 0>             \
                                : knn_leaf_sweep_kernel<T, QPB_, 0, 0>;                   \
    e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            (int)lds);                                                    \
    if (e != hipSuccess) return e;                                                        \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(QPB_), lds, s, a, (const int4 *)items, nitems, ct); \
    return hipGetLastError();                                                             \
  }
  GSKNN_LEAF_CASE(64)
  GSKNN_LEAF_CASE(32)
  GSKNN_LEAF_CASE(16)
#undef GSKNN_LEAF_CASE
  return hipErrorNotSupported;
}

hipError_t launch_leaf_sweep_f32(int d, int k, const float *X, const float *X2, const int *perm,
                                 const int *items, const int *nitems, float *D, int *I, int ldk,
                                 int qpb, int ct, int grid, hipStream_t s) {
  return launch_leaf_sweep_t<float>(d, k, X, X2, perm, items, nitems, D, I, ldk, qpb, ct, grid, s);
}

hipError_t launch_leaf_sweep_f64(int d, int k, const double *X, const double *X2, const int *perm,
                                 const int *items, const int *nitems, double *D, int *I, int ldk,
                                 int qpb, int ct, int grid, hipStream_t s) {
  return launch_leaf_sweep_t<double>(d, k, X, X2, perm, items, nitems, D, I, ldk, qpb, ct, grid, s);
}

size_t leaf_sweep_lds_bytes(int elem, int d, int k, int qpb, int ct) {
  return elem == 8 ? leaf_sweep_lds<double>(d, k, qpb, ct) : leaf_sweep_lds<float>(d, k, qpb, ct);
}


// ---------------------------------------------------------------------------
// FP32 MFMA form of the leaf sweep (d = 4 KS in {16, 32}, k <= kSweepMaxK): one
// wave per workgroup walks the work list; an item's Q = 16 QB queries (the
// leaf's points q0 .. q0 + Q - 1) stream the leaf's candidates in tiles of
// C = 16 CB through v_mfma_f32_16x16x4f32.  Lane group g carries dims
// g KS .. g KS + KS - 1 of its candidate and query in k-steps 0 .. KS - 1
// (A and B agree, so the sum is the dot product), so each lane reads KS
// contiguous floats of a row; the next tile's rows are read during the
// current tile's MFMAs.
//
// Exactness (as in the FP32 MFMA form of sgsknn, gsknn_f32.hip): the MFMA
// value t = fl(-2 dot~ + XA2[c]) and the oracle's distance d_or (dot summed
// in p order, product and sum rounded separately, x(-2), + XA2[c], + XB2[q])
// satisfy |t + XB2[q] - d_or| <= delta = c_d |a||b| + 4u (|XA2[c]| + |XB2[q]|),
// c_d = 1.01 (2 g_m + 2 g_o + 8.1 u), g_m = (2d+2) 2u / (1 - (2d+2) 2u),
// g_o = d u / (1 - d u), with |a|, |b| the points' Euclidean norms rounded up
// (xn).  A pair passes when t <= root - XB2[q] + dmx (dmx bounds delta over
// the table, plus 0.2 %), which every pair whose oracle distance can reach the
// row's root (the k-th entry under (distance, index)) satisfies.  Passing
// pairs wait in a per-lane queue of candidate ids; when a queue would
// overflow, every lane forms the oracle distances of its queued pairs (the
// candidate row from memory, the query row from LDS) and the four lane
// groups (each holding the wave's queries once) insert them into the LDS
// heaps in turn.  The heaps hold only oracle distances, so every row equals
// the composition of per-leaf oracle calls bit for bit.
constexpr int kSweepMaxK = 64;
constexpr int kSweepQR = 4;

struct SweepMfmaLayout {
  size_t hd, hi, qr, bytes;
  __host__ __device__ SweepMfmaLayout(int k, int Q, int d) {
    hd = 0;
    hi = al16(sizeof(float) * (size_t)k * Q);
    qr = al16(hi + sizeof(int) * (size_t)k * Q);
    bytes = al16(qr + sizeof(float) * (size_t)Q * d);
  }
};

struct SweepMfmaArgs {
  int d, k, ldk;
  const float *X, *X2;
  const float *xn;  // |X[i]| rounded up
  const float *mx;  // {max xn, max |X2|}
  const int *perm;
  float *D;
  int *I;
};

// The oracle's FP32 distance of candidate row xa to query row xq (LDS).
template <int DD>
__device__ __forceinline__ float sweep_oracle_dist(const float *xa, const float *xq, float nc,
                                                   float nq) {
  float v[DD];
#pragma unroll
  for (int p = 0; p < DD; p += 4) {
    const float4 f = *(const float4 *)(xa + p);
    v[p] = f.x;
    v[p + 1] = f.y;
    v[p + 2] = f.z;
    v[p + 3] = f.w;
  }
  float s = 0.0f;
#pragma unroll
  for (int p = 0; p < DD; ++p) s = xadd(s, xmul(v[p], xq[p]));
  return xadd(xadd(xmul(s, -2.0f), nc), nq);
}

template <int KS, int QB, int CB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void knn_leaf_sweep_mfma_kernel(SweepMfmaArgs a, const int4 *items,
                                                                 const int *nitems) {
  constexpr int Q = 16 * QB, C = 16 * CB, DD = 4 * KS, QR = kSweepQR;
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = a.k;
  const SweepMfmaLayout L(k, Q, DD);
  float *hd = (float *)(smem + L.hd);
  int *hi = (int *)(smem + L.hi);
  float *qrow = (float *)(smem + L.qr);
  const int lane = threadIdx.x, grp = lane >> 4, cl = lane & 15;

  const float u = 0x1p-24f;
  const double gmd = (2.0 * DD + 2) * 2.0 * 0x1p-24 / (1.0 - (2.0 * DD + 2) * 2.0 * 0x1p-24);
  const double god = DD * 0x1p-24 / (1.0 - DD * 0x1p-24);
  const float cdf = (float)((2.0 * gmd + 2.0 * god + 8.1 * 0x1p-24) * 1.01);
  const float mx0 = a.mx[0], mx1 = a.mx[1];
  const int count = *nitems;

  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int4 w = items[item];
    const int lstart = w.x, L_ = w.y, q0 = w.z;
    const int *leaf = a.perm + lstart;

    // Queries: fragments in registers, rows in LDS, heaps in LDS.
    float bq[QB][KS];
    float nq[QB], al[QB], cq[QB], dmx[QB];
    int ql[QB];
    bool qv[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
      ql[qb] = qb * 16 + cl;
      qv[qb] = q0 + ql[qb] < L_;
      const int pq = leaf[qv[qb] ? q0 + ql[qb] : L_ - 1];
      const float *xq = a.X + (size_t)pq * DD + grp * KS;
#pragma unroll
      for (int s = 0; s < KS; s += 4) {
        const float4 f = *(const float4 *)(xq + s);
        bq[qb][s] = qv[qb] ? f.x : 0.f;
        bq[qb][s + 1] = qv[qb] ? f.y : 0.f;
        bq[qb][s + 2] = qv[qb] ? f.z : 0.f;
        bq[qb][s + 3] = qv[qb] ? f.w : 0.f;
      }
#pragma unroll
      for (int s = 0; s < KS; ++s) qrow[ql[qb] * DD + grp * KS + s] = bq[qb][s];
      nq[qb] = a.X2[pq];
      al[qb] = cdf * a.xn[pq];
      cq[qb] = 4.0f * u * fabsf(nq[qb]);
      dmx[qb] = (al[qb] * mx0 + 4.0f * u * mx1 + cq[qb]) * 1.002f + 1e-30f;
    }
    if (lane < Q && q0 + lane < L_) {
      const size_t row = (size_t)leaf[q0 + lane] * a.ldk;
      for (int t = 0; t < k; ++t) {
        hd[t * Q + lane] = a.D[row + t];
        hi[t * Q + lane] = a.I[row + t];
      }
      for (int s = k / 2 - 1; s >= 0; --s)
        sift_down(hd, hi, Q, lane, s, k, hd[s * Q + lane], hi[s * Q + lane]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

    float thr[QB];
    auto refresh = [&]() __attribute__((always_inline)) {
#pragma unroll
      for (int qb = 0; qb < QB; ++qb)
        thr[qb] = qv[qb] ? f32_pass_thr(hd[ql[qb]], nq[qb], dmx[qb]) : -__builtin_inff();
    };
    refresh();

    // Candidate tiles: this lane's candidates cb 16 + cl, their dims
    // grp KS .. + KS; ids and norms for the epilogue (NaN norm: no candidate).
    const int ntiles = (L_ + C - 1) / C;
    float xa[CB][KS];
    int ci[CB];
    float cn[CB];
    auto load_tile = [&](int t, float (&xr)[CB][KS], int (&ids)[CB], float (&nc)[CB])
        __attribute__((always_inline)) {
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const int c = t * C + cb * 16 + cl;
        const int id = leaf[min(c, L_ - 1)];
        const float *xr0 = a.X + (size_t)id * DD + grp * KS;
#pragma unroll
        for (int s = 0; s < KS; s += 4) {
          const float4 f = *(const float4 *)(xr0 + s);
          xr[cb][s] = f.x;
          xr[cb][s + 1] = f.y;
          xr[cb][s + 2] = f.z;
          xr[cb][s + 3] = f.w;
        }
        ids[cb] = id;
        nc[cb] = c < L_ ? a.X2[id] : __builtin_nanf("");
      }
    };
    load_tile(0, xa, ci, cn);

    int pq_i[QR], pq_q[QR];
    int pq_n = 0;
#pragma unroll
    for (int r = 0; r < QR; ++r) pq_i[r] = pq_q[r] = 0;
    auto flush = [&]() __attribute__((always_inline)) {
      float ed[QR];
      float cnr[QR];
#pragma unroll
      for (int r = 0; r < QR; ++r) {
        const int id = pq_i[r];
        cnr[r] = a.X2[id];
      }
#pragma unroll
      for (int r = 0; r < QR; ++r) {
        int slot = ql[0];
        float qn = nq[0];
#pragma unroll
        for (int qb = 1; qb < QB; ++qb)
          if (pq_q[r] == qb) {
            slot = ql[qb];
            qn = nq[qb];
          }
        ed[r] = r < pq_n ? sweep_oracle_dist<DD>(a.X + (size_t)pq_i[r] * DD, qrow + slot * DD, cnr[r], qn)
                         : 0.0f;
      }
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
              if (cless(ed[r], pq_i[r], hd[slot], hi[slot]))
                sift_down(hd, hi, Q, slot, 0, k, ed[r], pq_i[r]);
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      pq_n = 0;
#pragma unroll
      for (int r = 0; r < QR; ++r) pq_i[r] = pq_q[r] = 0;
      refresh();
    };

    for (int t = 0; t < ntiles; ++t) {
      const bool more = t + 1 < ntiles;
      float xn_[CB][KS];
      int ci_n[CB];
      float cn_n[CB];
      if (more) load_tile(t + 1, xn_, ci_n, cn_n);
      f32x4 acc[QB][CB];
#pragma unroll
      for (int qb = 0; qb < QB; ++qb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) acc[qb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
          for (int qb = 0; qb < QB; ++qb)
            acc[qb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[cb][s], bq[qb][s], acc[qb][cb], 0, 0, 0);
      // Epilogue: lane l holds candidates 4 grp + v of each block against
      // query cl of each query block.
      float ctn[CB * 4];
      int cid[CB * 4];
#pragma unroll
      for (int b = 0; b < CB * 4; ++b) {
        ctn[b] = __shfl(cn[b >> 2], 4 * grp + (b & 3));
        cid[b] = __shfl(ci[b >> 2], 4 * grp + (b & 3));
      }
      unsigned pass[QB];
      bool anyp = false;
#pragma unroll
      for (int qb = 0; qb < QB; ++qb) {
        pass[qb] = 0u;
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float tt = __builtin_fmaf(-2.0f, acc[qb][cb][v], ctn[cb * 4 + v]);
            pass[qb] |= (tt <= thr[qb] ? 1u : 0u) << (cb * 4 + v);
          }
        anyp |= pass[qb] != 0u;
      }
      if (__any(anyp)) {
        // Queue the passing pairs; whenever some lane's queue is full,
        // apply the queues and drop the pairs that fail the new thresholds.
        while (true) {
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) {
            while (pass[qb] && pq_n < QR) {
              const int b = __builtin_ctz(pass[qb]);
              pass[qb] &= pass[qb] - 1;
              int idx = cid[0];
#pragma unroll
              for (int bb = 1; bb < CB * 4; ++bb)
                if (b == bb) idx = cid[bb];
#pragma unroll
              for (int r = 0; r < QR; ++r)
                if (pq_n == r) {
                  pq_i[r] = idx;
                  pq_q[r] = qb;
                }
              ++pq_n;
            }
          }
          bool left = false;
#pragma unroll
          for (int qb = 0; qb < QB; ++qb) left |= pass[qb] != 0u;
          if (!__any(left)) break;
          flush();
#pragma unroll
          for (int qb = 0; qb < QB; ++qb)
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const float tt = __builtin_fmaf(-2.0f, acc[qb][cb][v], ctn[cb * 4 + v]);
                if (!(tt <= thr[qb])) pass[qb] &= ~(1u << (cb * 4 + v));
              }
        }
      }
      if (more) {
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
#pragma unroll
          for (int s = 0; s < KS; ++s) xa[cb][s] = xn_[cb][s];
          ci[cb] = ci_n[cb];
          cn[cb] = cn_n[cb];
        }
      }
    }
    if (__any(pq_n > 0)) flush();
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < Q && q0 + lane < L_) {
      // The row is stored as the max-heap it is (the reference's row
      // format; the next iteration's load and the sorted-output call take
      // any heap order).
      const size_t row = (size_t)leaf[q0 + lane] * a.ldk;
      for (int t = 0; t < k; ++t) {
        a.D[row + t] = hd[t * Q + lane];
        a.I[row + t] = hi[t * Q + lane];
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // LDS reused by the next item
  }
}

// Norms for the MFMA sweep: xn[i] = |X[i]| (FP64 sum, rounded up) and the
// table maxima {max xn, max |X2|} (mx must hold two zeros beforehand).
__global__ void sweep_norms_kernel(int N, int d, const float *X, const float *X2, float *xn,
                                   float *mx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float *x = X + (size_t)i * d;
  double s = 0.0;
  for (int p = 0; p < d; ++p) s += (double)x[p] * (double)x[p];
  const float v = __double2float_ru(sqrt(s) * (1.0 + 0x1p-40));
  xn[i] = v;
  atomicMax((int *)&mx[0], __float_as_int(v));
  atomicMax((int *)&mx[1], __float_as_int(fabsf(X2[i])));
}

__global__ void sweep_zero2_kernel(float *mx) {
  if (threadIdx.x < 2) mx[threadIdx.x] = 0.0f;
}

hipError_t launch_sweep_norms(int N, int d, const float *X, const float *X2, float *xn, float *mx,
                              hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(sweep_zero2_kernel, dim3(1), dim3(64), 0, s, mx);
  hipLaunchKernelGGL(sweep_norms_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, d, X, X2, xn, mx);
  return hipGetLastError();
}

size_t leaf_sweep_mfma_lds(int d, int k, int qpb) { return SweepMfmaLayout(k, qpb, d).bytes; }

bool leaf_sweep_mfma_ok(int d, int k, int qpb) {
  return (d == 16 || d == 32) && k >= 1 && k <= kSweepMaxK && qpb == 32 &&
         leaf_sweep_mfma_lds(d, k, qpb) <= 64 * 1024;
}

hipError_t launch_leaf_sweep_mfma(int d, int k, const float *X, const float *X2, const float *xn,
                                  const float *mx, const int *perm, const int *items,
                                  const int *nitems, float *D, int *I, int ldk, int qpb, int grid,
                                  hipStream_t s) {
  if (!leaf_sweep_mfma_ok(d, k, qpb) || grid < 1) return hipErrorNotSupported;
  SweepMfmaArgs a{};
  a.d = d; a.k = k; a.ldk = ldk;
  a.X = X; a.X2 = X2; a.xn = xn; a.mx = mx; a.perm = perm; a.D = D; a.I = I;
  const size_t lds = leaf_sweep_mfma_lds(d, k, qpb);
  void (*kern)(SweepMfmaArgs, const int4 *, const int *) = nullptr;
#define GSKNN_SWEEP_MFMA(D_, QPB_) \
  if (d == D_ && qpb == QPB_) kern = knn_leaf_sweep_mfma_kernel<D_ / 4, QPB_ / 16, 4>;
  GSKNN_SWEEP_MFMA(16, 32)
  GSKNN_SWEEP_MFMA(32, 32)
#undef GSKNN_SWEEP_MFMA
  if (!kern) return hipErrorNotSupported;
  hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, s, a, (const int4 *)items, nitems);
  return hipGetLastError();
}

}  // namespace gsknn
