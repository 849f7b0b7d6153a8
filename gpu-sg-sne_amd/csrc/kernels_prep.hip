// kernels_prep.hip -- construction of P on the GPU: exact kNN, perplexity
// bisection, lambda rescaling, P + P^T, and the COO / hybrid format builds.
//
// Reference (CU = code/cuda/SG-tSNE-CUDA/src, HY = code/cuda/SG-tSNE-HYB/sgtsnepi/src):
//   kNN            CU/allknn.cuh:75-92 (FLANN, approximate; exact here)
//   perplexity     HY/sgtsne.cpp:201-254 (FP64 twin of CU/perplexity_search.cu:18-208)
//   lambda         HY/graph_rescaling.cpp:19-104 (CU/graph_rescaling.cu:11-107)
//   P + P^T        CU/graph_rescaling.cu:266-398, HY/sparsematrix.cpp:69-164,
//                  normalised to sum 1 as HY/sgtsne.cpp:78-85
//   COO rows       CU/prepareMatrix.cu:2-27
//   hybrid         CU/prepareMatrix.cu:56-97, CU/matrix_converter.cpp:106-248
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/sgtsne.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <climits>
#include <vector>

namespace sgt {

namespace {

// ---- exact kNN -----------------------------------------------------------
// Work-group tile: 64 queries x 32 references; 256 threads each own a 4x2
// micro-tile of (query, reference) distances accumulated over the dimensions
// in order with fma(diff, diff, acc).  Candidates below a query's current
// worst key (distance, index) are queued in LDS and merged by one thread per
// query into a max-heap of k entries kept in global memory.  Queries are the
// slots [0, nq): slot s is point qlist[s] (point s without a list); heaps are
// indexed by slot.
constexpr int kKnnQ = 64;
constexpr int kKnnR = 32;
constexpr int kKnnDC = 16;  // dimensions staged per round
constexpr int kKnnThreads = 256;

__device__ __forceinline__ bool key_less(double da, int ia, double db, int ib) {
  return da < db || (da == db && ia < ib);
}

__global__ __launch_bounds__(kKnnThreads) void knn_kernel(const double *__restrict__ X, int n,
                                                          int dim, int k, double *__restrict__ hd,
                                                          int *__restrict__ hi, const int *__restrict__ qlist,
                                                          int nq) {
  __shared__ double qs[kKnnDC][kKnnQ + 1];
  __shared__ double rs[kKnnDC][kKnnR + 1];
  __shared__ double qd[kKnnQ][kKnnR];
  __shared__ int qi[kKnnQ][kKnnR];
  __shared__ int qcnt[kKnnQ];
  __shared__ double thr_d[kKnnQ];
  __shared__ int thr_i[kKnnQ];
  __shared__ int qid[kKnnQ];  // point of each slot (-1: none)

  const int q0 = blockIdx.x * kKnnQ;
  const int tid = threadIdx.x;
  const int tq = tid >> 4;   // 0..15 -> queries tq*4 .. tq*4+3
  const int tr = tid & 15;   // 0..15 -> references tr + 16*c
  if (tid < kKnnQ) {
    qcnt[tid] = 0;
    thr_d[tid] = INFINITY;
    thr_i[tid] = INT_MAX;
    const int slot = q0 + tid;
    qid[tid] = slot < nq ? (qlist ? qlist[slot] : slot) : -1;
    if (slot < nq)
      for (int e = 0; e < k; ++e) {
        hd[(size_t)slot * k + e] = INFINITY;
        hi[(size_t)slot * k + e] = INT_MAX;
      }
  }
  __syncthreads();

  for (int r0 = 0; r0 < n; r0 += kKnnR) {
    double acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = 0.0;
    for (int l0 = 0; l0 < dim; l0 += kKnnDC) {
      for (int idx = tid; idx < kKnnDC * kKnnQ; idx += kKnnThreads) {
        const int qq = idx / kKnnDC, ll = idx % kKnnDC;
        const int q = qid[qq], l = l0 + ll;
        qs[ll][qq] = (q >= 0 && l < dim) ? X[(size_t)q * dim + l] : 0.0;
        if (qq < kKnnR) {
          const int r = r0 + qq;
          rs[ll][qq] = (r < n && l < dim) ? X[(size_t)r * dim + l] : 0.0;
        }
      }
      __syncthreads();
      const int lc = min(kKnnDC, dim - l0);
      for (int ll = 0; ll < lc; ++ll) {
        double qv[4], rv[2];
#pragma unroll
        for (int a = 0; a < 4; ++a) qv[a] = qs[ll][tq * 4 + a];
#pragma unroll
        for (int b = 0; b < 2; ++b) rv[b] = rs[ll][tr + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const double df = qv[a] - rv[b];
            acc[a][b] = fma(df, df, acc[a][b]);
          }
      }
      __syncthreads();
    }
    // queue candidates
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int ql = tq * 4 + a;
      const int q = qid[ql];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int r = r0 + tr + 16 * b;
        if (q >= 0 && r < n && r != q && key_less(acc[a][b], r, thr_d[ql], thr_i[ql])) {
          const int pos = atomicAdd(&qcnt[ql], 1);
          qd[ql][pos] = acc[a][b];
          qi[ql][pos] = r;
        }
      }
    }
    __syncthreads();
    // merge queued candidates into the heaps (one thread per query)
    if (tid < kKnnQ) {
      const int slot = q0 + tid;
      const int cnt = qcnt[tid];
      if (slot < nq && cnt > 0) {
        double *H = hd + (size_t)slot * k;
        int *HI = hi + (size_t)slot * k;
        double rd = H[0];
        int ri = HI[0];
        for (int c = 0; c < cnt; ++c) {
          const double cd = qd[tid][c];
          const int ci = qi[tid][c];
          if (!key_less(cd, ci, rd, ri)) continue;
          int pos = 0;
          for (;;) {
            const int l = 2 * pos + 1, rr = l + 1;
            int m = -1;
            double md = cd;
            int mi = ci;
            if (l < k && key_less(md, mi, H[l], HI[l])) { m = l; md = H[l]; mi = HI[l]; }
            if (rr < k && key_less(md, mi, H[rr], HI[rr])) { m = rr; md = H[rr]; mi = HI[rr]; }
            if (m < 0) break;
            H[pos] = md;
            HI[pos] = mi;
            pos = m;
          }
          H[pos] = cd;
          HI[pos] = ci;
          rd = H[0];
          ri = HI[0];
        }
        thr_d[tid] = rd;
        thr_i[tid] = ri;
      }
      qcnt[tid] = 0;
    }
    __syncthreads();
  }
}

// Heap -> ascending rows with the point itself first (slots as knn_kernel).
__global__ void knn_finalize_kernel(int nq, int k, double *__restrict__ hd, int *__restrict__ hi,
                                    const int *__restrict__ qlist, int32_t *__restrict__ I,
                                    double *__restrict__ D) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nq) return;
  const int q = qlist ? qlist[slot] : slot;
  double *H = hd + (size_t)slot * k;
  int *HI = hi + (size_t)slot * k;
  int32_t *Iq = I + (size_t)q * (k + 1);
  double *Dq = D + (size_t)q * (k + 1);
  Iq[0] = q;
  Dq[0] = 0.0;
  for (int e = k - 1; e >= 0; --e) {
    Iq[1 + e] = HI[0];
    Dq[1 + e] = H[0];
    const double cd = H[e];
    const int ci = HI[e];
    int pos = 0;
    for (;;) {
      const int l = 2 * pos + 1, rr = l + 1;
      int m = -1;
      double md = cd;
      int mi = ci;
      if (l < e && key_less(md, mi, H[l], HI[l])) { m = l; md = H[l]; mi = HI[l]; }
      if (rr < e && key_less(md, mi, H[rr], HI[rr])) { m = rr; md = H[rr]; mi = HI[rr]; }
      if (m < 0) break;
      H[pos] = md;
      HI[pos] = mi;
      pos = m;
    }
    H[pos] = cd;
    HI[pos] = ci;
  }
}

// ---- kNN through an FP32 matrix-core filter --------------------------------
// The exact distances above cost two FP64 vector operations per (pair,
// dimension).  Here the squared distances of the mean-centred points, rounded
// to FP32, come from matrix-core dot products, |x|^2 + |y|^2 - 2 x.y, and each
// query keeps its k + kFiltSlack nearest candidates under that approximation.
// A rigorous bound on the approximation error then certifies, per query, that
// no point outside the candidate list can be among the exact k nearest; the
// candidates' distances are recomputed with exactly the arithmetic of
// knn_kernel and the k smallest (distance, index) keys are kept, so a
// certified row equals knn_kernel's row bit for bit.  Rows that cannot be
// certified (ties or near-ties wider than the slack) go through knn_kernel.
constexpr int kFiltQ = 64;        // queries per work-group, 16 per wave
constexpr int kFiltR = 64;        // references per tile, 4 blocks of 16
constexpr int kFiltMaxDp = 128;   // padded dimensions (A operands held in registers)
constexpr int kFiltRS = 80;       // LDS row stride of the transposed tile (16 mod 32 banks)
constexpr int kFiltSlack = 16;    // candidates kept beyond k
constexpr int kRerank = 256;      // re-rank threads and sort width (k + slack <= 256)

typedef float fv4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool key_lessf(float da, int ia, float db, int ib) {
  return da < db || (da == db && ia < ib);
}

// A reference tile is kFiltR rows of NST float4.  Element t of the tile is
// row t % kFiltR, float4 column t / kFiltR: adjacent threads take adjacent
// rows, so the transposed LDS stores are conflict-free.  All of a thread's
// loads are issued before any is used.
template <int T, int NST>
__device__ __forceinline__ void tile_load(const float *__restrict__ Xc, const float *__restrict__ n32, int n, int r0,
                                          float4 (&buf)[kFiltR * NST / T], float &rnv) {
  static_assert((kFiltR * NST) % T == 0, "a tile must split evenly over the work-group");
  const float4 *X4 = reinterpret_cast<const float4 *>(Xc);
#pragma unroll
  for (int u = 0; u < kFiltR * NST / T; ++u) {
    const int t = threadIdx.x + u * T;
    const int j = t % kFiltR, c4 = t / kFiltR;
    buf[u] = r0 + j < n ? X4[(size_t)(r0 + j) * NST + c4] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  rnv = (threadIdx.x < kFiltR && r0 + (int)threadIdx.x < n) ? n32[r0 + threadIdx.x] : 0.0f;
}

// Stores a loaded tile transposed: rt[c * kFiltRS + j] = Xc[r0 + j][c].
template <int T, int NST>
__device__ __forceinline__ void tile_store(float *rt, float *rn, const float4 (&buf)[kFiltR * NST / T], float rnv) {
#pragma unroll
  for (int u = 0; u < kFiltR * NST / T; ++u) {
    const int t = threadIdx.x + u * T;
    const int j = t % kFiltR, c4 = t / kFiltR;
    float *col = rt + 4 * c4 * kFiltRS + j;
    col[0] = buf[u].x;
    col[kFiltRS] = buf[u].y;
    col[2 * kFiltRS] = buf[u].z;
    col[3 * kFiltRS] = buf[u].w;
  }
  if (threadIdx.x < kFiltR) rn[threadIdx.x] = rnv;
}

// The dot products of a wave's 16 queries with the tile's 64 references:
// acc[t][v] = x(query 16 w + 4 (lane / 16) + v) . x(reference 16 t + lane % 16).
template <int NST>
__device__ __forceinline__ void tile_dots(const float (&areg)[NST], const float *rt, int lane, fv4 (&acc)[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = fv4{0.0f, 0.0f, 0.0f, 0.0f};
  // B operand of step s, block t: row (dimension) 4 s + lane / 16, column
  // (reference) 16 t + lane % 16
  const float *rrow = rt + (lane >> 4) * kFiltRS + (lane & 15);
#pragma unroll
  for (int st = 0; st < NST; ++st)
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[st], rrow[4 * st * kFiltRS + 16 * t], acc[t], 0, 0, 0);
}

// Column means (one work-group of 256 threads per dimension).
__global__ __launch_bounds__(256) void knn_mean_kernel(const double *__restrict__ X, int n, int dim,
                                                       double *__restrict__ mu) {
  __shared__ double part[4];
  const int l = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += X[(size_t)i * dim + l];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) mu[l] = (part[0] + part[1] + part[2] + part[3]) / (double)n;
}

// Centred points rounded to FP32 and zero-padded to dp dimensions, their
// squared norms in FP32 (filter) and norms in FP64 (error bound); the
// largest norm goes to *rmax (non-negative doubles order as their bits).
__global__ void knn_center_kernel(const double *__restrict__ X, int n, int dim, int dp,
                                  const double *__restrict__ mu, float *__restrict__ Xc,
                                  float *__restrict__ n32, double *__restrict__ nrm,
                                  unsigned long long *__restrict__ rmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ss = 0.0;
  for (int l = 0; l < dp; ++l) {
    const float v = l < dim ? (float)(X[(size_t)i * dim + l] - mu[l]) : 0.0f;
    Xc[(size_t)i * dp + l] = v;
    ss += (double)v * (double)v;
  }
  n32[i] = (float)ss;
  const double r = sqrt(ss);
  nrm[i] = r;
  atomicMax(rmax, (unsigned long long)__double_as_longlong(r));
}

// ---- block-centred queries ----------------------------------------------------
// The filter's error bound grows with the size of the vectors whose dot
// product it takes, (|a| + |b|)^2, while neighbour gaps are set by local
// distances: with the points centred once globally, clusters far from the
// mean certify no row.  So queries are grouped into blocks of kF2Q that lie
// close together (in the order of a Morton code of a 3-D projection of the
// centred points), and each block is centred at its own mean c_b: the queries
// as fl32(x - c_b), the references of every tile as fl32(fl32(x - mu) -
// fl32(c_b - mu)), with their norms summed in the filter.  The certificate
// then involves only the query's distance to its block's mean.
__device__ __forceinline__ uint32_t spread_bits10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// +-1/sqrt(dim) entries of three fixed projection directions (a hash of the
// axis and the dimension), so every projection is at most the point's norm.
__device__ __forceinline__ float proj_sign(int axis, int l) {
  uint32_t h = (uint32_t)l * 2654435761u + (uint32_t)axis * 40503u + 0x9E3779B9u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  return (h & 1u) ? 1.0f : -1.0f;
}

// Cells of the point set: the nearest of P pivots (points (p n) / P, P <= 64;
// FP32 distances of the centred points, ties to the lower pivot), so that
// points of one cluster share cells.  Used only to order the points; any
// assignment gives the same kNN rows.
constexpr int kKnnMaxCells = 64;

__global__ __launch_bounds__(256) void knn_cell_kernel(const float *__restrict__ Xc, int n, int dp, int P,
                                                       int32_t *__restrict__ cell) {
  extern __shared__ float piv[];  // [P][dp]
  for (int t = threadIdx.x; t < P * dp; t += blockDim.x) {
    const int p = t / dp, l = t - p * dp;
    piv[t] = Xc[(size_t)(((int64_t)p * n) / P) * dp + l];
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 *x4 = reinterpret_cast<const float4 *>(Xc + (size_t)i * dp);
  float best = INFINITY;
  int bi = 0;
  for (int p = 0; p < P; ++p) {
    const float *pp = piv + p * dp;
    float d2 = 0.0f;
    for (int c = 0; c < dp / 4; ++c) {
      const float4 v = x4[c];
      const float a = v.x - pp[4 * c], b = v.y - pp[4 * c + 1], e = v.z - pp[4 * c + 2], f = v.w - pp[4 * c + 3];
      d2 = fmaf(a, a, d2);
      d2 = fmaf(b, b, d2);
      d2 = fmaf(e, e, d2);
      d2 = fmaf(f, f, d2);
    }
    if (d2 < best) {
      best = d2;
      bi = p;
    }
  }
  cell[i] = bi;
}

// Sort keys: the cell in the top 6 bits (when cells are given), then the
// Morton code of a 3-D projection of the centred point.
__global__ void knn_morton_key_kernel(const float *__restrict__ Xc, int n, int dim, int dp,
                                      const unsigned long long *__restrict__ rmax, const int32_t *__restrict__ cell,
                                      uint32_t *__restrict__ key, int32_t *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float R = (float)__longlong_as_double((long long)*rmax);
  const float scale = R > 0.0f ? 512.0f / (R * sqrtf((float)dim)) : 0.0f;
  float p[3] = {0.0f, 0.0f, 0.0f};
  for (int l = 0; l < dim; ++l) {
    const float v = Xc[(size_t)i * dp + l];
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] += v * proj_sign(a, l);
  }
  uint32_t k = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = fminf(fmaxf(p[a] * scale + 512.0f, 0.0f), 1023.0f);
    k |= spread_bits10((uint32_t)u) << a;
  }
  key[i] = cell ? ((uint32_t)cell[i] << 24) | (k >> 6) : k;
  idx[i] = i;
}

// Points per cell (the sorted order groups each cell's points together):
// a work-group histogram in LDS, then one global addition per cell.
__global__ __launch_bounds__(256) void knn_cell_count_kernel(const int32_t *__restrict__ cell, int n,
                                                             int *__restrict__ cnt) {
  __shared__ int h[kKnnMaxCells];
  if (threadIdx.x < kKnnMaxCells) h[threadIdx.x] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) atomicAdd(&h[cell[i]], 1);
  __syncthreads();
  if (threadIdx.x < kKnnMaxCells && h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

// First sorted position of each cell, and its first reference slot: every
// cell starts a fresh tile of kFiltR slots (the unused slots are padding).
__global__ void knn_cell_offsets_kernel(const int *__restrict__ cnt, int P, int *__restrict__ start,
                                        int *__restrict__ slot0) {
  if (threadIdx.x != 0) return;
  int a = 0, b = 0;
  for (int c = 0; c < P; ++c) {
    start[c] = a;
    slot0[c] = b;
    a += cnt[c];
    b += (cnt[c] + kFiltR - 1) / kFiltR * kFiltR;
  }
}

// The reference stream in cell order: slot of sorted position i, its point
// id (padding slots keep -1 and zero rows) and the id -> slot map.
__global__ void knn_ref_scatter_kernel(const float *__restrict__ Xc, int n, int dp,
                                       const int32_t *__restrict__ order, const int32_t *__restrict__ cell,
                                       const int *__restrict__ start, const int *__restrict__ slot0,
                                       float *__restrict__ Xr, int32_t *__restrict__ rid, int32_t *__restrict__ rinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int id = order[i], c = cell[id];
  const int sl = slot0[c] + (i - start[c]);
  rid[sl] = id;
  rinv[id] = sl;
  const float4 *src = reinterpret_cast<const float4 *>(Xc + (size_t)id * dp);
  float4 *dst = reinterpret_cast<float4 *>(Xr + (size_t)sl * dp);
  for (int c4 = 0; c4 < dp / 4; ++c4) dst[c4] = src[c4];
}

// Bounding sphere of each reference tile (one wave per tile), in the
// caller's coordinates (FP64): centre = mean of the tile's points, radius =
// the largest distance to it; -1 for a tile of padding only.
__global__ __launch_bounds__(64) void knn_tile_sphere_kernel(const double *__restrict__ X, int dim,
                                                             const int32_t *__restrict__ rid,
                                                             double *__restrict__ tcen, double *__restrict__ trad) {
  __shared__ double cen[kFiltMaxDp];
  const int t = blockIdx.x, lane = threadIdx.x;
  const int id = rid[(size_t)t * kFiltR + lane];
  const bool ok = id >= 0;
  const double cnt = wave_sum(ok ? 1.0 : 0.0);
  if (cnt == 0.0) {
    if (lane == 0) trad[t] = -1.0;
    return;
  }
  for (int l = 0; l < dim; ++l) {
    const double c = wave_sum(ok ? X[(size_t)id * dim + l] : 0.0) / cnt;
    if (lane == 0) {
      cen[l] = c;
      tcen[(size_t)t * dim + l] = c;
    }
  }
  __syncthreads();
  double r2 = 0.0;
  if (ok)
    for (int l = 0; l < dim; ++l) {
      const double v = X[(size_t)id * dim + l] - cen[l];
      r2 = fma(v, v, r2);
    }
  const double r = wave_max(sqrt(r2));
  if (lane == 0) trad[t] = r;
}

// One work-group per query block: the block's mean c_b (FP64) over its valid
// slots, delta_b = fl32(c_b - mu) padded to dp, and |delta_b|.
__global__ __launch_bounds__(256) void knn_block_centre_kernel(const double *__restrict__ X, int n, int dim, int dp,
                                                               const int32_t *__restrict__ qorder,
                                                               const double *__restrict__ mu,
                                                               double *__restrict__ cb, float *__restrict__ delta,
                                                               double *__restrict__ dn, int bq) {
  __shared__ double part[4];
  const int b = blockIdx.x;
  const int s0 = b * bq, s1 = min(n, s0 + bq);
  double ss = 0.0;
  for (int l = threadIdx.x; l < dp; l += blockDim.x) {
    float dv = 0.0f;
    if (l < dim) {
      double c = 0.0;
      for (int sl = s0; sl < s1; ++sl) c += X[(size_t)qorder[sl] * dim + l];
      c /= (double)(s1 - s0);
      cb[(size_t)b * dim + l] = c;
      dv = (float)(c - mu[l]);
    }
    delta[(size_t)b * dp + l] = dv;
    ss += (double)dv * (double)dv;
  }
  ss = wave_sum(ss);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) dn[b] = sqrt(part[0] + part[1] + part[2] + part[3]);
}

// Queries in slot order, centred at their block's mean: Xq[slot] =
// fl32(x_q - c_b) (zero-padded), its squared norm (FP32 of the FP64 sum) by
// slot, and by query id its norm |a| and |delta_b| for the certificate.
__global__ void knn_query_prep_kernel(const double *__restrict__ X, int n, int dim, int dp,
                                      const int32_t *__restrict__ qorder, const double *__restrict__ cb,
                                      const double *__restrict__ dn, float *__restrict__ Xq,
                                      float *__restrict__ qn32, double *__restrict__ qa, double *__restrict__ qdn,
                                      int bq) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= n) return;
  const int q = qorder[sl], b = sl / bq;
  double ss = 0.0;
  for (int l = 0; l < dp; ++l) {
    const float v = l < dim ? (float)(X[(size_t)q * dim + l] - cb[(size_t)b * dim + l]) : 0.0f;
    Xq[(size_t)sl * dp + l] = v;
    ss += (double)v * (double)v;
  }
  qn32[sl] = (float)ss;
  qa[q] = sqrt(ss);
  qdn[q] = dn[b];
}

// Candidate lists: a max-heap of kp (approximate distance, index) keys per
// query in global memory (lists longer than the LDS version holds).  Wave w
// owns queries q0 + 16w .. q0 + 16w + 15.  The points have 4 NST dimensions.
template <int NST>
__global__ __launch_bounds__(256) void knn_filter_kernel(const float *__restrict__ Xc,
                                                         const float *__restrict__ n32, int n, int kp,
                                                         float *__restrict__ hd, int *__restrict__ hi) {
  __shared__ float rt[4 * NST * kFiltRS];
  __shared__ float rn[kFiltR];
  __shared__ float qd[kFiltQ][kFiltR];
  __shared__ int qi[kFiltQ][kFiltR];
  __shared__ int qcnt[kFiltQ];
  __shared__ float thr_d[kFiltQ];
  __shared__ int thr_i[kFiltQ];
  __shared__ float qn[kFiltQ];

  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int q0 = blockIdx.x * kFiltQ;
  // A operand of step s: row (query) lane % 16, column (dimension) 4 s + lane / 16
  float areg[NST];
  {
    const int qa = q0 + 16 * w + (lane & 15);
#pragma unroll
    for (int st = 0; st < NST; ++st) areg[st] = qa < n ? Xc[(size_t)qa * 4 * NST + 4 * st + (lane >> 4)] : 0.0f;
  }
  if (tid < kFiltQ) {
    qcnt[tid] = 0;
    thr_d[tid] = INFINITY;
    thr_i[tid] = INT_MAX;
    const int q = q0 + tid;
    qn[tid] = q < n ? n32[q] : 0.0f;
    if (q < n)
      for (int e = 0; e < kp; ++e) {
        hd[(size_t)q * kp + e] = INFINITY;
        hi[(size_t)q * kp + e] = INT_MAX;
      }
  }
  float4 buf[kFiltR * NST / 256];
  float rnv;
  tile_load<256, NST>(Xc, n32, n, 0, buf, rnv);
  for (int r0 = 0; r0 < n; r0 += kFiltR) {
    __syncthreads();  // the previous tile's LDS is free
    tile_store<256, NST>(rt, rn, buf, rnv);
    __syncthreads();
    // the next tile's loads fly while this one is multiplied and merged
    if (r0 + kFiltR < n) tile_load<256, NST>(Xc, n32, n, r0 + kFiltR, buf, rnv);
    fv4 acc[4];
    tile_dots<NST>(areg, rt, lane, acc);
    // output v of block t: query 16 w + 4 (lane / 16) + v, reference 16 t + lane % 16
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int jl = 16 * t + (lane & 15), r = r0 + jl;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ql = 16 * w + 4 * (lane >> 4) + v, q = q0 + ql;
        const float dd = (qn[ql] + rn[jl]) - 2.0f * acc[t][v];
        if (q < n && r < n && r != q && key_lessf(dd, r, thr_d[ql], thr_i[ql])) {
          const int pos = atomicAdd(&qcnt[ql], 1);
          qd[ql][pos] = dd;
          qi[ql][pos] = r;
        }
      }
    }
    __syncthreads();
    if (tid < kFiltQ) {
      const int q = q0 + tid;
      const int cnt = qcnt[tid];
      if (q < n && cnt > 0) {
        float *H = hd + (size_t)q * kp;
        int *HI = hi + (size_t)q * kp;
        float rd = H[0];
        int ri = HI[0];
        for (int c = 0; c < cnt; ++c) {
          const float cd = qd[tid][c];
          const int ci = qi[tid][c];
          if (!key_lessf(cd, ci, rd, ri)) continue;
          int pos = 0;
          for (;;) {
            const int l = 2 * pos + 1, rr = l + 1;
            int m = -1;
            float md = cd;
            int mi = ci;
            if (l < kp && key_lessf(md, mi, H[l], HI[l])) { m = l; md = H[l]; mi = HI[l]; }
            if (rr < kp && key_lessf(md, mi, H[rr], HI[rr])) { m = rr; md = H[rr]; mi = HI[rr]; }
            if (m < 0) break;
            H[pos] = md;
            HI[pos] = mi;
            pos = m;
          }
          H[pos] = cd;
          HI[pos] = ci;
          rd = H[0];
          ri = HI[0];
        }
        thr_d[tid] = rd;
        thr_i[tid] = ri;
      }
      qcnt[tid] = 0;
    }
  }
}

// The same filter with the candidate heaps in LDS (lists of up to kF2MaxKp):
// 32 queries per work-group (two waves).  A tile's approximate distances go
// to LDS; each query's candidates of the tile form a 64-bit mask built from
// wave ballots (16 bits per 16-lane group), and the merge walks the mask in
// reference order and sifts into the LDS heap.
//
// The references come as tiles of a cell order (see knn_cell_kernel), each
// tile with a bounding sphere.  A block starts at the tiles around its own
// queries' (its lists fill with near candidates), then visits the others,
// 128 at a time, skipping every tile that cannot hold a candidate: with D
// the distance between the block's and the tile's spheres, every query a and
// reference b of them have a true distance of at least D, and the filter
// value of the pair is within 2^-12 S^2 of its squared distance (S = block
// radius + centre distance + tile radius + R + |c_b - mu|: more than eight
// times the certificate's error bounds for the rounding of the centred
// points and the FP32 dot products), so a tile with D^2 - 2^-12 S^2 above
// every query's current worst key holds no candidate.  Skipped tiles thus
// leave the lists as a full scan would: the k + kFiltSlack smallest keys.
constexpr int kF2Q = 32;
constexpr int kF2MaxKp = 128;
constexpr int kFiltHomeBack = 16;  // tiles visited before a block's own

// A reference tile shifted into a query block's frame: rt holds
// fl32(x~ - delta) transposed, and rnp[t] the partial squared norm of the
// row this thread stored (rows t % 64: two threads cover every row, so a
// row's norm is rnp[j] + rnp[j + 64]).
template <int NST>
__device__ __forceinline__ void tile_store_shift(float *rt, float *rnp, const float4 (&buf)[kFiltR * NST / 128],
                                                 const float4 (&dl)[kFiltR * NST / 128]) {
  float ss = 0.0f;
#pragma unroll
  for (int u = 0; u < kFiltR * NST / 128; ++u) {
    const int t = threadIdx.x + u * 128;
    const int j = t % kFiltR, c4 = t / kFiltR;
    const float4 v = make_float4(buf[u].x - dl[u].x, buf[u].y - dl[u].y, buf[u].z - dl[u].z, buf[u].w - dl[u].w);
    float *col = rt + 4 * c4 * kFiltRS + j;
    col[0] = v.x;
    col[kFiltRS] = v.y;
    col[2 * kFiltRS] = v.z;
    col[3 * kFiltRS] = v.w;
    ss = fmaf(v.x, v.x, ss);
    ss = fmaf(v.y, v.y, ss);
    ss = fmaf(v.z, v.z, ss);
    ss = fmaf(v.w, v.w, ss);
  }
  rnp[threadIdx.x] = ss;
}

// The reference stream of the LDS-heap filter: tiles of kFiltR slots in
// cell order (Xr rows, -1 ids for padding), their spheres, and the query
// blocks' spheres (centre cb, radius from the queries' norms qa, |c_b - mu|
// in dn; R the largest |x - mu|).
struct RefTiles {
  const float *Xr;
  const int32_t *rid, *rinv;
  int ntiles;
  const double *tcen, *trad;
  const double *cb, *qa, *dn;
  const unsigned long long *rmax;
  int dim;
  const int32_t *cell;  // point -> cell
  const int *slot0;     // cell -> its first slot (a tile boundary)
};

template <int NST>
__global__ __launch_bounds__(128) void knn_filter_lds_kernel(const RefTiles rf, const float *__restrict__ Xq,
                                                             const float *__restrict__ qn32,
                                                             const int32_t *__restrict__ qorder,
                                                             const float *__restrict__ delta, int n, int kp,
                                                             float *__restrict__ hd, int *__restrict__ hi) {
  static_assert((kFiltR * NST) % 128 == 0, "a tile must split evenly over the work-group");
  extern __shared__ float dyn[];  // heap keys [kF2Q][kp] | heap indices [kF2Q][kp]
  float *hk = dyn;
  int *hx = reinterpret_cast<int *>(hk + (size_t)kF2Q * kp);
  __shared__ float rt[4 * NST * kFiltRS];
  __shared__ float dist[kF2Q][kFiltR];
  __shared__ unsigned long long mask[kF2Q];
  __shared__ float rnp[128];
  __shared__ float qn[kF2Q];
  __shared__ int qid[kF2Q];
  __shared__ unsigned long long skipm[2];
  __shared__ float sthr;
  __shared__ double srho;

  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4;
  const int q0 = blockIdx.x * kF2Q;  // first slot of this query block
  const int ntiles = rf.ntiles;
  float areg[NST];
  {
    const int sa = q0 + 16 * w + (lane & 15);
#pragma unroll
    for (int st = 0; st < NST; ++st) areg[st] = sa < n ? Xq[(size_t)sa * 4 * NST + 4 * st + (lane >> 4)] : 0.0f;
  }
  // this thread's float4 columns of the block's shift
  float4 dl[kFiltR * NST / 128];
  {
    const float4 *D4 = reinterpret_cast<const float4 *>(delta) + (size_t)blockIdx.x * NST;
#pragma unroll
    for (int u = 0; u < kFiltR * NST / 128; ++u) dl[u] = D4[(threadIdx.x + u * 128) / kFiltR];
  }
  for (int t = tid; t < kF2Q * kp; t += 128) {
    hk[t] = INFINITY;
    hx[t] = INT_MAX;
  }
  if (tid < kF2Q) {
    qn[tid] = q0 + tid < n ? qn32[q0 + tid] : 0.0f;
    qid[tid] = q0 + tid < n ? qorder[q0 + tid] : -1;
  }
  if (w == 0) {  // the block's radius: its queries' largest |x - c_b| (|a| rounded up)
    const int q = lane < kF2Q && q0 + lane < n ? qorder[q0 + lane] : -1;
    const double r = wave_max(q >= 0 ? rf.qa[q] * (1.0 + 0x1p-20) : 0.0);
    if (lane == 0) srho = r;
  }
  // the first tile visited: kFiltHomeBack tiles before the block's own, but
  // not before its cell's first tile, so the lists fill from the block's
  // own cell and its threshold prunes the other cells' tiles early
  const int home = rf.rinv[qorder[q0]] / kFiltR;
  const int start = max(home - kFiltHomeBack, rf.slot0[rf.cell[qorder[q0]]] / kFiltR);
  const double R = __longlong_as_double((long long)*rf.rmax);
  const double dnb = rf.dn[blockIdx.x];
  const double *cbb = rf.cb + (size_t)blockIdx.x * rf.dim;
  auto visit = [&](int sq) { return (start + sq) % ntiles; };  // the sq-th tile visited
  float4 buf[kFiltR * NST / 128];
  float rnv;
  for (int g0 = 0; g0 < ntiles; g0 += 128) {
    __syncthreads();  // the previous tile's merge is done
    if (w == 0) {  // the worst key over the block's queries (infinite until every list is full)
      const float th = lane < kF2Q && qid[lane] >= 0 ? hk[lane * kp] : -INFINITY;
      const float m = (float)wave_max((double)th);
      if (lane == 0) sthr = m;
    }
    __syncthreads();
    bool skip = true;
    if (g0 + tid < ntiles) {
      const int t = visit(g0 + tid);
      const double rt_ = rf.trad[t];
      const double thr = (double)sthr;
      if (rt_ < 0.0) {
        skip = true;  // padding only
      } else if (!(thr < INFINITY)) {
        skip = false;
      } else {
        const double *tc = rf.tcen + (size_t)t * rf.dim;
        double d2 = 0.0;
        for (int l = 0; l < rf.dim; ++l) {
          const double v = cbb[l] - tc[l];
          d2 = fma(v, v, d2);
        }
        const double dcm = sqrt(d2), rho = srho;
        const double dlb = dcm - rho - rt_;
        const double S = rho + dcm + rt_ + R + dnb;
        skip = dlb > 0.0 && dlb * dlb - 0x1p-12 * S * S > thr;
      }
    }
    const unsigned long long sb = __ballot(skip);
    if (lane == 0) skipm[w] = sb;
    __syncthreads();
    const unsigned long long todo[2] = {~skipm[0], ~skipm[1]};
    auto next_tile = [&](int i) {  // the next visited index of this group after i, or -1
      const int k = i + 1;
      if (k >= 128) return -1;
      if (k < 64) {
        const unsigned long long m = todo[0] & (~0ull << k);
        if (m) return __builtin_ctzll(m);
      }
      const unsigned long long m = todo[1] & (k <= 64 ? ~0ull : ~0ull << (k - 64));
      return m ? 64 + __builtin_ctzll(m) : -1;
    };
    int i = next_tile(-1);
    if (i >= 0) tile_load<128, NST>(rf.Xr, rf.Xr, ntiles * kFiltR, visit(g0 + i) * kFiltR, buf, rnv);
    while (i >= 0) {
      const int r0 = visit(g0 + i) * kFiltR;
      const int inext = next_tile(i);
      __syncthreads();  // the previous tile's merge is done
      tile_store_shift<NST>(rt, rnp, buf, dl);
      __syncthreads();
      if (inext >= 0) tile_load<128, NST>(rf.Xr, rf.Xr, ntiles * kFiltR, visit(g0 + inext) * kFiltR, buf, rnv);
      int rr[4];  // ids of this lane's reference columns 16 t + lane % 16 (in flight over the dot products)
#pragma unroll
      for (int t = 0; t < 4; ++t) rr[t] = rf.rid[r0 + 16 * t + (lane & 15)];
      fv4 acc[4];
      tile_dots<NST>(areg, rt, lane, acc);
      // this lane's four query rows 16 w + 4 g + v: norms and current worst keys
      float qv[4], thd[4];
      int thi[4], qq[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ql = 16 * w + 4 * g + v;
        qv[v] = qn[ql];
        thd[v] = hk[ql * kp];
        thi[v] = hx[ql * kp];
        qq[v] = qid[ql];
      }
      unsigned long long mq[4] = {0ull, 0ull, 0ull, 0ull};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int jl = 16 * t + (lane & 15);
        const float rv = rnp[jl] + rnp[jl + 64];
        const int r = rr[t];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ql = 16 * w + 4 * g + v;
          const float dd = (qv[v] + rv) - 2.0f * acc[t][v];
          dist[ql][jl] = dd;
          const bool cand = qq[v] >= 0 && r >= 0 && r != qq[v] && key_lessf(dd, r, thd[v], thi[v]);
          const unsigned long long b = __ballot(cand);
          mq[v] |= ((b >> (16 * g)) & 0xFFFFull) << (16 * t);
        }
      }
      if ((lane & 15) == 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) mask[16 * w + 4 * g + v] = mq[v];
      __syncthreads();
      if (tid < kF2Q) {
        float *H = hk + tid * kp;
        int *HI = hx + tid * kp;
        unsigned long long m = mask[tid];
        while (m) {
          const int j = __ffsll((long long)m) - 1;
          m &= m - 1;
          const float cd = dist[tid][j];
          const int ci = rf.rid[r0 + j];
          if (!key_lessf(cd, ci, H[0], HI[0])) continue;
          int pos = 0;
          for (;;) {
            const int l = 2 * pos + 1, rr2 = l + 1;
            int mm = -1;
            float md = cd;
            int mi = ci;
            if (l < kp && key_lessf(md, mi, H[l], HI[l])) { mm = l; md = H[l]; mi = HI[l]; }
            if (rr2 < kp && key_lessf(md, mi, H[rr2], HI[rr2])) { mm = rr2; md = H[rr2]; mi = HI[rr2]; }
            if (mm < 0) break;
            H[pos] = md;
            HI[pos] = mi;
            pos = mm;
          }
          H[pos] = cd;
          HI[pos] = ci;
        }
      }
      i = inext;
    }
  }
  __syncthreads();
  for (int t = tid; t < kF2Q * kp; t += 128) {
    const int ql = t / kp, q = qid[ql];
    if (q >= 0) {
      hd[(size_t)q * kp + (t - ql * kp)] = hk[t];
      hi[(size_t)q * kp + (t - ql * kp)] = hx[t];
    }
  }
}

// Step counts (dimensions / 4, rounded up) the filters are built for.
// The block-centred arrays of the LDS-heap filter (kp <= kF2MaxKp).
struct BlockCentred {
  const float *Xq;         // queries in slot order, centred at their block's mean
  const float *qn32;       // their squared norms
  const int32_t *qorder;   // slot -> query id
  const float *delta;      // per block: fl32(c_b - mu), padded
  RefTiles rf;             // the references in cell order, with tile spheres
};

template <int NST>
int launch_filter(hipStream_t s, const float *Xc, const float *n32, int n, int kp, const BlockCentred &bc,
                  float *hd, int *hi) {
  if (kp <= kF2MaxKp) {
    const size_t lds = (sizeof(float) + sizeof(int)) * (size_t)kF2Q * kp;
    knn_filter_lds_kernel<NST><<<div_up(n, kF2Q), 128, lds, s>>>(bc.rf, bc.Xq, bc.qn32, bc.qorder, bc.delta, n, kp,
                                                                 hd, hi);
  } else {
    if constexpr ((kFiltR * NST) % 256 == 0) {
      knn_filter_kernel<NST><<<div_up(n, kFiltQ), 256, 0, s>>>(Xc, n32, n, kp, hd, hi);
    } else {
      SGT_CHECK(false, SGTSNE_ERR_ARG, "no global-heap filter of %d steps", NST);
    }
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

// float4 steps of the filter's padded dimension.  14 (dims 49-56: the usual
// 50-D PCA input) only with the LDS-heap kernel, whose 128-thread tile load
// splits 64 x 14 float4 evenly: its LDS then fits three work-groups per CU,
// where 16 steps fit two.  The zero padding adds exact zeros, so the FP32
// distances, the candidates and the rows are the same either way.
int filter_steps(int dim, int kp) {
  const int nst = (dim + 3) / 4;
  if (nst > 12 && nst <= 14 && kp <= kF2MaxKp) return 14;
  return nst <= 4 ? 4 : nst <= 8 ? 8 : nst <= 12 ? 12 : nst <= 16 ? 16 : nst <= 24 ? 24 : 32;
}

int launch_filter_for(hipStream_t s, int nst, const float *Xc, const float *n32, int n, int kp,
                      const BlockCentred &bc, float *hd, int *hi) {
  switch (nst) {
    case 4: return launch_filter<4>(s, Xc, n32, n, kp, bc, hd, hi);
    case 8: return launch_filter<8>(s, Xc, n32, n, kp, bc, hd, hi);
    case 12: return launch_filter<12>(s, Xc, n32, n, kp, bc, hd, hi);
    case 14: return launch_filter<14>(s, Xc, n32, n, kp, bc, hd, hi);
    case 16: return launch_filter<16>(s, Xc, n32, n, kp, bc, hd, hi);
    case 24: return launch_filter<24>(s, Xc, n32, n, kp, bc, hd, hi);
    default: return launch_filter<32>(s, Xc, n32, n, kp, bc, hd, hi);
  }
}

// Ascending bitonic sort of kRerank (key, index) pairs in LDS by the key, then
// the index.  Every thread of the work-group calls it.
template <class K>
__device__ void bitonic_sort(K *key, int *id) {
  const int t = threadIdx.x;
  for (int size = 2; size <= kRerank; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      const int p = t ^ stride;
      if (p > t) {
        const bool up = (t & size) == 0;
        const K a = key[t], b = key[p];
        const int ia = id[t], ib = id[p];
        const bool a_after_b = b < a || (b == a && ib < ia);
        if (a_after_b == up) {
          key[t] = b;
          key[p] = a;
          id[t] = ib;
          id[p] = ia;
        }
      }
    }
  __syncthreads();
}

// One work-group per query: certify the candidate list, recompute the
// candidates' distances as knn_kernel does (diff = x_q - x_c in FP64,
// acc = fma(diff, diff, acc) over the dimensions in order) and write the row.
//
// The filter compares a = fl32(x_q - c) with b = fl32(fl32(x_c - mu) -
// fl32(c - mu)) for the frame c of the query's block (c = mu, no shift, when
// qdn is 0): F = |a|^2 + |b|^2 - 2 a.b is within E = cE (|a| + |b|)^2 of
// |a - b|^2, and |a - b| is within e = ce (|a| + R + |c - mu| + |b|) of the
// true distance T^(1/2), R the largest |x - mu|.  The k nearest candidates
// (F <= F_k) lie within hi = sqrt(F_k + E) + e with |b| <= |a| + sqrt(F_k) + 1;
// a point outside the list within hi would have |b| <= |a| + hi + e + 1 and
// so F <= (hi + e)^2 + E, yet every non-candidate has F >= F_kp: when F_kp
// exceeds that bound (with a 1e-12 relative margin for the FP64 rounding of
// the recomputed distances), the list holds the k nearest.
__global__ __launch_bounds__(kRerank) void knn_rerank_kernel(
    const double *__restrict__ X, int n, int dim, int k, int kp, const float *__restrict__ hd,
    const int *__restrict__ hi, const double *__restrict__ qa, const double *__restrict__ qdn,
    const unsigned long long *__restrict__ rmax, double cE, double ce, int32_t *__restrict__ I,
    double *__restrict__ D, int *__restrict__ fb_list, int *__restrict__ fb_cnt) {
  __shared__ float sk[kRerank];
  __shared__ double sd[kRerank];
  __shared__ int si[kRerank];
  __shared__ int ok;
  const int q = blockIdx.x, t = threadIdx.x;
  sk[t] = t < kp ? hd[(size_t)q * kp + t] : INFINITY;
  si[t] = t < kp ? hi[(size_t)q * kp + t] : INT_MAX;
  bitonic_sort(sk, si);
  if (t == 0) {
    bool good = kp >= n - 1;  // every other point is a candidate
    if (!good) {
      const double R = __longlong_as_double((long long)*rmax);
      const double a = qa[q], dl = qdn[q];
      const double fk = fmax((double)sk[k - 1], 0.0), fkp = (double)sk[kp - 1];
      const double bk = a + sqrt(fk) + 1.0;
      const double ek = ce * (a + R + dl + bk);
      const double hik = (sqrt(fk + cE * (a + bk) * (a + bk)) + ek) * (1.0 + 1e-12) + 1e-300;
      const double bn = a + hik + ek + 1.0;
      const double en = ce * (a + R + dl + bn);
      const double fmaxn = (hik + en) * (hik + en) + cE * (a + bn) * (a + bn);
      good = fkp > fmaxn;  // false for NaN keys too
    }
    ok = good;
  }
  const int c = t < kp ? si[t] : INT_MAX;
  double acc = INFINITY;
  // a list left short (NaN distances never enter it) is not certified
  const bool invalid = t < kp && (c < 0 || c >= n);
  if (t < kp && c >= 0 && c < n) {
    acc = 0.0;
    const double *xq = X + (size_t)q * dim;
    const double *xc = X + (size_t)c * dim;
    for (int l = 0; l < dim; ++l) {
      const double df = xq[l] - xc[l];
      acc = fma(df, df, acc);
    }
  }
  __syncthreads();  // thread 0 has written ok
  if (invalid) ok = 0;
  sd[t] = acc;
  si[t] = c;
  bitonic_sort(sd, si);
  int32_t *Iq = I + (size_t)q * (k + 1);
  double *Dq = D + (size_t)q * (k + 1);
  if (t == 0) {
    Iq[0] = q;
    Dq[0] = 0.0;
    if (!ok) fb_list[atomicAdd(fb_cnt, 1)] = q;
  }
  if (t < k) {
    Iq[1 + t] = si[t];
    Dq[1 + t] = sd[t];
  }
}

// ---- perplexity bisection (HY/sgtsne.cpp:201-254), 16 lanes per point ------
// The reference's bisection on beta, step for step: the entropy H of the
// point's row at beta against log(u), beta doubled / halved until bracketed,
// then bisected, until |H - log u| < 1e-5 (at most 200 steps).  Each of the 16
// lanes of a point's group takes the neighbours m = lane, lane + 16, ...; a
// step's two sums (sum_m exp(-beta d_m) and sum_m d_m exp(-beta d_m)) are
// added over the group by a fixed butterfly (deterministic), so a step costs
// about k / 16 exponentials per lane instead of 2 k on one thread.  As in the
// reference, the values normalised at the end are those of the last
// evaluated beta (recomputed from it).
constexpr int kPerpLanes = 16;

__global__ __launch_bounds__(256) void perplexity_kernel2(const double *__restrict__ D, int n, int k, double u,
                                                          float *__restrict__ val, double *__restrict__ beta_out) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kPerpLanes;
  const int l = threadIdx.x & (kPerpLanes - 1);
  if (i >= n) return;  // whole groups leave together
  const double *dist = D + (size_t)i * (k + 1) + 1;
  const double tol = 1e-5;
  const double logu = log(u);
  double beta = 1.0, min_beta = -DBL_MAX, max_beta = DBL_MAX;
  double sum_P = 0.0, beta_eval = 1.0;
  bool found = false;
  int iter = 0;
  while (!found && iter < 200) {  // uniform over the group
    beta_eval = beta;
    double sp = 0.0, sh = 0.0;
    for (int m = l; m < k; m += kPerpLanes) {
      const double e = exp(-beta * dist[m]);
      sp += e;
      sh += dist[m] * e;
    }
#pragma unroll
    for (int o = kPerpLanes / 2; o > 0; o >>= 1) {
      sp += __shfl_xor(sp, o, kPerpLanes);
      sh += __shfl_xor(sh, o, kPerpLanes);
    }
    sum_P = DBL_MIN + sp;
    const double H = (beta * sh) / sum_P + log(sum_P);
    const double Hdiff = H - logu;
    if (Hdiff < tol && -Hdiff < tol) {
      found = true;
    } else if (Hdiff > 0) {
      min_beta = beta;
      beta = (max_beta == DBL_MAX || max_beta == -DBL_MAX) ? beta * 2.0 : (beta + max_beta) / 2.0;
    } else {
      max_beta = beta;
      beta = (min_beta == -DBL_MAX || min_beta == DBL_MAX) ? beta / 2.0 : (beta + min_beta) / 2.0;
    }
    ++iter;
  }
  for (int m = l; m < k; m += kPerpLanes) val[(size_t)i * k + m] = (float)(exp(-beta_eval * dist[m]) / sum_P);
  if (beta_out && l == 0) beta_out[i] = beta;
}

// ---- lambda rescaling (one thread per row, FP64) --------------------------
__global__ void lambda_kernel(int n, const int32_t *__restrict__ rowptr, float *__restrict__ val,
                              double lambda, int drop_leaf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tolBinary = 1e-5;
  const int maxIter = 100;
  double fval = 1.0 - lambda;
  double sig2 = 1.0, a = -1e3, c = INFINITY;
  int iter = 0;
  const int s = rowptr[i], e = rowptr[i + 1];
  while (fabs(fval) > tolBinary && iter < maxIter) {
    ++iter;
    if (fval > 0) {
      a = sig2;
      sig2 = isinf(c) ? 2 * a : 0.5 * (a + c);
    } else {
      c = sig2;
      sig2 = 0.5 * (a + c);
    }
    double sum_i = 0.0;
    for (int j = s; j < e; ++j) sum_i += exp(log((double)val[j]) * sig2);
    fval = sum_i - lambda;
  }
  double sum_i = 0.0;
  for (int j = s; j < e; ++j) sum_i += exp(log((double)val[j]) * sig2);
  for (int j = s; j < e; ++j) val[j] = (float)(exp(log((double)val[j]) * sig2) / sum_i);
  if (drop_leaf && e - s == 1) val[s] = 0.0f;
}

__global__ void rowids_kernel(int n, const int32_t *__restrict__ rowptr, int32_t *__restrict__ row) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int t = rowptr[i]; t < rowptr[i + 1]; ++t) row[t] = i;
}

// ---- P + P^T ----------------------------------------------------------------
__global__ void sym_keys_kernel(int n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                const float *__restrict__ val, int64_t nnz,
                                unsigned long long *__restrict__ keys, double *__restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int t = rowptr[i]; t < rowptr[i + 1]; ++t) {
    const unsigned long long r = (unsigned)i, c = (unsigned)col[t];
    keys[t] = (r << 32) | c;
    keys[nnz + t] = (c << 32) | r;
    vals[t] = (double)val[t];
    vals[nnz + t] = (double)val[t];
  }
}

__global__ void sym_finish_kernel(int64_t m, const unsigned long long *__restrict__ ukeys,
                                  const double *__restrict__ uvals, const double *__restrict__ total,
                                  int32_t *__restrict__ col, float *__restrict__ val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  col[t] = (int32_t)(ukeys[t] & 0xffffffffull);
  val[t] = (float)(uvals[t] / *total);
}

__global__ void sym_rowptr_kernel(int n, int64_t m, const unsigned long long *__restrict__ ukeys,
                                  int32_t *__restrict__ rowptr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  // first position whose row >= r
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(ukeys[mid] >> 32) < r) lo = mid + 1;
    else hi = mid;
  }
  rowptr[r] = (int32_t)lo;
}

// ---- hybrid ELL + COO tail ----------------------------------------------------
__global__ void build_ell_kernel(int n, int width, const int32_t *__restrict__ rowptr,
                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                 int32_t *__restrict__ ecol, float *__restrict__ eval,
                                 int32_t *__restrict__ tail_cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = rowptr[r], len = rowptr[r + 1] - s;
  for (int e = 0; e < width; ++e) {
    const int64_t o = (int64_t)e * n + r;
    if (e < len) { ecol[o] = col[s + e]; eval[o] = val[s + e]; }
    else { ecol[o] = r; eval[o] = 0.0f; }
  }
  tail_cnt[r] = len > width ? len - width : 0;
}

__global__ void build_tail_kernel(int n, int width, const int32_t *__restrict__ rowptr,
                                  const int32_t *__restrict__ col, const float *__restrict__ val,
                                  const int32_t *__restrict__ tail_off, int32_t *__restrict__ trow,
                                  int32_t *__restrict__ tcol, float *__restrict__ tval) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = rowptr[r], len = rowptr[r + 1] - s;
  int o = tail_off[r];
  for (int e = width; e < len; ++e, ++o) {
    trow[o] = r;
    tcol[o] = col[s + e];
    tval[o] = val[s + e];
  }
}

// ---- symmetric permutation of P: new row r = old row perm[r], columns
// relabelled by iperm (locality order; the values are untouched) -----------
__global__ void perm_len_kernel(int n, const int32_t *__restrict__ perm, const int32_t *__restrict__ rp,
                                int32_t *__restrict__ len) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int o = perm[r];
  len[r] = rp[o + 1] - rp[o];
}

__global__ void perm_fill_kernel(int n, const int32_t *__restrict__ perm, const int32_t *__restrict__ iperm,
                                 const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                 const float *__restrict__ val, const int32_t *__restrict__ rp2,
                                 int32_t *__restrict__ col2, float *__restrict__ val2) {
  // one wavefront per new row
  const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  const int o = perm[r];
  const int s0 = rp[o], len = rp[o + 1] - s0, d0 = rp2[r];
  for (int e = lane; e < len; e += 64) {
    col2[d0 + e] = iperm[col[s0 + e]];
    val2[d0 + e] = val[s0 + e];
  }
}

// Columns ascending within each row, in place, on the device alone: one
// work-group per row sorts its (column, value) pairs with a bitonic sort in
// LDS.  Rows longer than kRowSortMax keep their order (the order of a row
// only changes which lanes gather neighbouring points, never the result's
// definition), so nothing here needs the host.
constexpr int kRowSortMax = 4096;

__global__ __launch_bounds__(256) void row_sort_kernel(int n, const int32_t *__restrict__ rp, int32_t *__restrict__ col,
                                                       float *__restrict__ val) {
  __shared__ int32_t kc[kRowSortMax];
  __shared__ float kv[kRowSortMax];
  const int r = blockIdx.x;
  if (r >= n) return;
  const int b = rp[r], len = rp[r + 1] - b;
  if (len < 2 || len > kRowSortMax) return;  // the whole work-group leaves together
  int P = 2;
  while (P < len) P <<= 1;
  for (int t = threadIdx.x; t < P; t += blockDim.x) {
    kc[t] = t < len ? col[b + t] : INT_MAX;
    kv[t] = t < len ? val[b + t] : 0.0f;
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < P / 2; t += blockDim.x) {
        const int i = 2 * stride * (t / stride) + (t % stride), j = i + stride;
        const bool up = (i & size) == 0;
        const int ci = kc[i], cj = kc[j];
        if ((ci > cj) == up) {
          kc[i] = cj;
          kc[j] = ci;
          const float v = kv[i];
          kv[i] = kv[j];
          kv[j] = v;
        }
      }
      __syncthreads();
    }
  for (int t = threadIdx.x; t < len; t += blockDim.x) {
    col[b + t] = kc[t];
    val[b + t] = kv[t];
  }
}

// ---- spatial order of the embedding -------------------------------------------
// Morton key of each point's position on a 2^B-per-axis lattice over the
// bounding box (B = 16 / 10 / 32 bits per axis for d = 2 / 3 / 1).
__device__ __forceinline__ uint32_t spread_bits2(uint32_t x) {
  x &= 0xffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__device__ __forceinline__ uint32_t spread_bits3(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ void morton_kernel(int n, int d, int ds, const float *__restrict__ Y, float lo0, float lo1, float lo2,
                              float inv, uint32_t *__restrict__ key, int32_t *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float lo[3] = {lo0, lo1, lo2};
  uint32_t q[3] = {0, 0, 0};
  const float top = d == 1 ? 4294967040.0f : d == 2 ? 65535.0f : 1023.0f;
  for (int a = 0; a < d; ++a) {
    float t = (Y[(size_t)i * ds + a] - lo[a]) * inv * top;
    t = fminf(fmaxf(t, 0.0f), top);  // NaN -> 0
    q[a] = (uint32_t)t;
  }
  uint32_t k;
  if (d == 1) k = q[0];
  else if (d == 2) k = spread_bits2(q[0]) | (spread_bits2(q[1]) << 1);
  else k = spread_bits3(q[0]) | (spread_bits3(q[1]) << 1) | (spread_bits3(q[2]) << 2);
  key[i] = k;
  idx[i] = i;
}

__global__ void invert_perm_kernel(int n, const int32_t *__restrict__ order, int32_t *__restrict__ inv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) inv[order[r]] = r;
}

__global__ void gather_points_kernel(int n, int ds, const int32_t *__restrict__ order, const float *__restrict__ in,
                                     float *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * ds) return;
  const int r = (int)(t / ds), a = (int)(t % ds);
  out[t] = in[(int64_t)order[r] * ds + a];
}

__global__ void to_double_kernel(const float *__restrict__ in, double *__restrict__ out, int64_t m) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) out[t] = (double)in[t];
}

}  // namespace

namespace {

// Exact kNN of the query slots [0, nq) (points qlist[slot], or all points).
int knn_exact_rows(hipStream_t s, const double *X, int n, int dim, int k, const int *qlist, int nq, int32_t *I,
                   double *D) {
  double *hd = nullptr;
  int *hi = nullptr;
  SGT_HIP(hipMallocAsync((void **)&hd, sizeof(double) * (size_t)nq * k, s));
  SGT_HIP(hipMallocAsync((void **)&hi, sizeof(int) * (size_t)nq * k, s));
  knn_kernel<<<div_up(nq, kKnnQ), kKnnThreads, 0, s>>>(X, n, dim, k, hd, hi, qlist, nq);
  SGT_LAUNCH_CHECK();
  knn_finalize_kernel<<<div_up(nq, 256), 256, 0, s>>>(nq, k, hd, hi, qlist, I, D);
  SGT_LAUNCH_CHECK();
  SGT_HIP(hipFreeAsync(hd, s));
  SGT_HIP(hipFreeAsync(hi, s));
  return SGTSNE_OK;
}

}  // namespace

int knn_begin(hipStream_t s, const double *X, int n, int dim, int k, int32_t *I, double *D, int method,
              int *count_host, KnnPending *pend) {
  *pend = KnnPending{};
  SGT_CHECK(n > k && k >= 1 && dim >= 1, SGTSNE_ERR_ARG, "kNN needs n > k >= 1 (n=%d k=%d)", n, k);
  SGT_CHECK(k <= kKnnR * 64, SGTSNE_ERR_ARG, "k too large");
  const bool filter_ok = k + kFiltSlack <= kRerank && dim <= kFiltMaxDp && n >= kFiltQ;
  if (method == SGTSNE_KNN_EXACT_FP64 || (method == SGTSNE_KNN_AUTO && !filter_ok))
    return knn_exact_rows(s, X, n, dim, k, nullptr, n, I, D);  // nothing left for knn_end
  SGT_CHECK(filter_ok, SGTSNE_ERR_ARG, "the filtered kNN needs k <= %d, dim <= %d and n >= %d",
            kRerank - kFiltSlack, kFiltMaxDp, kFiltQ);
  const int kp = std::min(k + kFiltSlack, n - 1);
  const int nst = filter_steps(dim, kp), dp = 4 * nst;
  const bool blocked = kp <= kF2MaxKp;  // the LDS-heap filter centres each query block
  const int nb = div_up(n, kF2Q);
  double *mu = nullptr, *nrm = nullptr, *cb = nullptr, *dn = nullptr, *qa = nullptr, *qdn = nullptr;
  float *Xc = nullptr, *n32 = nullptr, *hd = nullptr, *Xq = nullptr, *qn32 = nullptr, *delta = nullptr;
  int *hi = nullptr, *fb = nullptr;
  int32_t *kidx = nullptr, *qorder = nullptr, *cell = nullptr, *rid = nullptr, *rinv = nullptr;
  int *ccnt = nullptr;  // per cell: count | first sorted position | first slot
  float *Xr = nullptr;
  double *tcen = nullptr, *trad = nullptr;
  uint32_t *key0 = nullptr, *key1 = nullptr;
  void *tmp = nullptr;
  unsigned long long *rmax = nullptr;
  SGT_HIP(hipMallocAsync((void **)&mu, sizeof(double) * dim, s));
  SGT_HIP(hipMallocAsync((void **)&nrm, sizeof(double) * (size_t)n, s));
  SGT_HIP(hipMallocAsync((void **)&Xc, sizeof(float) * (size_t)n * dp, s));
  SGT_HIP(hipMallocAsync((void **)&n32, sizeof(float) * (size_t)n, s));
  SGT_HIP(hipMallocAsync((void **)&hd, sizeof(float) * (size_t)n * kp, s));
  SGT_HIP(hipMallocAsync((void **)&hi, sizeof(int) * (size_t)n * kp, s));
  SGT_HIP(hipMallocAsync((void **)&fb, sizeof(int) * ((size_t)n + 1), s));  // [0]: count, then rows
  SGT_HIP(hipMallocAsync((void **)&rmax, sizeof(unsigned long long), s));
  SGT_HIP(hipMallocAsync((void **)&qdn, sizeof(double) * (size_t)n, s));
  SGT_HIP(hipMemsetAsync(rmax, 0, sizeof(unsigned long long), s));
  SGT_HIP(hipMemsetAsync(fb, 0, sizeof(int), s));
  knn_mean_kernel<<<dim, 256, 0, s>>>(X, n, dim, mu);
  SGT_LAUNCH_CHECK();
  knn_center_kernel<<<div_up(n, 256), 256, 0, s>>>(X, n, dim, dp, mu, Xc, n32, nrm, rmax);
  SGT_LAUNCH_CHECK();
  BlockCentred bc{};
  if (blocked) {
    SGT_HIP(hipMallocAsync((void **)&key0, sizeof(uint32_t) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&key1, sizeof(uint32_t) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&kidx, sizeof(int32_t) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&qorder, sizeof(int32_t) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&cb, sizeof(double) * (size_t)nb * dim, s));
    SGT_HIP(hipMallocAsync((void **)&delta, sizeof(float) * (size_t)nb * dp, s));
    SGT_HIP(hipMallocAsync((void **)&dn, sizeof(double) * (size_t)nb, s));
    SGT_HIP(hipMallocAsync((void **)&Xq, sizeof(float) * (size_t)n * dp, s));
    SGT_HIP(hipMallocAsync((void **)&qn32, sizeof(float) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&qa, sizeof(double) * (size_t)n, s));
    // point order: cell (nearest of P pivots), then the Morton code of a 3-D
    // projection of the centred points; the query blocks and the reference
    // tiles both follow it, the tiles padded so each lies in one cell
    const int P = std::min(kKnnMaxCells, std::max(1, n / 512));
    const int ntiles = div_up(n + (kFiltR - 1) * P, kFiltR);  // at least the padded slots
    SGT_HIP(hipMallocAsync((void **)&cell, sizeof(int32_t) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&ccnt, sizeof(int) * 3 * kKnnMaxCells, s));
    SGT_HIP(hipMallocAsync((void **)&rid, sizeof(int32_t) * (size_t)ntiles * kFiltR, s));
    SGT_HIP(hipMallocAsync((void **)&rinv, sizeof(int32_t) * (size_t)n, s));
    SGT_HIP(hipMallocAsync((void **)&Xr, sizeof(float) * (size_t)ntiles * kFiltR * dp, s));
    SGT_HIP(hipMallocAsync((void **)&tcen, sizeof(double) * (size_t)ntiles * dim, s));
    SGT_HIP(hipMallocAsync((void **)&trad, sizeof(double) * (size_t)ntiles, s));
    knn_cell_kernel<<<div_up(n, 256), 256, sizeof(float) * P * dp, s>>>(Xc, n, dp, P, cell);
    SGT_LAUNCH_CHECK();
    knn_morton_key_kernel<<<div_up(n, 256), 256, 0, s>>>(Xc, n, dim, dp, rmax, cell, key0, kidx);
    SGT_LAUNCH_CHECK();
    size_t tb = 0;
    SGT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key0, key1, kidx, qorder, n, 0, 30, s));
    SGT_HIP(hipMallocAsync(&tmp, tb, s));
    SGT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key0, key1, kidx, qorder, n, 0, 30, s));
    SGT_HIP(hipMemsetAsync(ccnt, 0, sizeof(int) * kKnnMaxCells, s));
    knn_cell_count_kernel<<<std::min(div_up(n, 256), 1024), 256, 0, s>>>(cell, n, ccnt);
    SGT_LAUNCH_CHECK();
    knn_cell_offsets_kernel<<<1, 64, 0, s>>>(ccnt, P, ccnt + kKnnMaxCells, ccnt + 2 * kKnnMaxCells);
    SGT_LAUNCH_CHECK();
    SGT_HIP(hipMemsetAsync(rid, 0xFF, sizeof(int32_t) * (size_t)ntiles * kFiltR, s));  // -1: padding
    SGT_HIP(hipMemsetAsync(Xr, 0, sizeof(float) * (size_t)ntiles * kFiltR * dp, s));
    knn_ref_scatter_kernel<<<div_up(n, 256), 256, 0, s>>>(Xc, n, dp, qorder, cell, ccnt + kKnnMaxCells,
                                                          ccnt + 2 * kKnnMaxCells, Xr, rid, rinv);
    SGT_LAUNCH_CHECK();
    knn_tile_sphere_kernel<<<ntiles, 64, 0, s>>>(X, dim, rid, tcen, trad);
    SGT_LAUNCH_CHECK();
    knn_block_centre_kernel<<<nb, 256, 0, s>>>(X, n, dim, dp, qorder, mu, cb, delta, dn, kF2Q);
    SGT_LAUNCH_CHECK();
    knn_query_prep_kernel<<<div_up(n, 256), 256, 0, s>>>(X, n, dim, dp, qorder, cb, dn, Xq, qn32, qa, qdn, kF2Q);
    SGT_LAUNCH_CHECK();
    bc = BlockCentred{Xq, qn32, qorder, delta,
                      RefTiles{Xr, rid, rinv, ntiles, tcen, trad, cb, qa, dn, rmax, dim, cell, ccnt + 2 * kKnnMaxCells}};
  } else {
    SGT_HIP(hipMemsetAsync(qdn, 0, sizeof(double) * (size_t)n, s));  // frame of every query: mu
  }
  SGT_TRY(launch_filter_for(s, nst, Xc, n32, n, kp, bc, hd, hi));
  // error constants (unit roundoff u = 2^-24): the squared distance from the
  // filter is within (2 dp + 3) u (|a| + |b|)^2 of that of the FP32 vectors
  // under round-to-nearest; allowing twice that covers directed rounding in
  // the accumulation, and 1e-30 covers underflow.  The FP32 roundings of the
  // centred points move a distance by at most 1.01 u (|a| + R + |c - mu| + |b|).
  const double u = std::ldexp(1.0, -24);
  const double cE = 2.0 * (2.0 * dp + 8.0) * u, ce = 1.05 * u;
  knn_rerank_kernel<<<n, kRerank, 0, s>>>(X, n, dim, k, kp, hd, hi, blocked ? qa : nrm, qdn, rmax, cE, ce, I, D,
                                          fb + 1, fb);
  SGT_LAUNCH_CHECK();
  SGT_HIP(hipMemcpyAsync(count_host, fb, sizeof(int), hipMemcpyDeviceToHost, s));
  void *bufs[] = {mu, nrm, Xc, n32, hd, hi, fb, rmax, qdn, key0, key1, kidx, qorder, cb, delta, dn, Xq, qn32, qa, tmp,
                  cell, ccnt, rid, rinv, Xr, tcen, trad};
  static_assert(sizeof(bufs) / sizeof(bufs[0]) <= sizeof(pend->bufs) / sizeof(pend->bufs[0]), "KnnPending::bufs");
  pend->open = true;
  pend->X = X;
  pend->n = n;
  pend->dim = dim;
  pend->k = k;
  pend->I = I;
  pend->D = D;
  pend->fb = fb;
  pend->count_host = count_host;
  for (void *p : bufs)
    if (p) pend->bufs[pend->nbufs++] = p;
  return SGTSNE_OK;
}

void knn_release(hipStream_t s, KnnPending *kp) {
  for (int i = 0; i < kp->nbufs; ++i) (void)hipFreeAsync(kp->bufs[i], s);
  *kp = KnnPending{};
}

int knn_end(hipStream_t s, KnnPending *kp, int *fallback_rows) {
  if (!kp->open) {  // exact rows throughout
    if (fallback_rows) *fallback_rows = -1;
    return SGTSNE_OK;
  }
  const int nfb = *kp->count_host;
  int st = SGTSNE_OK;
  if (nfb > 0) st = knn_exact_rows(s, kp->X, kp->n, kp->dim, kp->k, kp->fb + 1, nfb, kp->I, kp->D);
  if (fallback_rows) *fallback_rows = nfb;
  knn_release(s, kp);
  return st;
}

int launch_knn(hipStream_t s, const double *X, int n, int dim, int k, int32_t *I, double *D, int method,
               int *fallback_rows) {
  if (fallback_rows) *fallback_rows = -1;
  int nfb = 0;
  KnnPending kp;
  SGT_TRY(knn_begin(s, X, n, dim, k, I, D, method, &nfb, &kp));
  if (kp.open && hipStreamSynchronize(s) != hipSuccess) {
    knn_release(s, &kp);
    set_error("kNN failed");
    return SGTSNE_ERR_HIP;
  }
  return knn_end(s, &kp, fallback_rows);
}

int launch_knn_exact(hipStream_t s, const double *X, int n, int dim, int k, int32_t *I, double *D) {
  return launch_knn(s, X, n, dim, k, I, D, SGTSNE_KNN_AUTO, nullptr);
}

int launch_perplexity(hipStream_t s, const double *D, int n, int k, double u, float *val,
                      double *beta_out) {
  perplexity_kernel2<<<(int)div_up((int64_t)n * kPerpLanes, 256), 256, 0, s>>>(D, n, k, u, val, beta_out);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_lambda_rescale(hipStream_t s, int n, const int32_t *rowptr, float *val, double lambda,
                          int drop_leaf) {
  lambda_kernel<<<div_up(n, 128), 128, 0, s>>>(n, rowptr, val, lambda, drop_leaf);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_rowids(hipStream_t s, int n, const int32_t *rowptr, int32_t *row) {
  rowids_kernel<<<div_up(n, 256), 256, 0, s>>>(n, rowptr, row);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

// Keeps memory freed to the device's stream-ordered pool cached in the pool:
// with the default release threshold (0) every synchronisation returns it to
// the driver, and the next hipMallocAsync blocks the host to map it again.
// Cross-stream reuse is limited to what stream order and recorded events
// prove: a block freed on one stream is handed to another only through an
// event dependency the caller made, never because the runtime saw the free
// complete (opportunistic) or inserted a wait of its own (internal
// dependencies).
void retain_pool_memory() {
  int dev = 0;
  hipMemPool_t pool = nullptr;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
  uint64_t keep = UINT64_MAX;
  (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  int off = 0;
  (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowOpportunistic, &off);
  (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &off);
}

int device_exclusive_scan(hipStream_t s, const int32_t *in, int32_t *out, int n) {
  size_t tb = 0;
  SGT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, s));
  void *tmp = nullptr;
  SGT_HIP(hipMallocAsync(&tmp, tb > 0 ? tb : 1, s));
  SGT_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, s));
  SGT_HIP(hipFreeAsync(tmp, s));
  return SGTSNE_OK;
}

int sym_begin(hipStream_t s, int n, int64_t nnz, const int32_t *rowptr_d, const int32_t *col_d,
              const float *val_d, int *runs_host, SymPending *sp) {
  *sp = SymPending{};
  const int64_t m2 = 2 * nnz;
  SGT_CHECK(m2 < (int64_t)INT_MAX, SGTSNE_ERR_ARG, "2*nnz exceeds the int32 range");
  sp->n = n;
  sp->m2 = m2;
  sp->runs_host = runs_host;
  SGT_HIP(hipMallocAsync((void **)&sp->k0, sizeof(unsigned long long) * m2, s));
  SGT_HIP(hipMallocAsync((void **)&sp->k1, sizeof(unsigned long long) * m2, s));
  SGT_HIP(hipMallocAsync((void **)&sp->v0, sizeof(double) * m2, s));
  SGT_HIP(hipMallocAsync((void **)&sp->v1, sizeof(double) * m2, s));
  SGT_HIP(hipMallocAsync((void **)&sp->nruns, sizeof(int), s));
  SGT_HIP(hipMallocAsync((void **)&sp->total, sizeof(double), s));
  sp->open = true;
  sym_keys_kernel<<<div_up(n, 256), 256, 0, s>>>(n, rowptr_d, col_d, val_d, nnz, sp->k0, sp->v0);
  SGT_LAUNCH_CHECK();
  int nbits = 1;
  while ((1ll << nbits) < (long long)n) ++nbits;
  const int end_bit = 32 + nbits;
  size_t tb = 0;
  SGT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, sp->k0, sp->k1, sp->v0, sp->v1, (int)m2, 0, end_bit, s));
  void *tmp = nullptr;
  SGT_HIP(hipMallocAsync(&tmp, tb > 0 ? tb : 1, s));
  SGT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, sp->k0, sp->k1, sp->v0, sp->v1, (int)m2, 0, end_bit, s));
  SGT_HIP(hipFreeAsync(tmp, s));
  // unique keys into k0 / v0 (reuse), summing duplicates (P_ij + P_ji)
  tb = 0;
  SGT_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, tb, sp->k1, sp->k0, sp->v1, sp->v0, sp->nruns, hipcub::Sum(),
                                            (int)m2, s));
  SGT_HIP(hipMallocAsync(&tmp, tb > 0 ? tb : 1, s));
  SGT_HIP(hipcub::DeviceReduce::ReduceByKey(tmp, tb, sp->k1, sp->k0, sp->v1, sp->v0, sp->nruns, hipcub::Sum(),
                                            (int)m2, s));
  SGT_HIP(hipFreeAsync(tmp, s));
  SGT_HIP(hipMemcpyAsync(runs_host, sp->nruns, sizeof(int), hipMemcpyDeviceToHost, s));
  return SGTSNE_OK;
}

void sym_release(hipStream_t s, SymPending *sp) {
  void *bufs[] = {sp->k0, sp->k1, sp->v0, sp->v1, sp->nruns, sp->total};
  for (void *p : bufs)
    if (p) (void)hipFreeAsync(p, s);
  *sp = SymPending{};
}

int sym_end(hipStream_t s, SymPending *sp, int32_t **rowptr_out, int32_t **col_out, float **val_out,
            int64_t *nnz_out, bool stream_ordered) {
  SGT_CHECK(sp->open, SGTSNE_ERR_STATE, "no symmetrisation pending");
  const int n = sp->n;
  const int runs = *sp->runs_host;
  const unsigned long long *uk = sp->k0;
  const double *uv = sp->v0;
  size_t tb = 0;
  void *tmp = nullptr;
  SGT_HIP(hipcub::DeviceReduce::Sum(nullptr, tb, uv, sp->total, runs, s));
  SGT_HIP(hipMallocAsync(&tmp, tb > 0 ? tb : 1, s));
  SGT_HIP(hipcub::DeviceReduce::Sum(tmp, tb, uv, sp->total, runs, s));
  SGT_HIP(hipFreeAsync(tmp, s));
  int32_t *rp = nullptr, *cc = nullptr;
  float *vv = nullptr;
  // the outputs come from the stream-ordered pool when the caller releases
  // them with hipFreeAsync, else from hipMalloc (released with hipFree)
  if (stream_ordered) {
    SGT_HIP(hipMallocAsync((void **)&rp, sizeof(int32_t) * ((size_t)n + 1), s));
    SGT_HIP(hipMallocAsync((void **)&cc, sizeof(int32_t) * (size_t)(runs > 0 ? runs : 1), s));
    SGT_HIP(hipMallocAsync((void **)&vv, sizeof(float) * (size_t)(runs > 0 ? runs : 1), s));
  } else {
    SGT_HIP(hipMalloc((void **)&rp, sizeof(int32_t) * ((size_t)n + 1)));
    SGT_HIP(hipMalloc((void **)&cc, sizeof(int32_t) * (size_t)(runs > 0 ? runs : 1)));
    SGT_HIP(hipMalloc((void **)&vv, sizeof(float) * (size_t)(runs > 0 ? runs : 1)));
  }
  if (runs > 0) {
    sym_finish_kernel<<<div_up(runs, 256), 256, 0, s>>>(runs, uk, uv, sp->total, cc, vv);
    SGT_LAUNCH_CHECK();
  }
  sym_rowptr_kernel<<<div_up((int64_t)n + 1, 256), 256, 0, s>>>(n, runs, uk, rp);
  SGT_LAUNCH_CHECK();
  sym_release(s, sp);
  *rowptr_out = rp;
  *col_out = cc;
  *val_out = vv;
  *nnz_out = runs;
  return SGTSNE_OK;
}

int gpu_symmetrize(hipStream_t s, int n, int64_t nnz, const int32_t *rowptr_d, const int32_t *col_d,
                   const float *val_d, int32_t **rowptr_out, int32_t **col_out, float **val_out,
                   int64_t *nnz_out, bool stream_ordered) {
  int runs = 0;
  SymPending sp;
  SGT_TRY(sym_begin(s, n, nnz, rowptr_d, col_d, val_d, &runs, &sp));
  SGT_HIP(hipStreamSynchronize(s));
  SGT_TRY(sym_end(s, &sp, rowptr_out, col_out, val_out, nnz_out, stream_ordered));
  SGT_HIP(hipStreamSynchronize(s));
  return SGTSNE_OK;
}

int launch_permute_csr(hipStream_t s, int n, int64_t nnz, const int32_t *perm, const int32_t *iperm,
                       const int32_t *rp, const int32_t *col, const float *val, int32_t *rp2,
                       int32_t *col2, float *val2, bool sort_rows, bool device_sort) {
  int32_t *len = nullptr;
  SGT_HIP(hipMallocAsync((void **)&len, sizeof(int32_t) * ((size_t)n + 1), s));
  perm_len_kernel<<<div_up(n, 256), 256, 0, s>>>(n, perm, rp, len);
  SGT_LAUNCH_CHECK();
  SGT_HIP(hipMemsetAsync(len + n, 0, sizeof(int32_t), s));
  SGT_TRY(device_exclusive_scan(s, len, rp2, n + 1));
  if (!sort_rows || device_sort) {
    perm_fill_kernel<<<div_up((int64_t)n * 64, 256), 256, 0, s>>>(n, perm, iperm, rp, col, val, rp2, col2, val2);
    SGT_LAUNCH_CHECK();
    if (sort_rows) {  // in place, without the host (rows above kRowSortMax keep their order)
      row_sort_kernel<<<n, 256, 0, s>>>(n, rp2, col2, val2);
      SGT_LAUNCH_CHECK();
    }
    SGT_HIP(hipFreeAsync(len, s));
    return SGTSNE_OK;
  }
  int32_t *ctmp = nullptr;
  float *vtmp = nullptr;
  SGT_HIP(hipMallocAsync((void **)&ctmp, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1), s));
  SGT_HIP(hipMallocAsync((void **)&vtmp, sizeof(float) * (size_t)std::max<int64_t>(nnz, 1), s));
  perm_fill_kernel<<<div_up((int64_t)n * 64, 256), 256, 0, s>>>(n, perm, iperm, rp, col, val, rp2, ctmp, vtmp);
  SGT_LAUNCH_CHECK();
  // columns ascending within each row: neighbouring lanes gather neighbouring points
  int nbits = 1;
  while ((1ll << nbits) < (long long)n) ++nbits;
  size_t tb = 0;
  SGT_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb, ctmp, col2, vtmp, val2, (int)nnz, n, rp2,
                                                      rp2 + 1, 0, nbits, s));
  void *tmp = nullptr;
  SGT_HIP(hipMallocAsync(&tmp, tb > 0 ? tb : 1, s));
  SGT_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp, tb, ctmp, col2, vtmp, val2, (int)nnz, n, rp2,
                                                      rp2 + 1, 0, nbits, s));
  SGT_HIP(hipFreeAsync(tmp, s));
  SGT_HIP(hipFreeAsync(ctmp, s));
  SGT_HIP(hipFreeAsync(vtmp, s));
  SGT_HIP(hipFreeAsync(len, s));
  return SGTSNE_OK;
}

int spatial_order(hipStream_t s, int n, int d, int ds, const float *Y, const double *lo, double extent,
                  int32_t *order, int32_t *inv) {
  uint32_t *k0 = nullptr, *k1 = nullptr;
  int32_t *i0 = nullptr;
  SGT_HIP(hipMallocAsync((void **)&k0, sizeof(uint32_t) * (size_t)n, s));
  SGT_HIP(hipMallocAsync((void **)&k1, sizeof(uint32_t) * (size_t)n, s));
  SGT_HIP(hipMallocAsync((void **)&i0, sizeof(int32_t) * (size_t)n, s));
  const float invx = extent > 0 ? (float)(1.0 / extent) : 0.0f;
  morton_kernel<<<div_up(n, 256), 256, 0, s>>>(n, d, ds, Y, (float)lo[0], d > 1 ? (float)lo[1] : 0.f,
                                                d > 2 ? (float)lo[2] : 0.f, invx, k0, i0);
  SGT_LAUNCH_CHECK();
  size_t tb = 0;
  SGT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, i0, order, n, 0, 32, s));
  void *tmp = nullptr;
  SGT_HIP(hipMallocAsync(&tmp, tb > 0 ? tb : 1, s));
  SGT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, order, n, 0, 32, s));
  SGT_HIP(hipFreeAsync(tmp, s));
  invert_perm_kernel<<<div_up(n, 256), 256, 0, s>>>(n, order, inv);
  SGT_LAUNCH_CHECK();
  SGT_HIP(hipFreeAsync(k0, s));
  SGT_HIP(hipFreeAsync(k1, s));
  SGT_HIP(hipFreeAsync(i0, s));
  return SGTSNE_OK;
}

int launch_gather_points(hipStream_t s, int n, int ds, const int32_t *order, const float *in, float *out) {
  gather_points_kernel<<<div_up((int64_t)n * ds, 256), 256, 0, s>>>(n, ds, order, in, out);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_build_ell(hipStream_t s, int n, int width, const int32_t *rowptr, const int32_t *col,
                     const float *val, int32_t *ecol, float *eval, int32_t *tail_cnt) {
  build_ell_kernel<<<div_up(n, 256), 256, 0, s>>>(n, width, rowptr, col, val, ecol, eval, tail_cnt);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_build_tail(hipStream_t s, int n, int width, const int32_t *rowptr, const int32_t *col,
                      const float *val, const int32_t *tail_off, int32_t *trow, int32_t *tcol,
                      float *tval) {
  build_tail_kernel<<<div_up(n, 256), 256, 0, s>>>(n, width, rowptr, col, val, tail_off, trow, tcol, tval);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

}  // namespace sgt
