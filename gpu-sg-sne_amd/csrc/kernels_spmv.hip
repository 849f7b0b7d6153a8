// kernels_spmv.hip -- attractive term Fattr_i = sum_j P_ij (y_i - y_j) / (1 + |y_i - y_j|^2)
// over the symmetric P in four storage formats, and the KL cost.
//
// Reference: AttractiveEstimation CU/pq.cu:25-57 with pq_csr_naive
// (CU/pq.cuh:16-75, format 0), coo_pq_kernel (CU/pq.cuh:77-118, format 2),
// ell_spmv_kernel + coo_spmv_kernel (CU/hybridpq.cu:4-113, format 3); CPU
// HY/pq.cpp:15-57.  KL: ComputeKLkernel CU/pq.cu:61-102.
//
// Y is AoS with one aligned vector per point (float / float2 / float4), so
// every neighbour gather is a single load.
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/sgtsne.h"

#include <algorithm>
#include <cfloat>
#include <type_traits>

namespace sgt {

namespace {

constexpr int kBlock = 256;

// The pair term with every rounding spelled out (explicit fused multiply-adds,
// the first square rounded on its own): the compiler's choice of which
// product to fuse would otherwise depend on the code around each inlined copy,
// and the kernel forms, which add the same entries in the same order, would
// differ in the last bit.
template <int D>
__device__ __forceinline__ void accum_pair(const float (&yi)[D], const float (&yj)[D], float p,
                                           float (&acc)[D]) {
  float df[D];
#pragma unroll
  for (int a = 0; a < D; ++a) df[a] = yi[a] - yj[a];
  float dist = df[0] * df[0];
#pragma unroll
  for (int a = 1; a < D; ++a) dist = __builtin_fmaf(df[a], df[a], dist);
  const float pq = p / (1.0f + dist);
#pragma unroll
  for (int a = 0; a < D; ++a) acc[a] = __builtin_fmaf(pq, df[a], acc[a]);
}

// Format 4 (default): a group of G lanes per row walks the row's contiguous
// (col, val) stream; the group sum is a cross-lane butterfly.
template <int D, int G>
__global__ __launch_bounds__(kBlock) void attr_csr_seg_kernel(int n, const int32_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ col,
                                                              const float *__restrict__ val,
                                                              const float *__restrict__ Y,
                                                              float *__restrict__ F) {
  const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  const int ngroups = gridDim.x * blockDim.x / G;
  for (int row = gid; row < n; row += ngroups) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    const int e = rowptr[row + 1];
    for (int t = rowptr[row] + lane; t < e; t += G) {
      float yj[D];
      load_pt<D>(Y, col[t], yj);
      accum_pair<D>(yi, yj, val[t], acc);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, G);
    if (lane == 0) store_pt<D>(F, row, acc);
  }
}

// One lane's share of a row: entries t0, t0 + G, ... < e, added in that
// order.  U >= 8: masked steps, a lane's next U entries (short rows: all of
// them) have their column and value loads, then their gathers, in flight
// together.  U < 8: U strides per step, then the remainder one by one.
// (col, val) are streamed once per iteration with nontemporal loads and do
// not evict Y from the caches.
template <int D, int G, int U, class Gather>
__device__ __forceinline__ void row_accum_g(int t, const int e, const int32_t *__restrict__ col,
                                            const float *__restrict__ val, Gather gather,
                                            const float (&yi)[D], float (&acc)[D]) {
  if constexpr (U >= 8) {
    for (; t < e; t += U * G) {
      int c[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool in = t + u * G < e;
        c[u] = in ? __builtin_nontemporal_load(col + t + u * G) : -1;
        v[u] = in ? __builtin_nontemporal_load(val + t + u * G) : 0.0f;
      }
      float yj[U][D];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c[u] >= 0) gather(c[u], yj[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c[u] >= 0) accum_pair<D>(yi, yj[u], v[u], acc);
    }
  } else {
    for (; t + (U - 1) * G < e; t += U * G) {
      int c[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(col + t + u * G);
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(val + t + u * G);
      float yj[U][D];
#pragma unroll
      for (int u = 0; u < U; ++u) gather(c[u], yj[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) accum_pair<D>(yi, yj[u], v[u], acc);
    }
    for (; t < e; t += G) {
      float yj[D];
      gather(__builtin_nontemporal_load(col + t), yj);
      accum_pair<D>(yi, yj, __builtin_nontemporal_load(val + t), acc);
    }
  }
}

// The plain form: every neighbour from Y (global memory).
template <int D, int G, int U>
__device__ __forceinline__ void row_accum(int t, const int e, const int32_t *__restrict__ col,
                                          const float *__restrict__ val, const float *__restrict__ Y,
                                          const float (&yi)[D], float (&acc)[D]) {
  row_accum_g<D, G, U>(t, e, col, val, [&](int c, float (&yj)[D]) { load_pt<D>(Y, c, yj); }, yi, acc);
}

// Format 4, tuned form: as above, but each lane issues the (col, val) loads of
// U strides before the U neighbour gathers, so U independent loads per lane
// are in flight; (col, val) are streamed once per iteration with nontemporal
// loads and do not evict Y from the caches.
template <int D, int G, int U>
__global__ __launch_bounds__(kBlock) void attr_csr_seg2_kernel(int n, const int32_t *__restrict__ rowptr,
                                                               const int32_t *__restrict__ col,
                                                               const float *__restrict__ val,
                                                               const float *__restrict__ Y,
                                                               float *__restrict__ F) {
  const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  const int ngroups = gridDim.x * blockDim.x / G;
  for (int row = gid; row < n; row += ngroups) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    row_accum<D, G, U>(rowptr[row] + lane, rowptr[row + 1], col, val, Y, yi, acc);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, G);
    if (lane == 0) store_pt<D>(F, row, acc);
  }
}

// Format 4, blocked form: each work-group owns a contiguous range of rows,
// so the neighbour gathers of one CU cover one window of the (locality-
// ordered) points and hit its L1; ranges are dealt to XCDs in contiguous
// runs (work-groups are placed round-robin over the 8 XCDs) so that
// neighbouring ranges share an L2.
constexpr int kXcds = 8;

template <int D, int G, int U>
__global__ __launch_bounds__(kBlock) void attr_csr_blk_kernel(int n, int rows_per_block, int nranges,
                                                              const int32_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ col,
                                                              const float *__restrict__ val,
                                                              const float *__restrict__ Y,
                                                              float *__restrict__ F) {
  const int per_xcd = (gridDim.x + kXcds - 1) / kXcds;
  const int range = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (range >= nranges) return;
  const int r0 = range * rows_per_block;
  const int r1 = min(n, r0 + rows_per_block);
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  constexpr int NG = kBlock / G;
  for (int row = r0 + grp; row < r1; row += NG) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    row_accum<D, G, U>(rowptr[row] + lane, rowptr[row + 1], col, val, Y, yi, acc);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, G);
    if (lane == 0) store_pt<D>(F, row, acc);
  }
}

// Format 4, chunk-cached form: as the blocked form, with 1024-thread work-
// groups (one per CU) that first copy into LDS the chunks of Y (runs of 2^chs
// consecutive points) their rows' neighbours fall in most often, chosen per
// row range at upload (chunk_select_kernel), plus a chunk -> slot table.  A
// neighbour in a cached chunk is read from LDS, any other from Y.  Each lane
// sums the same entries in the same order as the other forms with G lanes per
// row (only the source of y_j changes), so the result is theirs bit for bit.
constexpr int kCcBlock = 1024;
constexpr size_t kCcLdsMax = 147456;  // cached points + slot table

template <int D, int G, int U>
__global__ __launch_bounds__(kCcBlock) void attr_csr_cc_kernel(int n, int rows_per_block, int nranges, int chs,
                                                               int nslots, int nchunks,
                                                               const int32_t *__restrict__ slots,
                                                               const int32_t *__restrict__ rowptr,
                                                               const int32_t *__restrict__ col,
                                                               const float *__restrict__ val,
                                                               const float *__restrict__ Y,
                                                               float *__restrict__ F) {
  using V = typename VecT<D>::T;
  extern __shared__ float4 craw[];
  V *cy = reinterpret_cast<V *>(craw);
  const int cpts = nslots << chs;
  int8_t *lut = reinterpret_cast<int8_t *>(cy + cpts);
  const V *__restrict__ Yv = reinterpret_cast<const V *>(Y);
  const int per_xcd = (gridDim.x + kXcds - 1) / kXcds;
  const int range = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (range >= nranges) return;  // the whole work-group leaves together
  const int32_t *__restrict__ sl = slots + (int64_t)range * nslots;
  for (int i = threadIdx.x; i < nchunks; i += kCcBlock) lut[i] = -1;
  __syncthreads();
  if ((int)threadIdx.x < nslots) {
    const int ch = sl[threadIdx.x];
    if (ch >= 0) lut[ch] = (int8_t)threadIdx.x;
  }
  // the cached chunks, eight loads in flight per thread
  const int cm = (1 << chs) - 1;
  for (int e0 = 0; e0 < cpts; e0 += kCcBlock * 8) {
    V tmp[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = e0 + k * kCcBlock + (int)threadIdx.x;
      if (e < cpts) {
        const int ch = sl[e >> chs];
        const int j = (ch << chs) + (e & cm);
        tmp[k] = (ch >= 0 && j < n) ? Yv[j] : V{};
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = e0 + k * kCcBlock + (int)threadIdx.x;
      if (e < cpts) cy[e] = tmp[k];
    }
  }
  __syncthreads();
  auto gather = [&](int c, float (&yj)[D]) {
    const int sl_of = lut[c >> chs];
    const V v = sl_of >= 0 ? cy[(sl_of << chs) | (c & cm)] : Yv[c];
    if constexpr (D == 1) {
      yj[0] = v;
    } else if constexpr (D == 2) {
      yj[0] = v.x; yj[1] = v.y;
    } else {
      yj[0] = v.x; yj[1] = v.y; yj[2] = v.z;
    }
  };
  const int r0 = range * rows_per_block;
  const int r1 = min(n, r0 + rows_per_block);
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  constexpr int NG = kCcBlock / G;
  for (int row = r0 + grp; row < r1; row += NG) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    row_accum_g<D, G, U>(rowptr[row] + lane, rowptr[row + 1], col, val, gather, yi, acc);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, G);
    if (lane == 0) store_pt<D>(F, row, acc);
  }
}

// Chunk histogram of the nonzeros per row range (one wave per row).
__global__ __launch_bounds__(kBlock) void chunk_hist_kernel(int n, int rows_per_block, int chs, int nchunks,
                                                            const int32_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ col,
                                                            int32_t *__restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < n; row += nw) {
    int32_t *h = counts + (int64_t)(row / rows_per_block) * nchunks;
    for (int t = rowptr[row] + lane; t < rowptr[row + 1]; t += 64) atomicAdd(&h[col[t] >> chs], 1);
  }
}

// Per row range, the nslots chunks with the most nonzeros (ties: lower chunk
// first; chunks without nonzeros are not taken, their slots get -1).  One
// work-group per range; the range's counts are consumed in place.
// The same counts with one work-group per row range accumulating in LDS,
// then writing every chunk's count (zeros included).
__global__ __launch_bounds__(kBlock) void chunk_hist_lds_kernel(int n, int rows_per_block, int chs, int nchunks,
                                                                const int32_t *__restrict__ rowptr,
                                                                const int32_t *__restrict__ col,
                                                                int32_t *__restrict__ counts) {
  extern __shared__ int32_t hist[];
  for (int i = threadIdx.x; i < nchunks; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int row = r0 + w; row < r1; row += nw)
    for (int t = rowptr[row] + lane; t < rowptr[row + 1]; t += 64) atomicAdd(&hist[col[t] >> chs], 1);
  __syncthreads();
  int32_t *h = counts + (int64_t)blockIdx.x * nchunks;
  for (int i = threadIdx.x; i < nchunks; i += blockDim.x) h[i] = hist[i];
}

__global__ __launch_bounds__(kBlock) void chunk_select_kernel(int nranges, int nchunks, int nslots,
                                                              int32_t *__restrict__ counts,
                                                              int32_t *__restrict__ slots) {
  __shared__ int64_t best[kBlock / 64];
  const int range = blockIdx.x;
  if (range >= nranges) return;
  int32_t *h = counts + (int64_t)range * nchunks;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int s = 0; s < nslots; ++s) {
    // key: count in the high word, inverted chunk index in the low word
    int64_t k = -1;
    for (int i = threadIdx.x; i < nchunks; i += blockDim.x) {
      const int64_t v = h[i] > 0 ? ((int64_t)h[i] << 32) | (int64_t)(0x7FFFFFFF - i) : -1;
      k = v > k ? v : k;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const int64_t other = __shfl_xor(k, o, 64);
      k = other > k ? other : k;
    }
    if (lane == 0) best[wid] = k;
    __syncthreads();
    int64_t b = best[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = best[w] > b ? best[w] : b;
    const int ch = b < 0 ? -1 : 0x7FFFFFFF - (int)(b & 0xFFFFFFFF);
    if (threadIdx.x == 0) {
      slots[(int64_t)range * nslots + s] = ch;
      if (ch >= 0) h[ch] = 0;
    }
    __syncthreads();
  }
}

// Format 4, windowed form: as the blocked form, but the work-group first
// copies the window of Y centred on its rows into LDS.  In the locality order
// most neighbours of a row range lie in that window, so most gathers become
// LDS reads and only the rest reach L2.  Each lane sums the same entries in
// the same order as the other forms (only the source of y_j changes), so the
// result is the same.
template <int D, int G, int U>
__global__ __launch_bounds__(kBlock) void attr_csr_win_kernel(int n, int rows_per_block, int nranges, int win,
                                                              const int32_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ col,
                                                              const float *__restrict__ val,
                                                              const float *__restrict__ Y,
                                                              float *__restrict__ F) {
  using V = typename VecT<D>::T;
  extern __shared__ float4 wraw[];
  V *wy = reinterpret_cast<V *>(wraw);
  const V *__restrict__ Yv = reinterpret_cast<const V *>(Y);
  const int per_xcd = (gridDim.x + kXcds - 1) / kXcds;
  const int range = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (range >= nranges) return;  // the whole work-group leaves together
  const int r0 = range * rows_per_block;
  const int r1 = min(n, r0 + rows_per_block);
  const int wn = min(win, n);
  const int c0 = max(0, min(((r0 + r1) >> 1) - (wn >> 1), n - wn));
  {
    // eight loads in flight per thread before the LDS writes
    constexpr int SU = 8;
    for (int i0 = 0; i0 < wn; i0 += kBlock * SU) {
      V tmp[SU];
#pragma unroll
      for (int k = 0; k < SU; ++k) {
        const int i = i0 + k * kBlock + (int)threadIdx.x;
        if (i < wn) tmp[k] = Yv[c0 + i];
      }
#pragma unroll
      for (int k = 0; k < SU; ++k) {
        const int i = i0 + k * kBlock + (int)threadIdx.x;
        if (i < wn) wy[i] = tmp[k];
      }
    }
  }
  __syncthreads();
  // one generic-address load per neighbour: the LDS copy when it is in the
  // window, Y otherwise; no branch, so the loads of an unrolled step overlap
  const V *wg = wy;
  auto src = [&](int c) -> const V * {
    const unsigned u = (unsigned)(c - c0);
    return u < (unsigned)wn ? wg + u : Yv + c;
  };
  auto unpack = [&](const V &v, float (&yj)[D]) {
    if constexpr (D == 1) {
      yj[0] = v;
    } else if constexpr (D == 2) {
      yj[0] = v.x; yj[1] = v.y;
    } else {
      yj[0] = v.x; yj[1] = v.y; yj[2] = v.z;
    }
  };
  auto gather = [&](int c, float (&yj)[D]) { unpack(*src(c), yj); };
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  constexpr int NG = kBlock / G;
  for (int row = r0 + grp; row < r1; row += NG) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    const int e = rowptr[row + 1];
    int t = rowptr[row] + lane;
    for (; t + (U - 1) * G < e; t += U * G) {
      int c[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(col + t + u * G);
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(val + t + u * G);
      V w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = *src(c[u]);
      float yj[U][D];
#pragma unroll
      for (int u = 0; u < U; ++u) unpack(w[u], yj[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) accum_pair<D>(yi, yj[u], v[u], acc);
    }
    for (; t < e; t += G) {
      float yj[D];
      gather(__builtin_nontemporal_load(col + t), yj);
      accum_pair<D>(yi, yj, __builtin_nontemporal_load(val + t), acc);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, G);
    if (lane == 0) store_pt<D>(F, row, acc);
  }
}

// Format 4, column-localised form.  Rows are cut into blocks whose distinct
// columns number at most `cap`; each block keeps its sorted distinct columns
// (ucol) and every nonzero a 16-bit index into them (lidx), both built once at
// upload.  A work-group first gathers Y at its block's distinct columns into
// LDS (each neighbour of the block read once, with coalesced loads where the
// columns are dense), then every nonzero reads its y_j from LDS.  Each lane
// sums the same entries in the same order as attr_csr_seg_kernel<D,G>, so the
// result is bit-identical to it.
template <int D, int G>
__global__ __launch_bounds__(kBlock) void attr_csr_loc_kernel(int nblk, const int32_t *__restrict__ brow,
                                                              const int32_t *__restrict__ bu,
                                                              const int32_t *__restrict__ ucol,
                                                              const int32_t *__restrict__ rowptr,
                                                              const uint16_t *__restrict__ lidx,
                                                              const float *__restrict__ val,
                                                              const float *__restrict__ Y,
                                                              float *__restrict__ F) {
  using V = typename VecT<D>::T;
  extern __shared__ float4 lraw[];
  V *ly = reinterpret_cast<V *>(lraw);
  const V *__restrict__ Yv = reinterpret_cast<const V *>(Y);
  const int per_xcd = (gridDim.x + kXcds - 1) / kXcds;
  const int b = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (b >= nblk) return;  // the whole work-group leaves together
  const int r0 = brow[b], r1 = brow[b + 1];
  const int u0 = bu[b], nu = bu[b + 1] - u0;
  {
    constexpr int SU = 8;
    for (int i0 = 0; i0 < nu; i0 += kBlock * SU) {
      int c[SU];
#pragma unroll
      for (int k = 0; k < SU; ++k) {
        const int i = i0 + k * kBlock + (int)threadIdx.x;
        c[k] = i < nu ? __builtin_nontemporal_load(ucol + u0 + i) : -1;
      }
      V tmp[SU];
#pragma unroll
      for (int k = 0; k < SU; ++k)
        if (c[k] >= 0) tmp[k] = Yv[c[k]];
#pragma unroll
      for (int k = 0; k < SU; ++k)
        if (c[k] >= 0) ly[i0 + k * kBlock + threadIdx.x] = tmp[k];
    }
  }
  __syncthreads();
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  constexpr int NG = kBlock / G;
  for (int row = r0 + grp; row < r1; row += NG) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    const int e = rowptr[row + 1];
    for (int t = rowptr[row] + lane; t < e; t += G) {
      const V v = ly[__builtin_nontemporal_load(lidx + t)];
      float yj[D];
      if constexpr (D == 1) {
        yj[0] = v;
      } else if constexpr (D == 2) {
        yj[0] = v.x; yj[1] = v.y;
      } else {
        yj[0] = v.x; yj[1] = v.y; yj[2] = v.z;
      }
      accum_pair<D>(yi, yj, __builtin_nontemporal_load(val + t), acc);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, G);
    if (lane == 0) store_pt<D>(F, row, acc);
  }
}

// Format 0: one thread per row (CU/pq.cuh:16-75).
template <int D>
__global__ __launch_bounds__(kBlock) void attr_csr_row_kernel(int n, const int32_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ col,
                                                              const float *__restrict__ val,
                                                              const float *__restrict__ Y,
                                                              float *__restrict__ F) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    for (int t = rowptr[row]; t < rowptr[row + 1]; ++t) {
      float yj[D];
      load_pt<D>(Y, col[t], yj);
      accum_pair<D>(yi, yj, val[t], acc);
    }
    store_pt<D>(F, row, acc);
  }
}

// Format 2: one thread per nonzero with atomic accumulation (CU/pq.cuh:77-118).
// Also the COO tail of format 3.  F must be zeroed (or hold the ELL part).
template <int D>
__global__ __launch_bounds__(kBlock) void attr_coo_kernel(int64_t nnz, const int32_t *__restrict__ row,
                                                          const int32_t *__restrict__ col,
                                                          const float *__restrict__ val,
                                                          const float *__restrict__ Y,
                                                          float *__restrict__ F) {
  const int ds = point_stride(D);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int r = row[t];
    float yi[D], yj[D], acc[D];
    load_pt<D>(Y, r, yi);
    load_pt<D>(Y, col[t], yj);
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    accum_pair<D>(yi, yj, val[t], acc);
#pragma unroll
    for (int a = 0; a < D; ++a) unsafeAtomicAdd(&F[(int64_t)r * ds + a], acc[a]);
  }
}

// Format 3, ELL part: one thread per row over a column-major [width][n]
// slab (CU/hybridpq.cu:4-69).  Padding entries point at the row itself.
template <int D>
__global__ __launch_bounds__(kBlock) void attr_ell_kernel(int n, int width, const int32_t *__restrict__ ecol,
                                                          const float *__restrict__ eval,
                                                          const float *__restrict__ Y,
                                                          float *__restrict__ F) {
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    float acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = 0.0f;
    for (int e = 0; e < width; ++e) {
      const int64_t o = (int64_t)e * n + row;
      float yj[D];
      load_pt<D>(Y, ecol[o], yj);
      accum_pair<D>(yi, yj, eval[o], acc);
    }
    store_pt<D>(F, row, acc);
  }
}

// KL = sum_nnz p log((p + FLT_MIN) / (q + FLT_MIN)), p = alpha P_ij,
// q = 1 / (Z (1 + d_ij^2)); FP64 accumulation into per-block partials.
template <int D>
__global__ __launch_bounds__(kBlock) void kl_kernel(int n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col,
                                                    const float *__restrict__ val,
                                                    const float *__restrict__ Y, double alpha,
                                                    const DevScalars *sc, double *part) {
  __shared__ double lds[4];
  const double Z = sc->Z;
  double c = 0.0;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
    float yi[D];
    load_pt<D>(Y, row, yi);
    for (int t = rowptr[row]; t < rowptr[row + 1]; ++t) {
      float yj[D];
      load_pt<D>(Y, col[t], yj);
      double dist = 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const double df = (double)yi[a] - (double)yj[a];
        dist += df * df;
      }
      const double p = (double)val[t] * alpha;
      const double q = 1.0 / (Z * (1.0 + dist));
      c += p * log((p + (double)FLT_MIN) / (q + (double)FLT_MIN));
    }
  }
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8] = lds[0] + lds[1] + lds[2] + lds[3];
}



}  // namespace

int launch_attr_csr_seg(hipStream_t s, int d, int group, int n, const int32_t *rowptr,
                        const int32_t *col, const float *val, const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  int blocks = div_up((int64_t)n * group, kBlock);
#define SEG_CASE(DD, GG)                                                                      \
  case GG:                                                                                    \
    attr_csr_seg_kernel<DD, GG><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, F);         \
    break;
#define SEG_D(DD)                                                                             \
  switch (group) {                                                                            \
    SEG_CASE(DD, 4) SEG_CASE(DD, 8) SEG_CASE(DD, 16) SEG_CASE(DD, 32) SEG_CASE(DD, 64)        \
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "row group must be 4..64");                     \
  }
  switch (d) {
    case 1: SEG_D(1) break;
    case 2: SEG_D(2) break;
    case 3: SEG_D(3) break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
#undef SEG_D
#undef SEG_CASE
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_csr_seg2(hipStream_t s, int d, int group, int unroll, int n, const int32_t *rowptr,
                         const int32_t *col, const float *val, const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  const int blocks = div_up((int64_t)n * group, kBlock);
  bool ok = false;
  auto go = [&](auto dd, auto gg, auto uu) {
    constexpr int DD = decltype(dd)::value, GG = decltype(gg)::value, UU = decltype(uu)::value;
    if (d == DD && group == GG && unroll == UU) {
      attr_csr_seg2_kernel<DD, GG, UU><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, F);
      ok = true;
    }
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  using I8 = std::integral_constant<int, 8>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  auto per_g = [&](auto dd, auto uu) {
    go(dd, I4{}, uu); go(dd, I8{}, uu); go(dd, I16{}, uu); go(dd, I32{}, uu); go(dd, I64{}, uu);
  };
  auto per_d = [&](auto dd) {
    per_g(dd, I1{}); per_g(dd, I2{}); per_g(dd, I4{});
    go(dd, I16{}, I8{}); go(dd, I32{}, I8{}); go(dd, I64{}, I8{}); go(dd, I16{}, I16{}); go(dd, I32{}, I16{});
  };
  per_d(I1{});
  per_d(I2{});
  per_d(I3{});
  SGT_CHECK(ok, SGTSNE_ERR_ARG, "no attractive kernel for d=%d group=%d unroll=%d", d, group, unroll);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_csr_blk(hipStream_t s, int d, int group, int unroll, int rows_per_block, int n,
                        const int32_t *rowptr, const int32_t *col, const float *val, const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  const int nranges = div_up(n, rows_per_block);
  const int blocks = div_up(nranges, kXcds) * kXcds;
  bool ok = false;
  auto go = [&](auto dd, auto gg, auto uu) {
    constexpr int DD = decltype(dd)::value, GG = decltype(gg)::value, UU = decltype(uu)::value;
    if (d == DD && group == GG && unroll == UU) {
      attr_csr_blk_kernel<DD, GG, UU><<<blocks, kBlock, 0, s>>>(n, rows_per_block, nranges, rowptr, col, val, Y, F);
      ok = true;
    }
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  using I8 = std::integral_constant<int, 8>;
  auto per_g = [&](auto dd, auto uu) { go(dd, I16{}, uu); go(dd, I32{}, uu); go(dd, I64{}, uu); };
  auto per_d = [&](auto dd) {
    per_g(dd, I1{});
    per_g(dd, I2{});
    go(dd, I16{}, I8{}); go(dd, I32{}, I8{}); go(dd, I16{}, I16{}); go(dd, I32{}, I16{});
  };
  per_d(I1{});
  per_d(I2{});
  per_d(I3{});
  SGT_CHECK(ok, SGTSNE_ERR_ARG, "no blocked attractive kernel for d=%d group=%d unroll=%d", d, group, unroll);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_csr_cc(hipStream_t s, int d, int group, int unroll, int rows_per_block, int chs, int nslots,
                       const int32_t *slots, int n, const int32_t *rowptr, const int32_t *col, const float *val,
                       const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  const int ds = d == 3 ? 4 : d;
  const int nchunks = div_up(n, 1 << chs);
  const size_t lds = ((size_t)nslots << chs) * ds * sizeof(float) + (size_t)nchunks;
  SGT_CHECK(slots != nullptr && rows_per_block > 0 && chs >= 6 && chs <= 16 && nslots >= 1 && nslots <= 127 &&
                lds <= kCcLdsMax,
            SGTSNE_ERR_ARG, "chunk cache of %d x 2^%d points does not fit LDS", nslots, chs);
  const int nranges = div_up(n, rows_per_block);
  const int blocks = div_up(nranges, kXcds) * kXcds;
  bool ok = false;
  auto go = [&](auto dd, auto gg, auto uu) {
    constexpr int DD = decltype(dd)::value, GG = decltype(gg)::value, UU = decltype(uu)::value;
    if (d == DD && group == GG && unroll == UU) {
      attr_csr_cc_kernel<DD, GG, UU><<<blocks, kCcBlock, lds, s>>>(n, rows_per_block, nranges, chs, nslots, nchunks,
                                                                  slots, rowptr, col, val, Y, F);
      ok = true;
    }
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I8 = std::integral_constant<int, 8>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  auto per_d = [&](auto dd) {
    go(dd, I16{}, I16{}); go(dd, I32{}, I8{}); go(dd, I32{}, I16{}); go(dd, I64{}, I8{});
  };
  per_d(I1{});
  per_d(I2{});
  per_d(I3{});
  SGT_CHECK(ok, SGTSNE_ERR_ARG, "no chunk-cached attractive kernel for d=%d group=%d unroll=%d", d, group, unroll);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_chunk_cache(hipStream_t s, int n, int rows_per_block, int chs, int nslots, const int32_t *rowptr,
                       const int32_t *col, int32_t *counts, int32_t *slots) {
  if (n <= 0) return SGTSNE_OK;
  const int nranges = div_up(n, rows_per_block);
  const int nchunks = div_up(n, 1 << chs);
  if ((size_t)nchunks * sizeof(int32_t) <= 65536) {
    // one work-group per row range, its histogram in LDS
    chunk_hist_lds_kernel<<<nranges, kBlock, (size_t)nchunks * sizeof(int32_t), s>>>(n, rows_per_block, chs, nchunks,
                                                                                    rowptr, col, counts);
  } else {
    SGT_HIP(hipMemsetAsync(counts, 0, sizeof(int32_t) * (size_t)nranges * nchunks, s));
    chunk_hist_kernel<<<(int)std::min<int64_t>(div_up((int64_t)n * 64, kBlock), 65536), kBlock, 0, s>>>(
        n, rows_per_block, chs, nchunks, rowptr, col, counts);
  }
  SGT_LAUNCH_CHECK();
  chunk_select_kernel<<<nranges, kBlock, 0, s>>>(nranges, nchunks, nslots, counts, slots);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_csr_win(hipStream_t s, int d, int group, int unroll, int rows_per_block, int win, int n,
                        const int32_t *rowptr, const int32_t *col, const float *val, const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  const int ds = d == 3 ? 4 : d;
  const int wn = std::min(win, n);
  const size_t lds = (size_t)wn * ds * sizeof(float);
  SGT_CHECK(win >= 64 && lds <= 65536, SGTSNE_ERR_ARG, "window of %d points does not fit LDS", win);
  const int nranges = div_up(n, rows_per_block);
  const int blocks = div_up(nranges, kXcds) * kXcds;
  bool ok = false;
  auto go = [&](auto dd, auto gg, auto uu) {
    constexpr int DD = decltype(dd)::value, GG = decltype(gg)::value, UU = decltype(uu)::value;
    if (d == DD && group == GG && unroll == UU) {
      attr_csr_win_kernel<DD, GG, UU><<<blocks, kBlock, lds, s>>>(n, rows_per_block, nranges, win, rowptr, col, val,
                                                                  Y, F);
      ok = true;
    }
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  auto per_g = [&](auto dd, auto uu) { go(dd, I16{}, uu); go(dd, I32{}, uu); };
  auto per_d = [&](auto dd) { per_g(dd, I1{}); per_g(dd, I2{}); };
  per_d(I1{});
  per_d(I2{});
  per_d(I3{});
  SGT_CHECK(ok, SGTSNE_ERR_ARG, "no windowed attractive kernel for d=%d group=%d unroll=%d", d, group, unroll);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_csr_loc(hipStream_t s, int d, int group, int nblk, int cap, const int32_t *brow, const int32_t *bu,
                        const int32_t *ucol, const int32_t *rowptr, const uint16_t *lidx, const float *val,
                        const float *Y, float *F) {
  if (nblk <= 0) return SGTSNE_OK;
  const int ds = d == 3 ? 4 : d;
  const size_t lds = (size_t)cap * ds * sizeof(float);
  SGT_CHECK(cap >= 1 && cap <= 65536 && lds <= 65536, SGTSNE_ERR_ARG, "column block of %d points does not fit LDS",
            cap);
  const int blocks = div_up(nblk, kXcds) * kXcds;
  bool ok = false;
  auto go = [&](auto dd, auto gg) {
    constexpr int DD = decltype(dd)::value, GG = decltype(gg)::value;
    if (d == DD && group == GG) {
      attr_csr_loc_kernel<DD, GG><<<blocks, kBlock, lds, s>>>(nblk, brow, bu, ucol, rowptr, lidx, val, Y, F);
      ok = true;
    }
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  auto per_d = [&](auto dd) { go(dd, I16{}); go(dd, I32{}); go(dd, I64{}); };
  per_d(I1{});
  per_d(I2{});
  per_d(I3{});
  SGT_CHECK(ok, SGTSNE_ERR_ARG, "no column-localised attractive kernel for d=%d group=%d", d, group);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_csr_row(hipStream_t s, int d, int n, const int32_t *rowptr, const int32_t *col,
                        const float *val, const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  int blocks = div_up(n, kBlock);
  switch (d) {
    case 1: attr_csr_row_kernel<1><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, F); break;
    case 2: attr_csr_row_kernel<2><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, F); break;
    case 3: attr_csr_row_kernel<3><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, F); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_coo(hipStream_t s, int d, int64_t nnz, const int32_t *row, const int32_t *col,
                    const float *val, const float *Y, float *F) {
  if (nnz <= 0) return SGTSNE_OK;
  int blocks = (int)std::min<int64_t>((nnz + kBlock - 1) / kBlock, 1 << 20);
  switch (d) {
    case 1: attr_coo_kernel<1><<<blocks, kBlock, 0, s>>>(nnz, row, col, val, Y, F); break;
    case 2: attr_coo_kernel<2><<<blocks, kBlock, 0, s>>>(nnz, row, col, val, Y, F); break;
    case 3: attr_coo_kernel<3><<<blocks, kBlock, 0, s>>>(nnz, row, col, val, Y, F); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_attr_ell(hipStream_t s, int d, int n, int width, const int32_t *ecol, const float *eval,
                    const float *Y, float *F) {
  if (n <= 0) return SGTSNE_OK;
  int blocks = div_up(n, kBlock);
  switch (d) {
    case 1: attr_ell_kernel<1><<<blocks, kBlock, 0, s>>>(n, width, ecol, eval, Y, F); break;
    case 2: attr_ell_kernel<2><<<blocks, kBlock, 0, s>>>(n, width, ecol, eval, Y, F); break;
    case 3: attr_ell_kernel<3><<<blocks, kBlock, 0, s>>>(n, width, ecol, eval, Y, F); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_kl(hipStream_t s, int d, int n, const int32_t *rowptr, const int32_t *col,
              const float *val, const float *Y, double alpha, const DevScalars *sc, RedWork rw,
              DevScalars *out) {
  int blocks = std::min(div_up(n > 0 ? n : 1, kBlock), std::min(rw.max_blocks, 2048));
  switch (d) {
    case 1: kl_kernel<1><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, alpha, sc, rw.part); break;
    case 2: kl_kernel<2><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, alpha, sc, rw.part); break;
    case 3: kl_kernel<3><<<blocks, kBlock, 0, s>>>(n, rowptr, col, val, Y, alpha, sc, rw.part); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return launch_finalize_kl(s, blocks, rw, out);
}

}  // namespace sgt
