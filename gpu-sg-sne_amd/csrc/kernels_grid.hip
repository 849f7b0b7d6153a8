// kernels_grid.hip -- point <-> grid kernels of the repulsive term, the fused
// update and the device-side reductions of one gradient iteration.
//
// Reference semantics (HY = code/cuda/SG-tSNE-HYB/sgtsnepi/src,
// CU = code/cuda/SG-tSNE-CUDA/src):
//   grid mapping      HY/nuconv.cpp:22-47, CU/nuconv.cu:9-35 (FP32 clamp)
//   Lagrange weights  HY/gridding.cpp:28-40
//   spread (S2G)      HY/gridding.cpp:164-211 / CU/gridding.cu:145-185
//   gather (G2S)      HY/gridding.cpp:508-560 / CU/gridding.cu:187-230
//   forces and Z      HY/qq.cpp:75-106 / CU/Frep.cu:43-80
//   update            CU/gradient_descend.cu:12-44, HY/gradient_descendHYB.cpp:29-78
//
// Reductions: every block writes its partials to rw.part[block * 8 + k]; a
// one-block finalize kernel combines them in a fixed order (deterministic),
// so no kernel needs a device-wide fence.
#include "common.hpp"
#include "kernels.hpp"
#include "grid_device.hpp"
#include "spread_fix.hpp"
#include "../../include/sgtsne.h"

namespace sgt {

namespace {

constexpr int kBlock = 256;

enum FinMode { FIN_MINMAX = 0, FIN_ZNEU = 1, FIN_MEAN = 2, FIN_ZEXACT = 3, FIN_KL = 4 };

// ---- plain streaming copy (bandwidth / counter calibration) ----------------------
__global__ __launch_bounds__(kBlock) void stream_copy_kernel(const float4 *__restrict__ in, float4 *__restrict__ out,
                                                             int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) out[i] = in[i];
}

// ---- fill ------------------------------------------------------------------
__global__ void fill_kernel(float *__restrict__ p, int64_t count, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// ---- centring + min/max ------------------------------------------------------
// Y -= mean (the mean of the last update, HY/gradient_descendHYB.cpp:64-76),
// then per-block min/max of the centred Y (HY/qq.cpp:147-163) and a count of
// non-finite coordinates.
template <int D>
__global__ __launch_bounds__(kBlock) void center_minmax_kernel(float *__restrict__ Y, int n,
                                                               int apply_mean, double *part,
                                                               const DevScalars *sc) {
  __shared__ double lds[(2 * D + 1) * 4];
  float m[D];
#pragma unroll
  for (int a = 0; a < D; ++a) m[a] = apply_mean ? (float)sc->mean[a] : 0.0f;
  float mn[D], mx[D];
  float bad = 0.0f;
#pragma unroll
  for (int a = 0; a < D; ++a) { mn[a] = INFINITY; mx[a] = -INFINITY; }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float y[D];
    load_pt<D>(Y, i, y);
    if (apply_mean) {
#pragma unroll
      for (int a = 0; a < D; ++a) y[a] -= m[a];
      store_pt<D>(Y, i, y);
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
      mn[a] = fminf(mn[a], y[a]);
      mx[a] = fmaxf(mx[a], y[a]);
      bad += isfinite(y[a]) ? 0.0f : 1.0f;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double vmin[D], vmax[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    vmin[a] = wave_min((double)mn[a]);
    vmax[a] = wave_max((double)mx[a]);
  }
  const double vbad = wave_sum((double)bad);
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      lds[a * 4 + wid] = vmin[a];
      lds[(D + a) * 4 + wid] = vmax[a];
    }
    lds[2 * D * 4 + wid] = vbad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *pb = part + (size_t)blockIdx.x * 8;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      pb[a] = fmin(fmin(lds[a * 4], lds[a * 4 + 1]), fmin(lds[a * 4 + 2], lds[a * 4 + 3]));
      pb[4 + a] = fmax(fmax(lds[(D + a) * 4], lds[(D + a) * 4 + 1]),
                       fmax(lds[(D + a) * 4 + 2], lds[(D + a) * 4 + 3]));
    }
    pb[3] = lds[2 * D * 4] + lds[2 * D * 4 + 1] + lds[2 * D * 4 + 2] + lds[2 * D * 4 + 3];
  }
}

// ---- S2G, LDS-privatised whole grid -------------------------------------------
// For grids that fit in LDS: each block spreads a contiguous chunk of points
// into a private LDS copy of the grid and writes it out; a second kernel sums
// the private grids.  No global atomics.  Lanes map to stencil nodes, not to
// points: the 4^d lanes of one point touch 4^d distinct cells, so an LDS
// atomic instruction conflicts only between the 64/4^d points of a wave
// (at most 4-way in 2-D, never in 3-D) even when every point of a clustered
// embedding falls into the same few cells.
constexpr int kS2GBlock = 1024;  // 16 waves share one private grid
constexpr int kS2GUnroll = 4;    // points per lane whose loads are issued together

template <int D>
__global__ __launch_bounds__(kS2GBlock) void s2g_lds_kernel(GridParams g, int n, const float *__restrict__ Y,
                                                            float *__restrict__ part, int chunk, int copies) {
  extern __shared__ float sgrid[];
  constexpr int M = D + 1;
  constexpr int PPW = NodeLanes<D>::PPW;
  constexpr int NODES = NodeLanes<D>::NODES;
  constexpr int U = kS2GUnroll;
  const int G = grid_cells<D>(g.ng);
  const int MG = M * G;
  for (int e = threadIdx.x; e < copies * MG; e += blockDim.x) sgrid[e] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int node = lane % NODES, slot = lane / NODES;
  const int i0 = blockIdx.x * chunk;
  const int i1 = min(n, i0 + chunk);
  // each point slot of each wave spreads a contiguous run of points, and a
  // lane adds its successive points in registers while they hit the same grid
  // node: in the locality order consecutive points mostly share a cell, so
  // most of the LDS atomics of the plain per-point form become register adds
  const int per_wave = (chunk + nwaves - 1) / nwaves;
  const int per_slot = (per_wave + PPW - 1) / PPW;
  const int w0 = min(i1, i0 + wave * per_wave + slot * per_slot);
  const int w1 = min(min(i1, i0 + (wave + 1) * per_wave), w0 + per_slot);
  float *cp = sgrid + (slot % copies) * MG;  // the point slots of one instruction use different copies
  int cur = -1;
  float acc[M];
#pragma unroll
  for (int a = 0; a < M; ++a) acc[a] = 0.0f;
  for (int base = w0; base < w1; base += U) {
    float y[U][D];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u;
      if (i < w1) load_pt<D>(Y, i, y[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u;
      if (i < w1) {
        int idx;
        float ww;
        node_of_lane<D>(g, y[u], node, idx, ww);
        if (idx != cur) {
          if (cur >= 0) {
#pragma unroll
            for (int a = 0; a < M; ++a) atomicAdd(&cp[a * G + cur], acc[a]);
          }
          cur = idx;
#pragma unroll
          for (int a = 0; a < M; ++a) acc[a] = 0.0f;
        }
        acc[0] += ww;
#pragma unroll
        for (int a = 0; a < D; ++a) acc[a + 1] += (y[u][a] - g.center[a]) * ww;
      }
    }
  }
  if (cur >= 0) {
#pragma unroll
    for (int a = 0; a < M; ++a) atomicAdd(&cp[a * G + cur], acc[a]);
  }
  __syncthreads();
  // sum of the copies -> this block's slice of the partial buffer (coalesced)
  float *dst = part + (int64_t)blockIdx.x * MG;
  for (int e = threadIdx.x; e < MG; e += blockDim.x) {
    float v = 0.0f;
    for (int k = 0; k < copies; ++k) v += sgrid[k * MG + e];
    dst[e] = v;
  }
}

// ---- S2G on the matrix cores (2-D grids of side <= 16 NT) -------------------------
// For 4 points k at a time, the spread of channel c onto a 16 x 16 tile of
// nodes is a rank-4 product, one v_mfma_f32_16x16x4f32:
//   V_c[gx][gy] += sum_k wx_k(gx) * (q_c,k * wy_k(gy))
// with A[gx][k] = wx_k(gx) (lane l holds gx = l % 16 of point k = l / 16),
// B[k][gy] = q_c,k wy_k(gy) (lane l holds gy = l % 16 of the same point), and
// D[gx][gy] in lane l, register v at gx = 4 (l / 16) + v, gy = l % 16.  A
// lane evaluates one Lagrange factor per axis and tile; nodes off a point's
// stencil get weight 0.  Nothing is contended: each wave keeps its points'
// sums in registers whatever the order of the points or how many share a
// cell, the block's 4 waves are added in LDS in a fixed order, and
// grid_reduce_kernel adds the blocks' grids in a fixed order (deterministic).
typedef float mfv4 __attribute__((ext_vector_type(4)));
constexpr int kMfmaBlock = 256;
// points per work-group: about 1024 work-groups whatever N, whole batches of
// 256 (one per wave), between 256 and 4096 points each
inline int mfma_chunk(int n) { return std::min(4096, std::max(256, div_up(div_up(n, 1024), 256) * 256)); }

// Weight of stencil node f + r of a point at f + 1 + dl (the node factor of
// lane_node), 0 when r is not 0..3.
__device__ __forceinline__ float lagrange_node(int r, float dl) {
  const float x = r == 0 ? 1.0f + dl : r == 1 ? dl : r == 2 ? 1.0f - dl : 2.0f - dl;
  const float w = (r == 1 || r == 2) ? lg1(x) : lg2(x);
  return (r >= 0 && r <= 3) ? w : 0.0f;
}

template <int NT>
__global__ __launch_bounds__(kMfmaBlock) void s2g_mfma_kernel(GridParams g, int n, const float *__restrict__ Y,
                                                              float *__restrict__ part, int chunk) {
  constexpr int T = 16 * NT;  // side of the tiled grid (>= ng)
  __shared__ float red[4][3][T * T];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = lane >> 4, node = lane & 15;
  const int i0 = blockIdx.x * chunk;
  const int i1 = min(n, i0 + chunk);
  mfv4 acc[3][NT * NT];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < NT * NT; ++t) acc[c][t] = mfv4{0.0f, 0.0f, 0.0f, 0.0f};
  const float2 *Y2 = reinterpret_cast<const float2 *>(Y);
  // each wave takes 64 consecutive points at a time with one coalesced load
  // (the next 64 are loaded while these are spread); step s of a batch hands
  // point 4 s + k to lanes 16 k .. 16 k + 15 by cross-lane reads
  int base = i0 + 64 * wave;
  float2 nxt = base + lane < i1 ? Y2[base + lane] : make_float2(g.minv[0], g.minv[1]);
  for (; base < i1; base += 256) {
    const float2 cur = nxt;
    const int nb = base + 256;
    nxt = nb + lane < i1 ? Y2[nb + lane] : make_float2(g.minv[0], g.minv[1]);
    const int cnt = min(64, i1 - base);
#pragma unroll 4
    for (int st = 0; st < 16; ++st) {
      if (4 * st >= cnt) break;  // wave-uniform
      const int src = 4 * st + k;
      const float yx = __shfl(cur.x, src), yy = __shfl(cur.y, src);
      const bool in = src < cnt;
      float t0 = (yx - g.minv[0]) * g.scale, t1 = (yy - g.minv[1]) * g.scale;
      t0 = fminf(fmaxf(t0, 0.0f), g.gmax);  // NaN -> 0: never leaves the grid
      t1 = fminf(fmaxf(t1, 0.0f), g.gmax);
      const int f0 = (int)t0, f1 = (int)t1;
      const float d0 = t0 - (float)f0, d1 = t1 - (float)f1;
      float wx[NT], wy[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        wx[t] = in ? lagrange_node(16 * t + node - f0, d0) : 0.0f;
        wy[t] = lagrange_node(16 * t + node - f1, d1);
      }
      const float q1 = yx - g.center[0], q2 = yy - g.center[1];
#pragma unroll
      for (int ty = 0; ty < NT; ++ty) {
        const float b0 = wy[ty], b1 = q1 * wy[ty], b2 = q2 * wy[ty];
#pragma unroll
        for (int tx = 0; tx < NT; ++tx) {
          acc[0][tx * NT + ty] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx[tx], b0, acc[0][tx * NT + ty], 0, 0, 0);
          acc[1][tx * NT + ty] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx[tx], b1, acc[1][tx * NT + ty], 0, 0, 0);
          acc[2][tx * NT + ty] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx[tx], b2, acc[2][tx * NT + ty], 0, 0, 0);
        }
      }
    }
  }
  // this wave's sums: node (gx, gy) of channel c at red[wave][c][gy * T + gx]
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int tx = 0; tx < NT; ++tx)
#pragma unroll
      for (int ty = 0; ty < NT; ++ty) {
        float *dst = &red[wave][c][(16 * ty + node) * T + 16 * tx + 4 * k];
        const mfv4 a = acc[c][tx * NT + ty];
        dst[0] = a[0];
        dst[1] = a[1];
        dst[2] = a[2];
        dst[3] = a[3];
      }
  __syncthreads();
  // the block's grid (channel-major, x fastest), waves added in a fixed order
  const int ng = g.ng, G = ng * ng;
  float *out = part + (int64_t)blockIdx.x * 3 * G;
  for (int e = threadIdx.x; e < 3 * G; e += blockDim.x) {
    const int c = e / G, cell = e - c * G;
    const int gy = cell / ng, gx = cell - gy * ng;
    const int sidx = gy * T + gx;
    out[e] = (red[0][c][sidx] + red[1][c][sidx]) + (red[2][c][sidx] + red[3][c][sidx]);
  }
}

// ---- 3-D S2G on the matrix cores (grids of side <= 16) ------------------------------
// A work-group spreads a chunk of points onto its own copy of the grid in LDS.
// It first sorts its points by the base z-cell of their stencils (a stable
// counting sort of (z-cell, position) keys: deterministic) and gives each of its 4
// waves a contiguous quarter of the sorted points, so the work is even however
// the points crowd a few cells.  A wave walks its points in runs of equal base
// z-cell f, 4 at a time; such a step touches slices f .. f + 3 only, and slice
// f + r of channel c gets the rank-4 product
//   V_c[gx][gy][f + r] += sum_k wx_k(gx) (q_c,k wy_k(gy) wz_k(r))
// (one v_mfma_f32_16x16x4f32, layouts as in the 2-D form) into a window of 4
// slice accumulators.  When the base cell moves up, the slices below it are
// final for this wave: they are stored into the LDS grid, which no other wave
// stores to (the waves' runs are disjoint), and the window slides.  The 4
// windows left at the end are added in wave order after a barrier, so the grid
// does not depend on timing; grid_reduce_kernel adds the work-groups' grids in
// a fixed order.
constexpr int kMfma3Chunk = 2048;  // points per work-group (a multiple of the work-group size)

__global__ __launch_bounds__(kMfmaBlock) void s2g_mfma3_kernel(GridParams g, int n, const float *__restrict__ Y,
                                                               float *__restrict__ part, int chunk) {
  __shared__ float grid[4 * 16 * 16 * 16];  // [channel][z][y][x], side ng
  __shared__ uint32_t key[kMfma3Chunk];
  __shared__ int zcnt[kMfma3Chunk / kMfmaBlock][4][16];
  __shared__ int zbase[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = lane >> 4, node = lane & 15;
  const int ng = g.ng, G = ng * ng * ng;
  const int i0 = blockIdx.x * chunk;
  const int cnt = min(n, i0 + chunk) - i0;
  const float4 *__restrict__ Y4 = reinterpret_cast<const float4 *>(Y) + i0;
  for (int e = tid; e < 4 * G; e += kMfmaBlock) grid[e] = 0.0f;
  // keys (base z-cell, position) of the chunk's points in ascending order, by
  // a stable counting sort over the at most 16 base z-cells: point i = 256 r
  // + 64 w + lane is ranked among the equal cells of its wave by a ballot,
  // and the (cell, round, wave) counts give its place
  constexpr int kRounds = kMfma3Chunk / kMfmaBlock;
  int fzr[kRounds], rk[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int i = r * kMfmaBlock + tid;
    fzr[r] = -1;
    if (i < cnt) fzr[r] = (int)fminf(fmaxf((Y4[i].z - g.minv[2]) * g.scale, 0.0f), g.gmax);
    rk[r] = 0;
    for (int v = 0; v < 16; ++v) {
      const unsigned long long m = __ballot(fzr[r] == v);
      if (fzr[r] == v) rk[r] = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) zcnt[r][wave][v] = __popcll(m);
    }
  }
  __syncthreads();
  if (tid < 16) {  // per cell: exclusive prefix over (round, wave), then the cell total
    int run = 0;
    for (int r = 0; r < kRounds; ++r)
      for (int w = 0; w < 4; ++w) {
        const int c = zcnt[r][w][tid];
        zcnt[r][w][tid] = run;
        run += c;
      }
    zbase[tid] = run;
  }
  __syncthreads();
  if (tid == 0) {  // cell totals -> first position of each cell
    int run = 0;
    for (int v = 0; v < 16; ++v) {
      const int c = zbase[v];
      zbase[v] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int i = r * kMfmaBlock + tid;
    const int v = fzr[r];
    if (v >= 0) key[zbase[v] + zcnt[r][wave][v] + rk[r]] = ((uint32_t)v << 16) | (uint32_t)i;
  }
  __syncthreads();
  const int pa = (cnt * wave) >> 2, pb = (cnt * (wave + 1)) >> 2;  // this wave's sorted positions
  mfv4 acc[4][4];  // [slice zb + r][channel]
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = mfv4{0.0f, 0.0f, 0.0f, 0.0f};
  // slice r of the window -> grid slice z (lane l holds x = 4 (l / 16) + v, y = l % 16)
  auto put = [&](int r, int z, bool add) {
    const int gx0 = 4 * k, gy = node;
    if (z >= ng || gy >= ng) return;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (gx0 + v < ng) {
          float *cell = &grid[((c * ng + z) * ng + gy) * ng + gx0 + v];
          *cell = add ? *cell + acc[r][c][v] : acc[r][c][v];
        }
  };
  auto clear = [&](int r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = mfv4{0.0f, 0.0f, 0.0f, 0.0f};
  };
  int zb = pa < pb ? (int)(key[pa] >> 16) : 0;  // window: slices zb .. zb + 3
  int p0 = pa;
  while (p0 < pb) {  // wave-uniform
    const int fz = (int)(key[p0] >> 16);
    if (fz != zb) {  // slices zb .. fz - 1 are final for this wave
      const int sh = fz - zb;
      if (sh == 1) {
        put(0, zb, false);
#pragma unroll
        for (int c = 0; c < 4; ++c) { acc[0][c] = acc[1][c]; acc[1][c] = acc[2][c]; acc[2][c] = acc[3][c]; }
        clear(3);
      } else if (sh == 2) {
        put(0, zb, false);
        put(1, zb + 1, false);
#pragma unroll
        for (int c = 0; c < 4; ++c) { acc[0][c] = acc[2][c]; acc[1][c] = acc[3][c]; }
        clear(2);
        clear(3);
      } else if (sh == 3) {
        put(0, zb, false);
        put(1, zb + 1, false);
        put(2, zb + 2, false);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[0][c] = acc[3][c];
        clear(1);
        clear(2);
        clear(3);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          put(r, zb + r, false);
          clear(r);
        }
      }
      zb = fz;
    }
    // end of the run of base cell fz within this wave's positions
    int lo = p0 + 1, hi = pb;
    const uint32_t nextz = (uint32_t)(fz + 1) << 16;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (key[m] < nextz) lo = m + 1;
      else hi = m;
    }
    const int re = lo;
    for (int ps = p0; ps < re; ps += 4) {
      const int src = ps + k;
      const bool in = src < re;
      const float4 y = Y4[key[in ? src : ps] & 0xFFFFu];
      float t0 = (y.x - g.minv[0]) * g.scale, t1 = (y.y - g.minv[1]) * g.scale, t2 = (y.z - g.minv[2]) * g.scale;
      t0 = fminf(fmaxf(t0, 0.0f), g.gmax);  // NaN -> 0: never leaves the grid
      t1 = fminf(fmaxf(t1, 0.0f), g.gmax);
      t2 = fminf(fmaxf(t2, 0.0f), g.gmax);
      const int f0 = (int)t0, f1 = (int)t1;
      const float d0 = t0 - (float)f0, d1 = t1 - (float)f1, d2 = t2 - (float)(int)t2;
      const float wx = in ? lagrange_node(node - f0, d0) : 0.0f;
      const float wy = lagrange_node(node - f1, d1);
      const float wz[4] = {lg2(1.0f + d2), lg1(d2), lg1(1.0f - d2), lg2(2.0f - d2)};
      const float q1 = y.x - g.center[0], q2 = y.y - g.center[1], q3 = y.z - g.center[2];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float b = wy * wz[r];
        acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx, b, acc[r][0], 0, 0, 0);
        acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx, q1 * b, acc[r][1], 0, 0, 0);
        acc[r][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx, q2 * b, acc[r][2], 0, 0, 0);
        acc[r][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx, q3 * b, acc[r][3], 0, 0, 0);
      }
    }
    p0 = re;  // the next run starts right after this one
  }
  // the windows left open, added in wave order
  for (int w = 0; w < 4; ++w) {
    __syncthreads();
    if (wave == w && pa < pb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) put(r, zb + r, true);
    }
  }
  __syncthreads();
  float *out = part + (int64_t)blockIdx.x * 4 * G;
  for (int e = tid; e < 4 * G; e += kMfmaBlock) out[e] = grid[e];
}

// ---- S2G with global float atomics (grids too large for LDS) -------------------
// Same lane-per-node mapping; the grid is large here, so contention is low.
template <int D>
__global__ __launch_bounds__(kBlock) void s2g_atomic_kernel(GridParams g, int n,
                                                            const float *__restrict__ Y,
                                                            float *__restrict__ V) {
  constexpr int PPW = NodeLanes<D>::PPW;
  constexpr int NODES = NodeLanes<D>::NODES;
  const int G = grid_cells<D>(g.ng);
  const int lane = threadIdx.x & 63;
  const int node = lane % NODES, slot = lane / NODES;
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int base = gw * PPW; base < n; base += nw * PPW) {
    const int i = base + slot;
    if (i < n) {
      float y[D];
      load_pt<D>(Y, i, y);
      int idx;
      float ww;
      node_of_lane<D>(g, y, node, idx, ww);
      unsafeAtomicAdd(&V[idx], ww);
#pragma unroll
      for (int a = 0; a < D; ++a) unsafeAtomicAdd(&V[(a + 1) * G + idx], (y[a] - g.center[a]) * ww);
    }
  }
}

// Sum of the per-block private grids (deterministic): 64 cells per block,
// 16 partial groups per cell summed in a fixed order, then combined in LDS.
__global__ __launch_bounds__(1024) void grid_reduce_kernel(const float *__restrict__ part, int nparts,
                                                           int cells, float *__restrict__ V) {
  __shared__ float red[16][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int gidx = threadIdx.x >> 6;
  float acc = 0.0f;
  if (c < cells) {
    int b = gidx;
    for (; b + 48 < nparts; b += 64) {
      const float v0 = part[(int64_t)b * cells + c], v1 = part[(int64_t)(b + 16) * cells + c];
      const float v2 = part[(int64_t)(b + 32) * cells + c], v3 = part[(int64_t)(b + 48) * cells + c];
      acc += (v0 + v1) + (v2 + v3);
    }
    for (; b < nparts; b += 16) acc += part[(int64_t)b * cells + c];
  }
  red[gidx][threadIdx.x & 63] = acc;
  __syncthreads();
  if (threadIdx.x < 64 && c < cells) {
    float t = 0.0f;
    for (int k = 0; k < 16; ++k) t += red[k][threadIdx.x];
    V[c] = t;
  }
}

// ---- G2S + forces + Z ------------------------------------------------------------
// Phi_c(y_i) by the same stencil; Frep_i (unnormalised) = ~y_i Phi0 - Phi_{1..d}.
// Z per point, compensated sum over points, FP64:
//   ZDIRECT: z_i = psi_i, the gathered convolution of the unit charges with
//            1/(1 + r^2) (channel D+1), so Z = sum_i psi_i - N;
//   else:    z_i = (1 + 2|~y_i|^2) Phi0 - 2 ~y_i . Phi_{1..d} (HY/qq.cpp:75-106).
// ~y is y minus the charge shift used by S2G (any shift leaves Frep and
// sum z_i unchanged).
template <int D, bool ZDIRECT>
__global__ __launch_bounds__(kBlock) void g2s_force_kernel(GridParams g, int n,
                                                           const float *__restrict__ Y,
                                                           const float *__restrict__ Phi,
                                                           float *__restrict__ Frep, double *part) {
  __shared__ double lds[8];
  const int ng = g.ng;
  const int G = grid_cells<D>(ng);
  constexpr int C = ZDIRECT ? D + 2 : D + 1;
  Neu z{0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float y[D];
    load_pt<D>(Y, i, y);
    int f[D];
    float w[D][4];
    stencil<D>(g, y, f, w);
    float ph[C];
#pragma unroll
    for (int c = 0; c < C; ++c) ph[c] = 0.0f;
    for_stencil<D>(ng, f, w, [&](int idx, float ww) {
#pragma unroll
      for (int c = 0; c < C; ++c) ph[c] = __builtin_fmaf(Phi[c * G + idx], ww, ph[c]);
    });
    float fr[D];
    double ysq = 0.0, ydot = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float yt = y[a] - g.center[a];
      fr[a] = yt * ph[0] - ph[a + 1];
      ysq += (double)yt * (double)yt;
      ydot += (double)yt * (double)ph[a + 1];
    }
    store_pt<D>(Frep, i, fr);
    if constexpr (ZDIRECT) neu_add(z, (double)ph[D + 1]);
    else neu_add(z, (1.0 + 2.0 * ysq) * (double)ph[0] - 2.0 * ydot);
  }
  z = wave_neu(z);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { lds[wid * 2] = z.s; lds[wid * 2 + 1] = z.c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu t{lds[0], lds[1]};
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) neu_merge(t, Neu{lds[2 * k], lds[2 * k + 1]});
    part[(size_t)blockIdx.x * 8 + 0] = t.s;
    part[(size_t)blockIdx.x * 8 + 1] = t.c;
  }
}

// ---- G2S + forces + Z with the grid in LDS (2-D side <= 64, 3-D side <= 16) --------
// The same arithmetic, in the same order, as g2s_force_kernel, with the 4^D
// stencil reads of a point served from LDS instead of L1/L2: each work-group
// (one per CU, 16 waves) first copies the grid into LDS, channels 0..3 of a
// node as one float4 and the fifth (Z-direct in 3-D) in a second array, then
// strides over the points.
constexpr int kG2sLdsCells = 4096;
constexpr int kG2sLdsBlock = 1024;

inline bool g2s_lds_fits(int d, int ng) { return (d == 2 && ng <= 64) || (d == 3 && ng <= 16); }

template <int D, bool ZDIRECT>
__global__ __launch_bounds__(kG2sLdsBlock) void g2s_lds_kernel(GridParams g, int n, const float *__restrict__ Y,
                                                               const float *__restrict__ Phi,
                                                               float *__restrict__ Frep, double *part) {
  constexpr int C = ZDIRECT ? D + 2 : D + 1;
  static_assert(C >= 3 && C <= 5, "2-D or 3-D only");
  __shared__ float4 q4[kG2sLdsCells];
  __shared__ float q5[C == 5 ? kG2sLdsCells : 1];
  __shared__ double lds[2 * (kG2sLdsBlock / 64)];
  const int ng = g.ng;
  const int G = grid_cells<D>(ng);
  for (int e = threadIdx.x; e < G; e += blockDim.x) {
    float4 v;
    v.x = Phi[e];
    v.y = Phi[G + e];
    v.z = Phi[2 * G + e];
    v.w = C > 3 ? Phi[3 * G + e] : 0.0f;
    q4[e] = v;
    if constexpr (C == 5) q5[e] = Phi[4 * G + e];
  }
  __syncthreads();
  Neu z{0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float y[D];
    load_pt<D>(Y, i, y);
    int f[D];
    float w[D][4];
    stencil<D>(g, y, f, w);
    float ph[C];
#pragma unroll
    for (int c = 0; c < C; ++c) ph[c] = 0.0f;
    auto node = [&](int idx, float ww) {
      const float4 v = q4[idx];
      ph[0] = __builtin_fmaf(v.x, ww, ph[0]);
      ph[1] = __builtin_fmaf(v.y, ww, ph[1]);
      ph[2] = __builtin_fmaf(v.z, ww, ph[2]);
      if constexpr (C > 3) ph[3] = __builtin_fmaf(v.w, ww, ph[3]);
      if constexpr (C == 5) ph[4] = __builtin_fmaf(q5[idx], ww, ph[4]);
    };
    if constexpr (D == 3) {
      // for_stencil's order and weights, one z-plane of 16 nodes at a time
      // (the whole stencil unrolled would spill)
#pragma unroll 1
      for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
        for (int i1 = 0; i1 < 4; ++i1) {
          const float w21 = w[2][i2] * w[1][i1];
#pragma unroll
          for (int i0 = 0; i0 < 4; ++i0) node(((f[2] + i2) * ng + f[1] + i1) * ng + f[0] + i0, w21 * w[0][i0]);
        }
    } else {
      for_stencil<D>(ng, f, w, node);
    }
    float fr[D];
    double ysq = 0.0, ydot = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float yt = y[a] - g.center[a];
      fr[a] = yt * ph[0] - ph[a + 1];
      ysq += (double)yt * (double)yt;
      ydot += (double)yt * (double)ph[a + 1];
    }
    store_pt<D>(Frep, i, fr);
    if constexpr (ZDIRECT) neu_add(z, (double)ph[D + 1]);
    else neu_add(z, (1.0 + 2.0 * ysq) * (double)ph[0] - 2.0 * ydot);
  }
  z = wave_neu(z);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { lds[wid * 2] = z.s; lds[wid * 2 + 1] = z.c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu t{lds[0], lds[1]};
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) neu_merge(t, Neu{lds[2 * k], lds[2 * k + 1]});
    part[(size_t)blockIdx.x * 8 + 0] = t.s;
    part[(size_t)blockIdx.x * 8 + 1] = t.c;
  }
}

// ---- fused gradient + gains + momentum update ----------------------------------
// dy = alpha Fattr - Frep / Z; gains, uy, y exactly as CU/gradient_descend.cu:29-44;
// per-block sums of the new Y feed the mean of the next centring.

template <int D>
__global__ __launch_bounds__(kBlock) void update_kernel(int n, float *__restrict__ Y,
                                                        float *__restrict__ uY,
                                                        float *__restrict__ gains,
                                                        const float *__restrict__ Fattr,
                                                        const float *__restrict__ Frep,
                                                        float alpha, float momentum, float eta,
                                                        const DevScalars *sc, double *part) {
  __shared__ double lds[D * 4];
  const float zinv = (float)(1.0 / sc->Z);
  double s[D];
#pragma unroll
  for (int a = 0; a < D; ++a) s[a] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float y[D], u[D], gn[D], fa[D], fr[D];
    load_pt<D>(Y, i, y);
    load_pt<D>(uY, i, u);
    load_pt<D>(gains, i, gn);
    load_pt<D>(Fattr, i, fa);
    load_pt<D>(Frep, i, fr);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float dy = alpha * fa[a] - fr[a] * zinv;
      float gg = (sgnf(dy) != sgnf(u[a])) ? (gn[a] + 0.2f) : (gn[a] * 0.8f);
      if (gg < 0.01f) gg = 0.01f;
      gn[a] = gg;
      u[a] = momentum * u[a] - eta * gg * dy;
      y[a] = y[a] + u[a];
      s[a] += (double)y[a];
    }
    store_pt<D>(Y, i, y);
    store_pt<D>(uY, i, u);
    store_pt<D>(gains, i, gn);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < D; ++a) s[a] = wave_sum(s[a]);
  if (lane == 0)
#pragma unroll
    for (int a = 0; a < D; ++a) lds[a * 4 + wid] = s[a];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int a = 0; a < D; ++a)
      part[(size_t)blockIdx.x * 8 + a] = lds[a * 4] + lds[a * 4 + 1] + lds[a * 4 + 2] + lds[a * 4 + 3];
}

// ---- exact O(N^2) repulsion (validation, CU/compute_error.cu:8-53) -------------
template <int D>
__global__ __launch_bounds__(kBlock) void rep_exact_kernel(int n, const float *__restrict__ Y,
                                                           float *__restrict__ Frep, double *part) {
  __shared__ float tile[kBlock][D];
  __shared__ double lds[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float yi[D];
#pragma unroll
  for (int a = 0; a < D; ++a) yi[a] = 0.0f;
  if (i < n) load_pt<D>(Y, i, yi);
  double f[D];
#pragma unroll
  for (int a = 0; a < D; ++a) f[a] = 0.0;
  double z = 0.0;
  for (int j0 = 0; j0 < n; j0 += kBlock) {
    const int j = j0 + threadIdx.x;
    float yj[D];
#pragma unroll
    for (int a = 0; a < D; ++a) yj[a] = 0.0f;
    if (j < n) load_pt<D>(Y, j, yj);
#pragma unroll
    for (int a = 0; a < D; ++a) tile[threadIdx.x][a] = yj[a];
    __syncthreads();
    const int cnt = min(kBlock, n - j0);
    for (int t = 0; t < cnt; ++t) {
      if (j0 + t == i) continue;
      double dist = 0.0, df[D];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        df[a] = (double)yi[a] - (double)tile[t][a];
        dist += df[a] * df[a];
      }
      const double q = 1.0 / (1.0 + dist);
#pragma unroll
      for (int a = 0; a < D; ++a) f[a] += df[a] * q * q;
      z += q;
    }
    __syncthreads();
  }
  if (i < n) {
    float fo[D];
#pragma unroll
    for (int a = 0; a < D; ++a) fo[a] = (float)f[a];
    store_pt<D>(Frep, i, fo);
  } else {
    z = 0.0;
  }
  z = wave_sum(z);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = z;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8] = lds[0] + lds[1] + lds[2] + lds[3];
}

// ---- one-block finalisation of the block partials (fixed order) ---------------
__global__ __launch_bounds__(kBlock) void finalize_kernel(int mode, int nblocks, int d, int n,
                                                          int apply_mean, const double *__restrict__ part,
                                                          DevScalars *sc) {
  __shared__ double red[8][kBlock];
  __shared__ double redc[kBlock];
  const int t = threadIdx.x;
  double acc[8];
  Neu zn{0.0, 0.0};
  for (int k = 0; k < 8; ++k) acc[k] = (mode == FIN_MINMAX) ? (k < 4 ? INFINITY : -INFINITY) : 0.0;
  if (mode == FIN_MINMAX) acc[3] = 0.0;
  for (int b = t; b < nblocks; b += kBlock) {
    const double *pb = part + (size_t)b * 8;
    if (mode == FIN_MINMAX) {
      for (int a = 0; a < d; ++a) {
        acc[a] = fmin(acc[a], pb[a]);
        acc[4 + a] = fmax(acc[4 + a], pb[4 + a]);
      }
      acc[3] += pb[3];
    } else if (mode == FIN_ZNEU) {
      neu_merge(zn, Neu{pb[0], pb[1]});
    } else {
      for (int a = 0; a < 4; ++a) acc[a] += pb[a];
    }
  }
  if (mode == FIN_ZNEU) { acc[0] = zn.s; acc[1] = zn.c; }
  for (int k = 0; k < 8; ++k) red[k][t] = acc[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) {
      if (mode == FIN_MINMAX) {
        for (int a = 0; a < 3; ++a) {
          red[a][t] = fmin(red[a][t], red[a][t + s]);
          red[4 + a][t] = fmax(red[4 + a][t], red[4 + a][t + s]);
        }
        red[3][t] += red[3][t + s];
      } else if (mode == FIN_ZNEU) {
        Neu x{red[0][t], red[1][t]};
        neu_merge(x, Neu{red[0][t + s], red[1][t + s]});
        red[0][t] = x.s;
        red[1][t] = x.c;
      } else {
        for (int a = 0; a < 4; ++a) red[a][t] += red[a][t + s];
      }
    }
    __syncthreads();
  }
  (void)redc;
  if (t == 0) {
    if (mode == FIN_MINMAX) {
      const bool bad = red[3][0] > 0.0;
      for (int a = 0; a < d; ++a) {
        sc->minv[a] = bad ? NAN : red[a][0];  // a non-finite coordinate poisons the extent
        sc->maxv[a] = red[4 + a][0];
        if (apply_mean) sc->mean[a] = 0.0;
      }
    } else if (mode == FIN_ZNEU) {
      sc->Z = (red[0][0] + red[1][0]) - (double)n;
    } else if (mode == FIN_MEAN) {
      for (int a = 0; a < d; ++a) {
        sc->sum[a] = red[a][0];
        sc->mean[a] = red[a][0] / (double)n;
      }
    } else if (mode == FIN_ZEXACT) {
      sc->zexact = red[0][0];
    } else if (mode == FIN_KL) {
      sc->kl = red[0][0];
    }
  }
}

int grid_blocks(int n, int max_blocks) {
  int b = div_up(n > 0 ? n : 1, kBlock);
  return b < max_blocks ? b : max_blocks;
}

int finalize(hipStream_t s, int mode, int nblocks, int d, int n, int apply_mean, RedWork rw, DevScalars *sc) {
  finalize_kernel<<<1, kBlock, 0, s>>>(mode, nblocks, d, n, apply_mean, rw.part, sc);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

// Large grids (those no LDS copy or matrix-core tile holds): the points in
// runs of `unit` consecutive ones (the engine's locality order), each run's
// stencil window accumulated in LDS as 64-bit fixed point and added to the
// fixed-point grid Vf (spread_fix.hpp), so the grid is exact and the same
// whatever the order of the atomics.
template <int D>
__global__ __launch_bounds__(s2g_threads<D>()) void s2g_fix_kernel(GridParams g, float fsf, int n,
                                                                   const float *__restrict__ Y,
                                                                   unsigned long long *__restrict__ Vf, int unit,
                                                                   int lds_slots) {
  extern __shared__ unsigned long long win[];
  __shared__ int sbox[6];
  const int nunits = (n + unit - 1) / unit;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x)
    spread_fix_unit<D>(g, fsf, Vf, u * unit, min(n, (u + 1) * unit), Y, win, lds_slots, sbox);
}

// V = Vf * 2^-e in FP32 (the FP64 product rounded once), Vf cleared.
__global__ void fix_to_float_kernel(long long *__restrict__ Vf, float *__restrict__ V, int64_t count, double inv) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    V[i] = (float)((double)Vf[i] * inv);
    Vf[i] = 0;
  }
}

}  // namespace

int launch_stream_copy(hipStream_t s, const float4 *in, float4 *out, int64_t count) {
  if (count <= 0) return SGTSNE_OK;
  stream_copy_kernel<<<2048, kBlock, 0, s>>>(in, out, count);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_fill(hipStream_t s, float *p, int64_t count, float v) {
  if (count <= 0) return SGTSNE_OK;
  int blocks = (int)std::min<int64_t>((count + kBlock - 1) / kBlock, 4096);
  fill_kernel<<<blocks, kBlock, 0, s>>>(p, count, v);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_center_minmax(hipStream_t s, int d, int n, float *Y, int apply_mean, RedWork rw,
                         DevScalars *sc) {
  const int blocks = grid_blocks(n, std::min(rw.max_blocks, 1024));
  switch (d) {
    case 1: center_minmax_kernel<1><<<blocks, kBlock, 0, s>>>(Y, n, apply_mean, rw.part, sc); break;
    case 2: center_minmax_kernel<2><<<blocks, kBlock, 0, s>>>(Y, n, apply_mean, rw.part, sc); break;
    case 3: center_minmax_kernel<3><<<blocks, kBlock, 0, s>>>(Y, n, apply_mean, rw.part, sc); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return finalize(s, FIN_MINMAX, blocks, d, n, apply_mean, rw, sc);
}

// LDS grid copies for the spread: one per point slot of a wave when they fit.
int s2g_lds_fits(int d, int ng, size_t *lds_bytes, int *copies) {
  const size_t G = d == 1 ? (size_t)ng : d == 2 ? (size_t)ng * ng : (size_t)ng * ng * ng;
  const size_t one = G * (size_t)(d + 1) * sizeof(float);
  int c = d == 1 ? 16 : d == 2 ? 4 : 1;
  while (c > 1 && one * c > 65536) c >>= 1;
  *copies = c;
  *lds_bytes = one * c;
  return *lds_bytes <= 65536;
}

int64_t s2g_scratch_need(const GridParams &g, int n, int mode) {
  int64_t need = 0;
  (void)launch_s2g(nullptr, g, -n, nullptr, nullptr, nullptr, 0, &need, mode);
  return need;
}

int launch_s2g(hipStream_t s, const GridParams &g, int n, const float *Y, float *V, float *scratch,
               int64_t scratch_floats, int64_t *needed, int mode) {
  const bool query = n < 0;  // size query only
  if (query) n = -n;
  if (n > 0 && g.d == 2 && g.ng <= 32 && mode != SGTSNE_SPREAD_ATOMIC) {
    // matrix cores: no contention however the points crowd the few cells
    const int64_t cells = 3ll * g.ng * g.ng;
    const int chunk = mfma_chunk(n);
    const int blocks = div_up(n, chunk);
    if (needed) *needed = (int64_t)blocks * cells;
    if (query) return SGTSNE_OK;
    SGT_CHECK((int64_t)blocks * cells <= scratch_floats, SGTSNE_ERR_STATE, "spread scratch too small");
    if (g.ng <= 16)
      s2g_mfma_kernel<1><<<blocks, kMfmaBlock, 0, s>>>(g, n, Y, scratch, chunk);
    else
      s2g_mfma_kernel<2><<<blocks, kMfmaBlock, 0, s>>>(g, n, Y, scratch, chunk);
    SGT_LAUNCH_CHECK();
    grid_reduce_kernel<<<div_up(cells, 64), 1024, 0, s>>>(scratch, blocks, (int)cells, V);
    SGT_LAUNCH_CHECK();
    return SGTSNE_OK;
  }
  if (n > 0 && g.d == 3 && g.ng <= 16 && mode != SGTSNE_SPREAD_ATOMIC) {
    const int64_t cells = 4ll * g.ng * g.ng * g.ng;
    const int blocks = div_up(n, kMfma3Chunk);
    if (needed) *needed = (int64_t)blocks * cells;
    if (query) return SGTSNE_OK;
    SGT_CHECK((int64_t)blocks * cells <= scratch_floats, SGTSNE_ERR_STATE, "spread scratch too small");
    s2g_mfma3_kernel<<<blocks, kMfmaBlock, 0, s>>>(g, n, Y, scratch, kMfma3Chunk);
    SGT_LAUNCH_CHECK();
    grid_reduce_kernel<<<div_up(cells, 64), 1024, 0, s>>>(scratch, blocks, (int)cells, V);
    SGT_LAUNCH_CHECK();
    return SGTSNE_OK;
  }
  size_t lds = 0;
  int copies = 1;
  const int64_t cells = (int64_t)(g.d == 1 ? g.ng : g.d == 2 ? g.ng * g.ng : g.ng * g.ng * g.ng) * (g.d + 1);
  if (n > 0 && g.d >= 2 && mode != SGTSNE_SPREAD_ATOMIC && !s2g_lds_fits(g.d, g.ng, &lds, &copies)) {
    // fixed point in the scratch (2 floats per 64-bit cell): cleared, spread,
    // converted into V and cleared again
    if (needed) *needed = 2 * cells;
    if (query) return SGTSNE_OK;
    SGT_CHECK(2 * cells <= scratch_floats, SGTSNE_ERR_STATE, "spread scratch too small");
    // no cell can exceed 2^61 in magnitude: |weight| <= 1 and the charges
    // are within half the extent of the box centre
    const double maxy = (double)(g.nGrid - 1) / (double)g.scale * (1.0 + 1e-6);
    const double bound = 2.0 * (double)n * (0.5 * maxy + 1.0);
    const int e = std::min(60, std::max(1, 61 - std::ilogb(bound)));
    unsigned long long *Vf = reinterpret_cast<unsigned long long *>(scratch);
    SGT_TRY(launch_fill(s, scratch, 2 * cells, 0.0f));
    const int unit = g.d == 3 ? kFixUnit3 : 1024;
    const int lds_b = 64512;  // two work-groups per CU
    const int blocks = (int)std::min<int64_t>(div_up(n, unit), 2048);
    if (g.d == 2)
      s2g_fix_kernel<2><<<blocks, s2g_threads<2>(), lds_b, s>>>(g, std::ldexp(1.0f, e), n, Y, Vf, unit, lds_b / 8);
    else
      s2g_fix_kernel<3><<<blocks, s2g_threads<3>(), lds_b, s>>>(g, std::ldexp(1.0f, e), n, Y, Vf, unit, lds_b / 8);
    SGT_LAUNCH_CHECK();
    fix_to_float_kernel<<<(int)std::min<int64_t>(div_up(cells, 256), 4096), 256, 0, s>>>(
        reinterpret_cast<long long *>(Vf), V, cells, std::ldexp(1.0, -e));
    SGT_LAUNCH_CHECK();
    return SGTSNE_OK;
  }
  int blocks = 0, chunk = 0;
  if (n > 0 && s2g_lds_fits(g.d, g.ng, &lds, &copies)) {
    // about 64 points per wave, at most 1024 private grids and 64 MiB of them
    int64_t b = std::max<int64_t>(1, div_up(n, (kS2GBlock / 64) * 64));
    b = std::min<int64_t>(b, 1024);
    b = std::min<int64_t>(b, std::max<int64_t>(1, (int64_t)(64ll << 20) / (cells * 4)));
    chunk = div_up(n, b);
    blocks = div_up(n, chunk);
  }
  if (needed) *needed = blocks > 0 ? (int64_t)blocks * cells : 0;
  if (query) return SGTSNE_OK;
  if (blocks > 0 && (int64_t)blocks * cells <= scratch_floats) {
    switch (g.d) {
      case 1: s2g_lds_kernel<1><<<blocks, kS2GBlock, lds, s>>>(g, n, Y, scratch, chunk, copies); break;
      case 2: s2g_lds_kernel<2><<<blocks, kS2GBlock, lds, s>>>(g, n, Y, scratch, chunk, copies); break;
      case 3: s2g_lds_kernel<3><<<blocks, kS2GBlock, lds, s>>>(g, n, Y, scratch, chunk, copies); break;
      default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
    }
    SGT_LAUNCH_CHECK();
    grid_reduce_kernel<<<div_up(cells, 64), 1024, 0, s>>>(scratch, blocks, (int)cells, V);
  } else {
    SGT_TRY(launch_fill(s, V, cells, 0.0f));
    const int ppb = (kBlock / 64) * (g.d == 1 ? 16 : g.d == 2 ? 4 : 1);  // points per block per step
    const int ab = (int)std::min<int64_t>(div_up(n > 0 ? n : 1, ppb), 65536);
    switch (g.d) {
      case 1: s2g_atomic_kernel<1><<<ab, kBlock, 0, s>>>(g, n, Y, V); break;
      case 2: s2g_atomic_kernel<2><<<ab, kBlock, 0, s>>>(g, n, Y, V); break;
      case 3: s2g_atomic_kernel<3><<<ab, kBlock, 0, s>>>(g, n, Y, V); break;
      default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
    }
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_g2s_force(hipStream_t s, const GridParams &g, int n, const float *Y, const float *Phi,
                     float *Frep, int zdirect, RedWork rw, DevScalars *sc) {
  if (g2s_lds_fits(g.d, g.ng)) {
    // one 16-wave work-group per CU (the grid copy takes half of its LDS)
    const int lb = std::max(1, std::min({rw.max_blocks, 256, div_up(n > 0 ? n : 1, kG2sLdsBlock)}));
    if (g.d == 2) {
      if (zdirect) g2s_lds_kernel<2, true><<<lb, kG2sLdsBlock, 0, s>>>(g, n, Y, Phi, Frep, rw.part);
      else g2s_lds_kernel<2, false><<<lb, kG2sLdsBlock, 0, s>>>(g, n, Y, Phi, Frep, rw.part);
    } else {
      if (zdirect) g2s_lds_kernel<3, true><<<lb, kG2sLdsBlock, 0, s>>>(g, n, Y, Phi, Frep, rw.part);
      else g2s_lds_kernel<3, false><<<lb, kG2sLdsBlock, 0, s>>>(g, n, Y, Phi, Frep, rw.part);
    }
    SGT_LAUNCH_CHECK();
    return finalize(s, FIN_ZNEU, lb, g.d, n, 0, rw, sc);
  }
  const int blocks = grid_blocks(n, std::min(rw.max_blocks, 2048));
#define G2S_CASE(DD)                                                                               \
  case DD:                                                                                         \
    if (zdirect) g2s_force_kernel<DD, true><<<blocks, kBlock, 0, s>>>(g, n, Y, Phi, Frep, rw.part); \
    else g2s_force_kernel<DD, false><<<blocks, kBlock, 0, s>>>(g, n, Y, Phi, Frep, rw.part);        \
    break;
  switch (g.d) {
    G2S_CASE(1)
    G2S_CASE(2)
    G2S_CASE(3)
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
#undef G2S_CASE
  SGT_LAUNCH_CHECK();
  return finalize(s, FIN_ZNEU, blocks, g.d, n, 0, rw, sc);
}

int launch_finalize_kl(hipStream_t s, int nblocks, RedWork rw, DevScalars *sc) {
  return finalize(s, FIN_KL, nblocks, 1, 0, 0, rw, sc);
}

int launch_update(hipStream_t s, int d, int n, float *Y, float *uY, float *gains,
                  const float *Fattr, const float *Frep, float alpha, float momentum, float eta,
                  RedWork rw, DevScalars *sc) {
  const int blocks = grid_blocks(n, std::min(rw.max_blocks, 2048));
  switch (d) {
    case 1: update_kernel<1><<<blocks, kBlock, 0, s>>>(n, Y, uY, gains, Fattr, Frep, alpha, momentum, eta, sc, rw.part); break;
    case 2: update_kernel<2><<<blocks, kBlock, 0, s>>>(n, Y, uY, gains, Fattr, Frep, alpha, momentum, eta, sc, rw.part); break;
    case 3: update_kernel<3><<<blocks, kBlock, 0, s>>>(n, Y, uY, gains, Fattr, Frep, alpha, momentum, eta, sc, rw.part); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return finalize(s, FIN_MEAN, blocks, d, n, 0, rw, sc);
}

int launch_rep_exact(hipStream_t s, int d, int n, const float *Y, float *Frep, RedWork rw,
                     DevScalars *sc) {
  const int blocks = div_up(n > 0 ? n : 1, kBlock);
  SGT_CHECK(blocks <= rw.max_blocks, SGTSNE_ERR_ARG, "exact repulsion: n too large for workspace");
  switch (d) {
    case 1: rep_exact_kernel<1><<<blocks, kBlock, 0, s>>>(n, Y, Frep, rw.part); break;
    case 2: rep_exact_kernel<2><<<blocks, kBlock, 0, s>>>(n, Y, Frep, rw.part); break;
    case 3: rep_exact_kernel<3><<<blocks, kBlock, 0, s>>>(n, Y, Frep, rw.part); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return finalize(s, FIN_ZEXACT, blocks, d, n, 0, rw, sc);
}

}  // namespace sgt
