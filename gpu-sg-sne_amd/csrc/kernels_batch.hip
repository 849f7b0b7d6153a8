// kernels_batch.hip -- one launch per phase for every job of a batch.
//
// The phases and their reference semantics are those of the single-job
// kernels (kernels_grid.hip, kernels_fft.hip): spread HY/gridding.cpp:164-211,
// convolution HY/non_periodic_conv.cpp, gather + forces HY/qq.cpp:75-106,
// update CU/gradient_descend.cu:12-44 with the schedule of
// CU/gradient_descend.cu:102-245, centring HY/gradient_descendHYB.cpp:64-76,
// grid mapping HY/qq.cpp:147-166.
//
// Work is cut into items of at most kBatchChunk consecutive points of one
// job; work-groups are persistent and stride over the items (or, for the
// FFT passes, over the line blocks of every running job, found by a binary
// search in a per-pass prefix table the plan kernel writes).
#include "common.hpp"
#include "kernels.hpp"
#include "grid_device.hpp"
#include "fft_device.hpp"
#include "spread_fix.hpp"
#include "batch.hpp"
#include "../../include/sgtsne.h"

#include <climits>

namespace sgt {

namespace {

using namespace fft;

constexpr int kBT = 256;  // threads of the point kernels

// threads of the gather: in 3-D one per point of a work item, so the 64-cell
// stencil reads of a whole item are in flight at once (measured: 181 -> 120
// us on the served mix); in 1-D and 2-D the 256-thread form is faster
template <int D>
constexpr int g2s_threads() { return D == 3 ? kBatchChunk : kBT; }

__constant__ int kGridSizesDev[] = {SGT_GRID_SIZE_LIST};
constexpr int kGridSizesHost[] = {SGT_GRID_SIZE_LIST};
constexpr int kGridSizesCount = sizeof(kGridSizesHost) / sizeof(kGridSizesHost[0]);

// ---- S2G: 64-bit fixed point, LDS window of the item's stencils -------------
// (spread_fix.hpp).  Work unit = one sub-chunk of an item: 3-D stencil windows
// grow fast with the point count, so 3-D items are spread in quarters, each
// window its own.
static_assert(kBatchChunk / 4 <= kFixUnit3, "a 3-D work unit exceeds the spread's batch");

template <int D>
__global__ __launch_bounds__(s2g_threads<D>()) void bs2g_kernel(const BatchJob *__restrict__ jobs,
                                                                const BatchItem *__restrict__ items, int nitems,
                                                                const float *__restrict__ Y, int lds_slots) {
  extern __shared__ unsigned long long win[];
  __shared__ int sbox[6];
  constexpr int SUB = D == 3 ? 4 : 1;
  for (int wu = blockIdx.x; wu < nitems * SUB; wu += gridDim.x) {
    const int it = wu / SUB, part = wu - it * SUB;
    BatchItem itm = items[it];
    const BatchJob *J = jobs + itm.job;
    if (J->status != BJ_RUN) continue;  // the same value for every thread of the block
    if (SUB > 1) {
      const int len = itm.end - itm.begin, step = (len + SUB - 1) / SUB;
      itm.begin += part * step;
      itm.end = min(itm.end, itm.begin + step);
      if (itm.begin >= itm.end) continue;
    }
    spread_fix_unit<D>(J->g, (float)J->fix_scale, reinterpret_cast<unsigned long long *>(J->vfix), itm.begin,
                       itm.end, Y, win, lds_slots, sbox);
  }
}

// ---- FFT passes over every running job -----------------------------------------
template <int IN, int OUT, int MID, bool EC>
__global__ __launch_bounds__(kFftThreads) void bpass_kernel(const PassArgs *__restrict__ args,
                                                            const int *__restrict__ prefix, int njobs,
                                                            int npasses, int pass) {
  extern __shared__ float2 lds[];
  const int *pf = prefix + (size_t)pass * (njobs + 1);
  const int total = pf[njobs];
  for (int u = blockIdx.x; u < total; u += gridDim.x) {
    int lo = 0, hi = njobs;  // pf[lo] <= u < pf[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pf[mid] <= u) lo = mid;
      else hi = mid;
    }
    // read in place: the block is uniform, so its fields come in through
    // scalar loads instead of a per-lane private copy
    const PassArgs &a = args[(size_t)lo * npasses + pass];
    pass_body<IN, OUT, MID, EC>(a, u - pf[lo], blockIdx.y, lds);
  }
}

// ---- G2S + forces + Z partials -------------------------------------------------
template <int D, bool ZDIRECT>
__global__ __launch_bounds__(g2s_threads<D>()) void bg2s_kernel(const BatchJob *__restrict__ jobs,
                                                   const BatchItem *__restrict__ items, int nitems,
                                                   const float *__restrict__ Y, float *__restrict__ Frep,
                                                   double *__restrict__ zpart) {
  __shared__ double lds[2 * (g2s_threads<D>() / 64)];
  constexpr int C = ZDIRECT ? D + 2 : D + 1;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem itm = items[it];
    const BatchJob *J = jobs + itm.job;
    if (J->status != BJ_RUN) continue;
    const GridParams g = J->g;
    // potentials node-major: channels 0-3 of a node as one float4, channel 4
    // (the direct Z term in 3-D) as a plane after them
    const float4 *Phi4 = reinterpret_cast<const float4 *>(J->phi);
    const int ng = g.ng;
    const int G = grid_cells<D>(ng);
    const float *Phi1 = J->phi + 4 * (size_t)G;
    Neu z{0.0, 0.0};
    for (int i = itm.begin + threadIdx.x; i < itm.end; i += blockDim.x) {
      float y[D];
      load_pt<D>(Y, i, y);
      int f[D];
      float w[D][4];
      stencil<D>(g, y, f, w);
      float ph[C];
#pragma unroll
      for (int c = 0; c < C; ++c) ph[c] = 0.0f;
      for_stencil<D>(ng, f, w, [&](int idx, float ww) {
        const float4 q = Phi4[idx];
        const float qc[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int c = 0; c < C; ++c) ph[c] += (c < 4 ? qc[c] : Phi1[idx]) * ww;
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
      zpart[(size_t)it * 2 + 0] = t.s;
      zpart[(size_t)it * 2 + 1] = t.c;
    }
    __syncthreads();
  }
}

// ---- fused update (Z of the job from its item partials) --------------------------
// Z of one job from its item partials: lane-strided compensated sums, then a
// fixed butterfly.  Called by the 64 lanes of one wavefront.
__device__ inline double job_zeta(const BatchJob *J, const double *__restrict__ zpart) {
  Neu t{0.0, 0.0};
  for (int k = threadIdx.x & 63; k < J->nitems; k += 64) {
    const size_t q = (size_t)(J->item0 + k) * 2;
    neu_merge(t, Neu{zpart[q], zpart[q + 1]});
  }
  t = wave_neu(t);
  return (t.s + t.c) - (double)J->n;
}

__global__ void bzeta_kernel(const BatchJob *__restrict__ jobs, int job, const double *__restrict__ zpart,
                             double *__restrict__ out) {
  const double Z = job_zeta(jobs + job, zpart);
  if (threadIdx.x == 0) *out = Z;
}

__device__ __forceinline__ float bsgn(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }

template <int D>
__global__ __launch_bounds__(kBT) void bupdate_kernel(BatchJob *__restrict__ jobs,
                                                      const BatchItem *__restrict__ items, int nitems,
                                                      float *__restrict__ Y, float *__restrict__ uY,
                                                      float *__restrict__ gains, const float *__restrict__ Fattr,
                                                      const float *__restrict__ Frep,
                                                      const double *__restrict__ zpart,
                                                      double *__restrict__ spart) {
  __shared__ double lds[D * 4];
  __shared__ double sZ;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem itm = items[it];
    BatchJob *J = jobs + itm.job;
    if (J->status != BJ_RUN) continue;
    if (threadIdx.x < 64) {
      const double Z = job_zeta(J, zpart);
      if (threadIdx.x == 0) {
        sZ = Z;
        if (it == J->item0) J->Z = Z;
      }
    }
    __syncthreads();
    const float zinv = (float)(1.0 / sZ);
    const int iter = J->iter;
    const float alpha = iter <= J->early_iter ? J->alpha0 : 1.0f;
    const float momentum = iter <= 250 ? 0.5f : 0.8f;
    const float eta = 200.0f;
    double s[D];
#pragma unroll
    for (int a = 0; a < D; ++a) s[a] = 0.0;
    for (int i = itm.begin + threadIdx.x; i < itm.end; i += blockDim.x) {
      float y[D], u[D], gn[D], fa[D], fr[D];
      load_pt<D>(Y, i, y);
      load_pt<D>(uY, i, u);
      load_pt<D>(gains, i, gn);
      load_pt<D>(Fattr, i, fa);
      load_pt<D>(Frep, i, fr);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const float dy = alpha * fa[a] - fr[a] * zinv;
        float gg = (bsgn(dy) != bsgn(u[a])) ? (gn[a] + 0.2f) : (gn[a] * 0.8f);
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
        spart[(size_t)it * 4 + a] = (lds[a * 4] + lds[a * 4 + 1]) + (lds[a * 4 + 2] + lds[a * 4 + 3]);
    __syncthreads();
  }
}

// ---- centring + bounding-box partials -------------------------------------------
template <int D>
__global__ __launch_bounds__(kBT) void bcenter_kernel(const BatchJob *__restrict__ jobs,
                                                      const BatchItem *__restrict__ items, int nitems,
                                                      float *__restrict__ Y, const double *__restrict__ spart,
                                                      double *__restrict__ mpart) {
  __shared__ double lds[(2 * D + 1) * 4];
  __shared__ float smean[4];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem itm = items[it];
    const BatchJob *J = jobs + itm.job;
    const int st = J->status;
    if (st != BJ_RUN && st != BJ_INIT) continue;
    if (threadIdx.x < 64) {
      double m[D];
      if (st == BJ_RUN) {
#pragma unroll
        for (int a = 0; a < D; ++a) m[a] = 0.0;
        for (int k = threadIdx.x; k < J->nitems; k += 64)
#pragma unroll
          for (int a = 0; a < D; ++a) m[a] += spart[(size_t)(J->item0 + k) * 4 + a];
#pragma unroll
        for (int a = 0; a < D; ++a) m[a] = wave_sum(m[a]) / (double)J->n;
      } else {
#pragma unroll
        for (int a = 0; a < D; ++a) m[a] = (double)J->mean0[a];
      }
      if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < D; ++a) smean[a] = (float)m[a];
    }
    __syncthreads();
    float m[D], mn[D], mx[D];
    float bad = 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a) { m[a] = smean[a]; mn[a] = INFINITY; mx[a] = -INFINITY; }
    for (int i = itm.begin + threadIdx.x; i < itm.end; i += blockDim.x) {
      float y[D];
      load_pt<D>(Y, i, y);
#pragma unroll
      for (int a = 0; a < D; ++a) y[a] -= m[a];
      store_pt<D>(Y, i, y);
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
      double *pb = mpart + (size_t)it * 8;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        pb[a] = fmin(fmin(lds[a * 4], lds[a * 4 + 1]), fmin(lds[a * 4 + 2], lds[a * 4 + 3]));
        pb[4 + a] = fmax(fmax(lds[(D + a) * 4], lds[(D + a) * 4 + 1]),
                         fmax(lds[(D + a) * 4 + 2], lds[(D + a) * 4 + 3]));
      }
      pb[3] = lds[2 * D * 4] + lds[2 * D * 4 + 1] + lds[2 * D * 4 + 2] + lds[2 * D * 4 + 3];
    }
    __syncthreads();
  }
}

// ---- plan: one wavefront per job ---------------------------------------------------
__device__ inline int pass_reads_grid(int kind) { return kind == PK_V_W || kind == PK_V_PHI; }

__global__ __launch_bounds__(256) void bplan_kernel(int d, BatchJob *__restrict__ jobs, int njobs,
                                                    const double *__restrict__ mpart, PassArgs *__restrict__ args,
                                                    int npasses, const float2 *__restrict__ tw) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= njobs) return;
  BatchJob *J = jobs + j;
  const int st = J->status;
  if (st != BJ_RUN && st != BJ_INIT) return;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double bad = 0.0;
  for (int k = lane; k < J->nitems; k += 64) {
    const double *pb = mpart + (size_t)(J->item0 + k) * 8;
    for (int a = 0; a < d; ++a) {
      mn[a] = fmin(mn[a], pb[a]);
      mx[a] = fmax(mx[a], pb[4 + a]);
    }
    bad += pb[3];
  }
  for (int a = 0; a < d; ++a) {
    mn[a] = wave_min(mn[a]);
    mx[a] = wave_max(mx[a]);
  }
  bad = wave_sum(bad);
  // the butterflies leave the same values in every lane, so every lane takes
  // the same decisions below: lane 0 writes the job record, lane p the
  // arguments of pass p (straight into the table, no private copy)
  if (st == BJ_RUN) {
    const int it = J->iter + 1;
    if (lane == 0) J->iter = it;
    if (it >= J->max_iter) {
      if (lane == 0) J->status = BJ_DONE;
      return;
    }
  } else if (lane == 0) {
    J->status = BJ_RUN;
  }
  double maxy = 0.0;
  bool finite = bad == 0.0;
  for (int a = 0; a < d; ++a) {
    finite = finite && isfinite(mn[a]) && isfinite(mx[a]);
    maxy = fmax(maxy, mx[a] - mn[a]);
    if (lane == 0) {
      J->minv[a] = mn[a];
      J->maxv[a] = mx[a];
    }
  }
  if (!finite || !(maxy > 0.0)) {
    if (lane == 0) J->status = BJ_ERR_NUMERIC;
    return;
  }
  int nGrid;
  if (J->grid_mode == SGTSNE_GRID_FIXED) {
    nGrid = best_grid_size_in(kGridSizesDev, kGridSizesCount, J->ng_fixed);
  } else {
    nGrid = max((int)ceil(maxy / J->h), 14);
    nGrid = best_grid_size_in(kGridSizesDev, kGridSizesCount, nGrid);
  }
  if (nGrid + 2 > J->ng_cap) {
    if (lane == 0) J->status = BJ_ERR_GRID;
    return;
  }
  GridParams g;
  grid_from_box(d, nGrid, mn, mx, maxy, g);
  // fixed point: |charge| <= maxy / 2 + 1 per point, |weight| <= 1, n points per cell
  const double bound = 2.0 * (double)J->n * (0.5 * maxy + 1.0);
  const int e = 61 - ilogb(bound);
  const double fix_inv = ldexp(1.0, -e);
  if (lane == 0) {
    J->g = g;
    J->fix_scale = ldexp(1.0, e);
    J->fix_inv = fix_inv;
  }
  const int np = conv_npasses(d);
  if (lane < np) {
    int kinds[kMaxPasses];
    pass_kinds(d, kinds);
    ConvPtrs ptr{J->vfix, J->phi, J->W, J->Kt1, J->Kt2, J->Kh, tw};
    PassArgs &a = args[(size_t)j * npasses + lane];
    conv_pass(g, J->zdirect != 0, ptr, lane, a);
    if (pass_reads_grid(kinds[lane])) a.fix_inv = fix_inv;
  }
}

// Prefix sums of the line blocks of every pass over the running jobs.
__global__ __launch_bounds__(256) void bprefix_kernel(const BatchJob *__restrict__ jobs, int njobs,
                                                      const PassArgs *__restrict__ args, int npasses,
                                                      int *__restrict__ prefix) {
  __shared__ int cnt[kMaxPasses][1024];
  for (int j = threadIdx.x; j < njobs; j += blockDim.x) {
    const bool run = jobs[j].status == BJ_RUN;
    for (int p = 0; p < npasses; ++p) cnt[p][j] = run ? pass_blocks(args[(size_t)j * npasses + p]) : 0;
  }
  __syncthreads();
  if (threadIdx.x < npasses) {
    const int p = threadIdx.x;
    int acc = 0;
    int *pf = prefix + (size_t)p * (njobs + 1);
    for (int j = 0; j < njobs; ++j) {
      pf[j] = acc;
      acc += cnt[p][j];
    }
    pf[njobs] = acc;
  }
}

// ---- locality order of one job ------------------------------------------------------
__device__ __forceinline__ uint32_t bspread2(uint32_t x) {
  x &= 0xffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__device__ __forceinline__ uint32_t bspread3(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ void bmorton_kernel(int d, const BatchJob *__restrict__ jobs, int job, int off, int n,
                               const float *__restrict__ Y, uint32_t *__restrict__ key, int32_t *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BatchJob *J = jobs + job;
  double ext = 0.0;
  for (int a = 0; a < d; ++a) ext = fmax(ext, J->maxv[a] - J->minv[a]);
  const float inv = ext > 0.0 ? (float)(1.0 / ext) : 0.0f;
  const float top = d == 1 ? 4294967040.0f : d == 2 ? 65535.0f : 1023.0f;
  const int ds = point_stride(d);
  uint32_t q[3] = {0, 0, 0};
  for (int a = 0; a < d; ++a) {
    float t = (Y[(int64_t)(off + i) * ds + a] - (float)J->minv[a]) * inv * top;
    t = fminf(fmaxf(t, 0.0f), top);  // NaN -> 0
    q[a] = (uint32_t)t;
  }
  uint32_t k;
  if (d == 1) k = q[0];
  else if (d == 2) k = bspread2(q[0]) | (bspread2(q[1]) << 1);
  else k = bspread3(q[0]) | (bspread3(q[1]) << 1) | (bspread3(q[2]) << 2);
  key[i] = k;
  idx[i] = i;
}

// kNN ids [n][k+1] (self first) -> CSR columns without the self column
// (CU/perplexity_search.cu:100-112); rowptr[i] = i k.
__global__ void knn_csr_kernel(int n, int k, const int32_t *__restrict__ I, int32_t *__restrict__ rowptr,
                               int32_t *__restrict__ col) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= n) rowptr[t] = (int32_t)(t * k);
  if (t < (int64_t)n * k) {
    const int64_t i = t / k, e = t - i * k;
    col[t] = I[i * (k + 1) + e + 1];
  }
}

__global__ void gather_i32_kernel(int n, const int32_t *__restrict__ order, const int32_t *__restrict__ in,
                                  int32_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[order[i]];
}

__global__ void invert_kernel(int n, const int32_t *__restrict__ order, int32_t *__restrict__ inv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) inv[order[r]] = r;
}

__global__ void iota_kernel(int n, int32_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

__global__ void to_double_kernel_b(const float *__restrict__ in, double *__restrict__ out, int64_t m) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) out[t] = (double)in[t];
}

__global__ void offset_copy_kernel(const int32_t *__restrict__ in, int32_t *__restrict__ out, int64_t count,
                                   int32_t delta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = t; e < count; e += stride) out[e] = in[e] + delta;
}

template <int D>
int bs2g_dispatch(hipStream_t s, const BatchJob *jobs, const BatchItem *items, int nitems, const float *Y, int blocks,
                  int lds_bytes) {
  bs2g_kernel<D><<<blocks, s2g_threads<D>(), lds_bytes, s>>>(jobs, items, nitems, Y, lds_bytes / 8);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

template <int IN, int OUT, int MID, bool EC>
int bpass(hipStream_t s, const PassArgs *args, const int *prefix, int njobs, int npasses, int pass, int pairs,
          int blocks) {
  dim3 grid(blocks, pairs);
  hipLaunchKernelGGL((bpass_kernel<IN, OUT, MID, EC>), grid, dim3(kFftThreads), kLdsBytes, s, args, prefix, njobs,
                     npasses, pass);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

}  // namespace

size_t batch_grid_bytes(int d, int ng_cap, size_t *vfix_b, size_t *phi_b, size_t *w_b, size_t *k_b) {
  const int L = fft_length(2 * ng_cap - 1);  // FFT lengths grow with ng: the largest one
  const size_t G = d == 1 ? (size_t)ng_cap : d == 2 ? (size_t)ng_cap * ng_cap : (size_t)ng_cap * ng_cap * ng_cap;
  *vfix_b = G * (size_t)(d + 1) * sizeof(long long);
  *phi_b = G * (size_t)(d == 3 ? 5 : 4) * sizeof(float);  // node-major float4 (+ a plane in 3-D)
  conv_workspace_bytes(d, ng_cap, L, w_b, k_b);
  if (*w_b == 0) *w_b = sizeof(float2);
  return *vfix_b + *phi_b + *w_b + 3 * *k_b;
}

size_t batch_pass_args_bytes() { return sizeof(PassArgs); }

int batch_npasses(int d) {
  int kinds[kMaxPasses];
  return pass_kinds(d, kinds);
}

int launch_bs2g(hipStream_t s, int d, const BatchJob *jobs, const BatchItem *items, int nitems, const float *Y,
                int blocks, int lds_bytes) {
  if (nitems <= 0) return SGTSNE_OK;
  SGT_CHECK(lds_bytes >= 8 && lds_bytes <= 65536, SGTSNE_ERR_ARG, "bad spread LDS size");
  switch (d) {
    case 1: return bs2g_dispatch<1>(s, jobs, items, nitems, Y, blocks, lds_bytes);
    case 2: return bs2g_dispatch<2>(s, jobs, items, nitems, Y, blocks, lds_bytes);
    case 3: return bs2g_dispatch<3>(s, jobs, items, nitems, Y, blocks, lds_bytes);
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
}

int launch_bconv(hipStream_t s, int d, int zdirect, const void *args_v, const int *prefix, int njobs, int blocks) {
  if (njobs <= 0) return SGTSNE_OK;
  const PassArgs *args = static_cast<const PassArgs *>(args_v);
  int kinds[kMaxPasses];
  const int np = pass_kinds(d, kinds);
  const int pairs = make_plan(d, zdirect != 0).pairs;
  for (int p = 0; p < np; ++p) {
    const int pr = (kinds[p] == PK_KGEN || kinds[p] == PK_KT_T || (kinds[p] == PK_W_W_CONV && d == 3)) ? 1 : pairs;
    switch (kinds[p]) {
      case PK_KGEN: SGT_TRY((bpass<IN_KGEN, OUT_KT, MID_FWD, true>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_KT_T: SGT_TRY((bpass<IN_KT, OUT_KT, MID_FWD, false>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_V_W: SGT_TRY((bpass<IN_VFIX, OUT_W, MID_FWD, true>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_W_W_CONV: SGT_TRY((bpass<IN_W, OUT_W, MID_CONV, false>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_W_PHI: SGT_TRY((bpass<IN_W, OUT_PHIQ, MID_INV, true>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_V_PHI: SGT_TRY((bpass<IN_VFIX, OUT_PHIQ, MID_CONV, true>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_W_W_FWD: SGT_TRY((bpass<IN_W, OUT_W, MID_FWD, false>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      case PK_W_W_INV: SGT_TRY((bpass<IN_W, OUT_W, MID_INV, false>(s, args, prefix, njobs, np, p, pr, blocks))); break;
      default: SGT_CHECK(false, SGTSNE_ERR_ARG, "bad FFT pass kind");
    }
  }
  return SGTSNE_OK;
}

int launch_bg2s(hipStream_t s, int d, int zdirect, const BatchJob *jobs, const BatchItem *items, int nitems,
                const float *Y, float *Frep, double *zpart, int blocks) {
  if (nitems <= 0) return SGTSNE_OK;
#define BG2S_CASE(DD)                                                                                     \
  case DD:                                                                                                \
    if (zdirect) bg2s_kernel<DD, true><<<blocks, g2s_threads<DD>(), 0, s>>>(jobs, items, nitems, Y, Frep, zpart); \
    else bg2s_kernel<DD, false><<<blocks, g2s_threads<DD>(), 0, s>>>(jobs, items, nitems, Y, Frep, zpart);        \
    break;
  switch (d) {
    BG2S_CASE(1)
    BG2S_CASE(2)
    BG2S_CASE(3)
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
#undef BG2S_CASE
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_bupdate(hipStream_t s, int d, BatchJob *jobs, const BatchItem *items, int nitems, float *Y, float *uY,
                   float *gains, const float *Fattr, const float *Frep, const double *zpart, double *spart,
                   int blocks) {
  if (nitems <= 0) return SGTSNE_OK;
  switch (d) {
    case 1: bupdate_kernel<1><<<blocks, kBT, 0, s>>>(jobs, items, nitems, Y, uY, gains, Fattr, Frep, zpart, spart); break;
    case 2: bupdate_kernel<2><<<blocks, kBT, 0, s>>>(jobs, items, nitems, Y, uY, gains, Fattr, Frep, zpart, spart); break;
    case 3: bupdate_kernel<3><<<blocks, kBT, 0, s>>>(jobs, items, nitems, Y, uY, gains, Fattr, Frep, zpart, spart); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_bcenter(hipStream_t s, int d, const BatchJob *jobs, const BatchItem *items, int nitems, float *Y,
                   const double *spart, double *mpart, int blocks) {
  if (nitems <= 0) return SGTSNE_OK;
  switch (d) {
    case 1: bcenter_kernel<1><<<blocks, kBT, 0, s>>>(jobs, items, nitems, Y, spart, mpart); break;
    case 2: bcenter_kernel<2><<<blocks, kBT, 0, s>>>(jobs, items, nitems, Y, spart, mpart); break;
    case 3: bcenter_kernel<3><<<blocks, kBT, 0, s>>>(jobs, items, nitems, Y, spart, mpart); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_bplan(hipStream_t s, int d, BatchJob *jobs, int njobs, const double *mpart, void *args_v, int *prefix,
                 const float2 *tw) {
  if (njobs <= 0) return SGTSNE_OK;
  SGT_CHECK(njobs <= 1024, SGTSNE_ERR_ARG, "at most 1024 job slots");
  PassArgs *args = static_cast<PassArgs *>(args_v);
  const int np = batch_npasses(d);
  bplan_kernel<<<div_up(njobs, 4), 256, 0, s>>>(d, jobs, njobs, mpart, args, np, tw);
  SGT_LAUNCH_CHECK();
  bprefix_kernel<<<1, 256, 0, s>>>(jobs, njobs, args, np, prefix);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_bmorton(hipStream_t s, int d, const BatchJob *jobs, int job, int off, int n, const float *Y,
                   uint32_t *keys, int32_t *vals) {
  if (n <= 0) return SGTSNE_OK;
  bmorton_kernel<<<div_up(n, 256), 256, 0, s>>>(d, jobs, job, off, n, Y, keys, vals);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_bzeta(hipStream_t s, const BatchJob *jobs, int job, const double *zpart, double *out) {
  bzeta_kernel<<<1, 64, 0, s>>>(jobs, job, zpart, out);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_knn_csr(hipStream_t s, int n, int k, const int32_t *I, int32_t *rowptr, int32_t *col) {
  knn_csr_kernel<<<div_up((int64_t)n * k + 1, 256), 256, 0, s>>>(n, k, I, rowptr, col);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_gather_i32(hipStream_t s, int n, const int32_t *order, const int32_t *in, int32_t *out) {
  if (n <= 0) return SGTSNE_OK;
  gather_i32_kernel<<<div_up(n, 256), 256, 0, s>>>(n, order, in, out);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_invert(hipStream_t s, int n, const int32_t *order, int32_t *inv) {
  if (n <= 0) return SGTSNE_OK;
  invert_kernel<<<div_up(n, 256), 256, 0, s>>>(n, order, inv);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_iota(hipStream_t s, int n, int32_t *out) {
  if (n <= 0) return SGTSNE_OK;
  iota_kernel<<<div_up(n, 256), 256, 0, s>>>(n, out);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_to_double(hipStream_t s, const float *in, double *out, int64_t m) {
  if (m <= 0) return SGTSNE_OK;
  to_double_kernel_b<<<div_up(m, 256), 256, 0, s>>>(in, out, m);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_offset_copy_i32(hipStream_t s, const int32_t *in, int32_t *out, int64_t count, int32_t delta) {
  if (count <= 0) return SGTSNE_OK;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 8192);
  offset_copy_kernel<<<blocks, 256, 0, s>>>(in, out, count, delta);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

}  // namespace sgt
