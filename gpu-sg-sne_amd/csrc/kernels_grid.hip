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
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/sgtsne.h"

namespace sgt {

namespace {

constexpr int kBlock = 256;

// ---- partial reduction helpers -------------------------------------------
// Each block writes `cnt` doubles to rw.part[block*8 ...]; the block that
// takes the last ticket sees all partials (release/acquire through the
// ticket counter at agent scope) and finalises.
__device__ __forceinline__ void store_partial(double *slot, double v) {
  __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_partial(const double *slot) {
  return __hip_atomic_load(const_cast<double *>(slot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ bool take_last_ticket(unsigned *ticket) {
  __shared__ int am_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (am_last) __threadfence();
  return am_last != 0;
}

__device__ __forceinline__ void reset_ticket(unsigned *ticket) {
  if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int K>
__device__ void block_reduce_sum(double (&v)[K], double *lds /* [K*4] */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * 4 + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = lds[k * 4 + 0] + lds[k * 4 + 1] + lds[k * 4 + 2] + lds[k * 4 + 3];
  __syncthreads();
}

// Lagrange cubic weights for the 4 stencil nodes f..f+3 of a point at
// f + 1 + delta (HY/gridding.cpp:28-40: g1 at distance < 1, g2 at 1..2).
__device__ __forceinline__ float lg1(float t) { return ((0.5f * t - 1.0f) * t - 0.5f) * t + 1.0f; }
__device__ __forceinline__ float lg2(float t) {
  return ((-(1.0f / 6.0f) * t + 1.0f) * t - (11.0f / 6.0f)) * t + 1.0f;
}

template <int D>
__device__ __forceinline__ void stencil(const GridParams &g, const float (&y)[D], int (&f)[D],
                                        float (&w)[D][4]) {
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float t = (y[a] - g.minv[a]) * g.scale;
    t = fminf(fmaxf(t, 0.0f), g.gmax);  // NaN -> 0: never leaves the grid
    int fi = (int)t;
    float dl = t - (float)fi;
    f[a] = fi;
    w[a][0] = lg2(1.0f + dl);
    w[a][1] = lg1(dl);
    w[a][2] = lg1(1.0f - dl);
    w[a][3] = lg2(2.0f - dl);
  }
}

// ---- fill ------------------------------------------------------------------
__global__ void fill_kernel(float *__restrict__ p, int64_t count, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// ---- centring + min/max ------------------------------------------------------
// Y -= mean (the mean of the last update, HY/gradient_descendHYB.cpp:64-76),
// then min/max per dimension of the centred Y (HY/qq.cpp:147-163).
template <int D>
__global__ __launch_bounds__(kBlock) void center_minmax_kernel(float *__restrict__ Y, int n,
                                                               int apply_mean, RedWork rw,
                                                               DevScalars *sc) {
  __shared__ double lds[2 * D * 4 + 4];
  float m[D];
#pragma unroll
  for (int a = 0; a < D; ++a) m[a] = apply_mean ? (float)sc->mean[a] : 0.0f;
  float mn[D], mx[D];
  float bad = 0.0f;  // count of non-finite coordinates
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
  const double vbad = wave_sum((double)bad);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double vmin[D], vmax[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    vmin[a] = wave_min((double)mn[a]);
    vmax[a] = wave_max((double)mx[a]);
  }
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
    store_partial(&rw.part[blockIdx.x * 8 + 3],
                  lds[2 * D * 4] + lds[2 * D * 4 + 1] + lds[2 * D * 4 + 2] + lds[2 * D * 4 + 3]);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double lo = fmin(fmin(lds[a * 4], lds[a * 4 + 1]), fmin(lds[a * 4 + 2], lds[a * 4 + 3]));
      double hi = fmax(fmax(lds[(D + a) * 4], lds[(D + a) * 4 + 1]),
                       fmax(lds[(D + a) * 4 + 2], lds[(D + a) * 4 + 3]));
      store_partial(&rw.part[blockIdx.x * 8 + a], lo);
      store_partial(&rw.part[blockIdx.x * 8 + 4 + a], hi);
    }
  }
  if (take_last_ticket(rw.ticket)) {
    if (threadIdx.x == 0) {
      double nbad = 0.0;
      for (int b = 0; b < (int)gridDim.x; ++b) nbad += load_partial(&rw.part[b * 8 + 3]);
      for (int a = 0; a < D; ++a) {
        double lo = INFINITY, hi = -INFINITY;
        for (int b = 0; b < (int)gridDim.x; ++b) {
          lo = fmin(lo, load_partial(&rw.part[b * 8 + a]));
          hi = fmax(hi, load_partial(&rw.part[b * 8 + 4 + a]));
        }
        // a non-finite coordinate poisons the extent: the host rejects it
        sc->minv[a] = nbad > 0.0 ? NAN : lo;
        sc->maxv[a] = hi;
        if (apply_mean) sc->mean[a] = 0.0;
      }
    }
    reset_ticket(rw.ticket);
  }
}

// ---- S2G with global float atomics (baseline form) ----------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void s2g_atomic_kernel(GridParams g, int n,
                                                            const float *__restrict__ Y,
                                                            float *__restrict__ V) {
  const int ng = g.ng;
  const int G = D == 1 ? ng : D == 2 ? ng * ng : ng * ng * ng;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float y[D];
    load_pt<D>(Y, i, y);
    int f[D];
    float w[D][4];
    stencil<D>(g, y, f, w);
    float q[D + 1];
    q[0] = 1.0f;
#pragma unroll
    for (int a = 0; a < D; ++a) q[a + 1] = y[a] - g.center[a];
    if constexpr (D == 1) {
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) {
        int idx = f[0] + i0;
#pragma unroll
        for (int c = 0; c <= D; ++c) unsafeAtomicAdd(&V[c * G + idx], q[c] * w[0][i0]);
      }
    } else if constexpr (D == 2) {
#pragma unroll
      for (int i1 = 0; i1 < 4; ++i1)
#pragma unroll
        for (int i0 = 0; i0 < 4; ++i0) {
          float ww = w[1][i1] * w[0][i0];
          int idx = (f[1] + i1) * ng + f[0] + i0;
#pragma unroll
          for (int c = 0; c <= D; ++c) unsafeAtomicAdd(&V[c * G + idx], q[c] * ww);
        }
    } else {
#pragma unroll
      for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
        for (int i1 = 0; i1 < 4; ++i1)
#pragma unroll
          for (int i0 = 0; i0 < 4; ++i0) {
            float ww = w[2][i2] * w[1][i1] * w[0][i0];
            int idx = ((f[2] + i2) * ng + f[1] + i1) * ng + f[0] + i0;
#pragma unroll
            for (int c = 0; c <= D; ++c) unsafeAtomicAdd(&V[c * G + idx], q[c] * ww);
          }
    }
  }
}

// ---- G2S + forces + Z ------------------------------------------------------------
// Phi_c(y_i) by the same stencil; Frep_i (unnormalised) = ~y_i Phi0 - Phi_{1..d};
// z_i = (1 + 2|~y_i|^2) Phi0 - 2 ~y_i . Phi_{1..d} in FP64 and a compensated
// sum; the last block writes Z = sum z_i - N (HY/qq.cpp:75-106).  ~y is y minus
// the charge shift used by S2G (any shift leaves Frep and sum z_i unchanged).
template <int D>
__global__ __launch_bounds__(kBlock) void g2s_force_kernel(GridParams g, int n,
                                                           const float *__restrict__ Y,
                                                           const float *__restrict__ Phi,
                                                           float *__restrict__ Frep, RedWork rw,
                                                           DevScalars *sc) {
  __shared__ double lds[8];
  const int ng = g.ng;
  const int G = D == 1 ? ng : D == 2 ? ng * ng : ng * ng * ng;
  Neu z{0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float y[D];
    load_pt<D>(Y, i, y);
    int f[D];
    float w[D][4];
    stencil<D>(g, y, f, w);
    float ph[D + 1];
#pragma unroll
    for (int c = 0; c <= D; ++c) ph[c] = 0.0f;
    if constexpr (D == 1) {
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) {
        int idx = f[0] + i0;
#pragma unroll
        for (int c = 0; c <= D; ++c) ph[c] += Phi[c * G + idx] * w[0][i0];
      }
    } else if constexpr (D == 2) {
#pragma unroll
      for (int i1 = 0; i1 < 4; ++i1)
#pragma unroll
        for (int i0 = 0; i0 < 4; ++i0) {
          float ww = w[1][i1] * w[0][i0];
          int idx = (f[1] + i1) * ng + f[0] + i0;
#pragma unroll
          for (int c = 0; c <= D; ++c) ph[c] += Phi[c * G + idx] * ww;
        }
    } else {
#pragma unroll
      for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
        for (int i1 = 0; i1 < 4; ++i1) {
          float w21 = w[2][i2] * w[1][i1];
#pragma unroll
          for (int i0 = 0; i0 < 4; ++i0) {
            float ww = w21 * w[0][i0];
            int idx = ((f[2] + i2) * ng + f[1] + i1) * ng + f[0] + i0;
#pragma unroll
            for (int c = 0; c <= D; ++c) ph[c] += Phi[c * G + idx] * ww;
          }
        }
    }
    float fr[D];
    double ysq = 0.0, ydot = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float yt = y[a] - g.center[a];
      fr[a] = yt * ph[0] - ph[a + 1];
      ysq += (double)yt * (double)yt;
      ydot += (double)yt * (double)ph[a + 1];
    }
    store_pt<D>(Frep, i, fr);
    neu_add(z, (1.0 + 2.0 * ysq) * (double)ph[0] - 2.0 * ydot);
  }
  z = wave_neu(z);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { lds[wid * 2] = z.s; lds[wid * 2 + 1] = z.c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    Neu t{lds[0], lds[1]};
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) neu_merge(t, Neu{lds[2 * k], lds[2 * k + 1]});
    store_partial(&rw.part[blockIdx.x * 8 + 0], t.s);
    store_partial(&rw.part[blockIdx.x * 8 + 1], t.c);
  }
  if (take_last_ticket(rw.ticket)) {
    if (threadIdx.x == 0) {
      Neu t{0.0, 0.0};
      for (int b = 0; b < (int)gridDim.x; ++b)
        neu_merge(t, Neu{load_partial(&rw.part[b * 8]), load_partial(&rw.part[b * 8 + 1])});
      sc->Z = (t.s + t.c) - (double)n;
    }
    reset_ticket(rw.ticket);
  }
}

// ---- fused gradient + gains + momentum update ----------------------------------
// dy = alpha Fattr - Frep / Z; gains, uy, y exactly as CU/gradient_descend.cu:29-44;
// partial sums of the new Y feed the mean of the next centring.
__device__ __forceinline__ float sgnf(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }

template <int D>
__global__ __launch_bounds__(kBlock) void update_kernel(int n, float *__restrict__ Y,
                                                        float *__restrict__ uY,
                                                        float *__restrict__ gains,
                                                        const float *__restrict__ Fattr,
                                                        const float *__restrict__ Frep,
                                                        float alpha, float momentum, float eta,
                                                        RedWork rw, DevScalars *sc) {
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
      float dy = alpha * fa[a] - fr[a] * zinv;
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
  block_reduce_sum<D>(s, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int a = 0; a < D; ++a) store_partial(&rw.part[blockIdx.x * 8 + a], s[a]);
  if (take_last_ticket(rw.ticket)) {
    if (threadIdx.x == 0) {
      for (int a = 0; a < D; ++a) {
        double t = 0.0;
        for (int b = 0; b < (int)gridDim.x; ++b) t += load_partial(&rw.part[b * 8 + a]);
        sc->sum[a] = t;
        sc->mean[a] = t / (double)n;
      }
    }
    reset_ticket(rw.ticket);
  }
}

// ---- exact O(N^2) repulsion (validation, CU/compute_error.cu:8-53) -------------
template <int D>
__global__ __launch_bounds__(kBlock) void rep_exact_kernel(int n, const float *__restrict__ Y,
                                                           float *__restrict__ Frep, RedWork rw,
                                                           DevScalars *sc) {
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
    int j = j0 + threadIdx.x;
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
      double q = 1.0 / (1.0 + dist);
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
  double zz[1] = {z};
  block_reduce_sum<1>(zz, lds);
  if (threadIdx.x == 0) store_partial(&rw.part[blockIdx.x * 8], zz[0]);
  if (take_last_ticket(rw.ticket)) {
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int b = 0; b < (int)gridDim.x; ++b) t += load_partial(&rw.part[b * 8]);
      sc->zexact = t;
    }
    reset_ticket(rw.ticket);
  }
}

int grid_blocks(int n, int max_blocks) {
  int b = div_up(n > 0 ? n : 1, kBlock);
  return b < max_blocks ? b : max_blocks;
}

}  // namespace

int launch_fill(hipStream_t s, float *p, int64_t count, float v) {
  if (count <= 0) return SGTSNE_OK;
  int blocks = (int)std::min<int64_t>((count + kBlock - 1) / kBlock, 4096);
  fill_kernel<<<blocks, kBlock, 0, s>>>(p, count, v);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_center_minmax(hipStream_t s, int d, int n, float *Y, int apply_mean, RedWork rw,
                         DevScalars *sc) {
  int blocks = grid_blocks(n, std::min(rw.max_blocks, 1024));
  switch (d) {
    case 1: center_minmax_kernel<1><<<blocks, kBlock, 0, s>>>(Y, n, apply_mean, rw, sc); break;
    case 2: center_minmax_kernel<2><<<blocks, kBlock, 0, s>>>(Y, n, apply_mean, rw, sc); break;
    case 3: center_minmax_kernel<3><<<blocks, kBlock, 0, s>>>(Y, n, apply_mean, rw, sc); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_s2g_atomic(hipStream_t s, const GridParams &g, int n, const float *Y, float *V) {
  int blocks = div_up(n > 0 ? n : 1, kBlock);
  switch (g.d) {
    case 1: s2g_atomic_kernel<1><<<blocks, kBlock, 0, s>>>(g, n, Y, V); break;
    case 2: s2g_atomic_kernel<2><<<blocks, kBlock, 0, s>>>(g, n, Y, V); break;
    case 3: s2g_atomic_kernel<3><<<blocks, kBlock, 0, s>>>(g, n, Y, V); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_g2s_force(hipStream_t s, const GridParams &g, int n, const float *Y, const float *Phi,
                     float *Frep, RedWork rw, DevScalars *sc) {
  int blocks = grid_blocks(n, rw.max_blocks);
  switch (g.d) {
    case 1: g2s_force_kernel<1><<<blocks, kBlock, 0, s>>>(g, n, Y, Phi, Frep, rw, sc); break;
    case 2: g2s_force_kernel<2><<<blocks, kBlock, 0, s>>>(g, n, Y, Phi, Frep, rw, sc); break;
    case 3: g2s_force_kernel<3><<<blocks, kBlock, 0, s>>>(g, n, Y, Phi, Frep, rw, sc); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_update(hipStream_t s, int d, int n, float *Y, float *uY, float *gains,
                  const float *Fattr, const float *Frep, float alpha, float momentum, float eta,
                  RedWork rw, DevScalars *sc) {
  int blocks = grid_blocks(n, rw.max_blocks);
  switch (d) {
    case 1: update_kernel<1><<<blocks, kBlock, 0, s>>>(n, Y, uY, gains, Fattr, Frep, alpha, momentum, eta, rw, sc); break;
    case 2: update_kernel<2><<<blocks, kBlock, 0, s>>>(n, Y, uY, gains, Fattr, Frep, alpha, momentum, eta, rw, sc); break;
    case 3: update_kernel<3><<<blocks, kBlock, 0, s>>>(n, Y, uY, gains, Fattr, Frep, alpha, momentum, eta, rw, sc); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int launch_rep_exact(hipStream_t s, int d, int n, const float *Y, float *Frep, RedWork rw,
                     DevScalars *sc) {
  int blocks = div_up(n > 0 ? n : 1, kBlock);
  SGT_CHECK(blocks <= rw.max_blocks, SGTSNE_ERR_ARG, "exact repulsion: n too large for workspace");
  switch (d) {
    case 1: rep_exact_kernel<1><<<blocks, kBlock, 0, s>>>(n, Y, Frep, rw, sc); break;
    case 2: rep_exact_kernel<2><<<blocks, kBlock, 0, s>>>(n, Y, Frep, rw, sc); break;
    case 3: rep_exact_kernel<3><<<blocks, kBlock, 0, s>>>(n, Y, Frep, rw, sc); break;
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "d must be 1..3");
  }
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

}  // namespace sgt
