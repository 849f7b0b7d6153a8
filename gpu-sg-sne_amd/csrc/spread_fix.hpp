// spread_fix.hpp -- the spread of a run of points onto a 64-bit fixed-point
// grid through an LDS window of their stencils (HY/gridding.cpp:164-211 for
// the weights and channels), shared by the batched engine and the single-job
// engine's large grids.
//
// The grid is spread as integers q * w * 2^e (e chosen per job and iteration
// so that no cell can overflow), so the result is exact and does not depend
// on the order of the additions.  A run of consecutive points (in a locality
// order) usually covers a small window of the grid: the window is
// accumulated in LDS and added to the grid with one global atomic per
// non-zero cell; runs whose window does not fit add directly.
#pragma once

#include "common.hpp"
#include "grid_device.hpp"

#include <climits>

namespace sgt {

// x * 2^e rounded to the nearest integer (ties to even), for the spread's
// fixed-point scale 2^e (1 <= e <= 60) and |x * 2^e| < 2^62: the product is
// exact in FP32, so this is the integer llrint((double)x * 2^e) gives,
// without FP64 arithmetic.
__device__ __forceinline__ unsigned long long fix_round(float x, float scale) {
  return (unsigned long long)llrintf(x * scale);
}
// threads of a spread work-group: in 3-D a work unit (at most kFixUnit3
// points) is split over 8 waves, each with one batch of at most 64 points
constexpr int kFixUnit3 = 256;
template <int D>
constexpr int s2g_threads() { return D == 3 ? 512 : 256; }
static_assert((kFixUnit3 + s2g_threads<3>() / 64 - 1) / (s2g_threads<3>() / 64) <= 64,
              "a 3-D spread wave takes one batch of at most 64 points");

// min and max over the wave's 64 lanes
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Spreads points [begin, end) (3-D: at most kFixUnit3 of them) onto the
// fixed-point grid V of g with scale fsf = 2^e.  Every thread of the
// work-group calls it; win holds lds_slots 64-bit cells, sbox 6 ints.
template <int D>
__device__ __forceinline__ void spread_fix_unit(const GridParams &g, float fsf, unsigned long long *V, int begin,
                                                int end, const float *__restrict__ Y, unsigned long long *win,
                                                int lds_slots, int *sbox) {
  constexpr int M = D + 1;
  constexpr int NODES = NodeLanes<D>::NODES;
  constexpr int PPW = NodeLanes<D>::PPW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int node = lane % NODES, slot = lane / NODES;
  const int a0 = node & 3, a1 = (node >> 2) & 3, a2 = node >> 4;
  const int ng = g.ng;
  const int G = grid_cells<D>(ng);
  // each point slot of each wave takes a contiguous run of the work unit's
  // points; a lane adds its points' integer contributions in registers while
  // they land on the same node (consecutive points mostly share a cell) and
  // issues one atomic per run.  Integer sums are exact, so the grid is the
  // same bits as with one atomic per point.  A slot's points arrive NODES at
  // a time: lane `node` loads point b + node of the run (one coalesced load
  // for the slot) and every lane of the slot reads them in order from the
  // loading lanes, so no lane waits on a global load per point.  In 3-D a
  // wave's run is one such batch, loaded once for the window and the spread.
  const int len = end - begin;
  const int per_wave = (len + nwaves - 1) / nwaves, per_slot = (per_wave + PPW - 1) / PPW;
  const int r0 = min(end, begin + wave * per_wave + slot * per_slot);
  const int r1 = min(min(end, begin + (wave + 1) * per_wave), r0 + per_slot);
  const int cnt = r1 - r0;  // this slot's points (<= 0: none)
  float y3[D];
#pragma unroll
  for (int a = 0; a < D; ++a) y3[a] = 0.0f;
  if (D == 3 && node < cnt) load_pt<D>(Y, r0 + node, y3);
  if (threadIdx.x < 3) {
    sbox[threadIdx.x] = INT_MAX;
    sbox[3 + threadIdx.x] = INT_MIN;
  }
  __syncthreads();
  {
    int lo[D], hi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) { lo[a] = INT_MAX; hi[a] = INT_MIN; }
    auto widen = [&](const float (&y)[D]) {
      int f[D];
      float w[D][4];
      stencil<D>(g, y, f, w);
#pragma unroll
      for (int a = 0; a < D; ++a) { lo[a] = min(lo[a], f[a]); hi[a] = max(hi[a], f[a]); }
    };
    if constexpr (D == 3) {
      if (node < cnt) widen(y3);
    } else {
      for (int i = begin + threadIdx.x; i < end; i += blockDim.x) {
        float y[D];
        load_pt<D>(Y, i, y);
        widen(y);
      }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
      lo[a] = wave_min_i(lo[a]);
      hi[a] = wave_max_i(hi[a]);
    }
    if (lane == 0) {
#pragma unroll
      for (int a = 0; a < D; ++a) {
        atomicMin(&sbox[a], lo[a]);
        atomicMax(&sbox[3 + a], hi[a]);
      }
    }
  }
  __syncthreads();
  int wlo[3] = {0, 0, 0}, wd[3] = {1, 1, 1};
  int wcells = 1;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    wlo[a] = sbox[a];
    wd[a] = sbox[3 + a] - sbox[a] + 4;
    wcells *= wd[a];
  }
  const bool in_lds = (int64_t)wcells * M <= (int64_t)lds_slots;
  __syncthreads();  // sbox is reset by the next item
  auto for_points = [&](auto &&visit) {
    if constexpr (D == 3) {  // one slot per wave: the source lane is uniform
      for (int k = 0; k < cnt; ++k) {
        float y[D];
#pragma unroll
        for (int a = 0; a < D; ++a) y[a] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(y3[a]), k));
        visit(y);
      }
    } else {
      for (int b = 0; b < per_slot; b += NODES) {
        float yv[D];
        if (b + node < cnt) {
          load_pt<D>(Y, r0 + b + node, yv);
        } else {
#pragma unroll
          for (int a = 0; a < D; ++a) yv[a] = 0.0f;
        }
        const int kn = min(NODES, per_slot - b);
        for (int k = 0; k < kn; ++k) {
          float y[D];
#pragma unroll
          for (int a = 0; a < D; ++a) y[a] = __shfl(yv[a], slot * NODES + k, 64);
          if (b + k < cnt) visit(y);
        }
      }
    }
  };
  unsigned long long acc[M];
  int cur = -1;
  if (in_lds) {
    for (int e = threadIdx.x; e < wcells * M; e += blockDim.x) win[e] = 0ull;
    __syncthreads();
    for_points([&](const float (&y)[D]) {
      int f[D];
      float ww;
      lane_node<D>(g, y, a0, a1, a2, f, ww);
      int loc;
      if constexpr (D == 1) {
        loc = f[0] + a0 - wlo[0];
      } else if constexpr (D == 2) {
        loc = (f[1] + a1 - wlo[1]) * wd[0] + f[0] + a0 - wlo[0];
      } else {
        loc = ((f[2] + a2 - wlo[2]) * wd[1] + f[1] + a1 - wlo[1]) * wd[0] + f[0] + a0 - wlo[0];
      }
      if (loc != cur) {
        if (cur >= 0) {
#pragma unroll
          for (int m = 0; m < M; ++m) atomicAdd(&win[m * wcells + cur], acc[m]);
        }
        cur = loc;
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] = 0ull;
      }
      acc[0] += fix_round(ww, fsf);
#pragma unroll
      for (int a = 0; a < D; ++a) acc[a + 1] += fix_round((y[a] - g.center[a]) * ww, fsf);
    });
    if (cur >= 0) {
#pragma unroll
      for (int m = 0; m < M; ++m) atomicAdd(&win[m * wcells + cur], acc[m]);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < wcells; e += blockDim.x) {
      int gi;
      if constexpr (D == 1) {
        gi = wlo[0] + e;
      } else if constexpr (D == 2) {
        const int e1 = e / wd[0], e0 = e - e1 * wd[0];
        gi = (wlo[1] + e1) * ng + wlo[0] + e0;
      } else {
        const int e2 = e / (wd[0] * wd[1]);
        const int r = e - e2 * wd[0] * wd[1];
        const int e1 = r / wd[0], e0 = r - e1 * wd[0];
        gi = ((wlo[2] + e2) * ng + wlo[1] + e1) * ng + wlo[0] + e0;
      }
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const unsigned long long v = win[m * wcells + e];
        if (v != 0ull) atomicAdd(&V[(int64_t)m * G + gi], v);
      }
    }
    __syncthreads();  // the window is cleared by the next item
  } else {
    for_points([&](const float (&y)[D]) {
      int gi;
      float ww;
      node_of_lane<D>(g, y, node, gi, ww);
      if (gi != cur) {
        if (cur >= 0) {
#pragma unroll
          for (int m = 0; m < M; ++m) atomicAdd(&V[(int64_t)m * G + cur], acc[m]);
        }
        cur = gi;
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] = 0ull;
      }
      acc[0] += fix_round(ww, fsf);
#pragma unroll
      for (int a = 0; a < D; ++a) acc[a + 1] += fix_round((y[a] - g.center[a]) * ww, fsf);
    });
    if (cur >= 0) {
#pragma unroll
      for (int m = 0; m < M; ++m) atomicAdd(&V[(int64_t)m * G + cur], acc[m]);
    }
  }
}

}  // namespace sgt
