// grid_device.hpp -- the grid mapping and Lagrange stencil shared by the
// single-job kernels (kernels_grid.hip) and the batched engine
// (kernels_batch.hip).
#pragma once

#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"
#include "fft_device.hpp"

namespace sgt {

// getBestGridSize (HY/qq.cpp:108-130, CU/utils.cu:194-215): the FFT-friendly
// sizes the interpolation grid side (nGrid + 2) is rounded up to.
#define SGT_GRID_SIZE_LIST                                                                    \
  8, 9, 10, 11, 12, 13, 14, 15, 16, 20, 25, 26, 28, 32, 33, 35, 36, 39, 40, 42, 44, 45, 48, 49, 50, \
      52, 54, 55, 56, 60, 63, 64, 65, 66, 70, 72, 75, 77, 78, 80, 84, 88, 90, 91, 96, 98, 99, 100,  \
      104, 105, 108, 110, 112, 117, 120, 125, 126, 130, 132, 135, 140, 144, 147, 150, 154, 156,     \
      160, 165, 168, 175, 176, 180, 182, 189, 192, 195, 196, 198, 200, 208, 210, 216, 220, 224,     \
      225, 231, 234, 240, 245, 250, 252, 260, 264, 270, 273, 275, 280, 288, 294, 297, 300, 308,     \
      312, 315, 320, 325, 330, 336, 343, 350, 351, 352, 360, 364, 375, 378, 385, 390, 392, 396,     \
      400, 416, 420, 432, 440, 441, 448, 450, 455, 462, 468, 480, 490, 495, 500, 504, 512

__host__ __device__ inline int best_grid_size_in(const int *sizes, int count, int nGrid) {
  for (int i = 0; i < count; i++)
    if (nGrid + 2 <= sizes[i]) return sizes[i] - 2;
  return sizes[count - 1] - 2;
}

// Grid mapping of one iteration from the bounding box of the centred Y
// (HY/qq.cpp:147-166, HY/nuconv.cpp:28-47; FP32 clamp as CU/nuconv.cu:9-35).
__host__ __device__ inline void grid_from_box(int d, int nGrid, const double *minv, const double *maxv, double maxy,
                                              GridParams &g) {
  g.d = d;
  g.nGrid = nGrid;
  g.ng = nGrid + 2;
  g.L = fft::fft_length(2 * g.ng - 1);
  for (int a = 0; a < 3; a++) {
    g.minv[a] = a < d ? (float)minv[a] : 0.0f;
    g.center[a] = a < d ? (float)(0.5 * (minv[a] + maxv[a])) : 0.0f;
  }
  g.scale = (float)((double)(nGrid - 1) / maxy);
  g.gmax = nextafterf((float)(nGrid - 1), 0.0f);
  g.hgrid = maxy / ((double)(nGrid - 1) - DBL_EPSILON);
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

// Visits the 4^D stencil nodes of one point: fn(linear node index, weight).
template <int D, class Fn>
__device__ __forceinline__ void for_stencil(int ng, const int (&f)[D], const float (&w)[D][4], Fn fn) {
  if constexpr (D == 1) {
#pragma unroll
    for (int i0 = 0; i0 < 4; ++i0) fn(f[0] + i0, w[0][i0]);
  } else if constexpr (D == 2) {
#pragma unroll
    for (int i1 = 0; i1 < 4; ++i1)
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) fn((f[1] + i1) * ng + f[0] + i0, w[1][i1] * w[0][i0]);
  } else {
#pragma unroll
    for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
      for (int i1 = 0; i1 < 4; ++i1) {
        const float w21 = w[2][i2] * w[1][i1];
#pragma unroll
        for (int i0 = 0; i0 < 4; ++i0) fn(((f[2] + i2) * ng + f[1] + i1) * ng + f[0] + i0, w21 * w[0][i0]);
      }
  }
}

template <int D>
__device__ __forceinline__ int grid_cells(int ng) {
  return D == 1 ? ng : D == 2 ? ng * ng : ng * ng * ng;
}

__device__ __forceinline__ float sgnf(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }

template <int D>
struct NodeLanes {
  static constexpr int NODES = D == 1 ? 4 : D == 2 ? 16 : 64;
  static constexpr int PPW = 64 / NODES;  // points per wave per step
};

// The base cell f of a point's stencil and the weight of its node
// (a0, a1, a2), evaluating only that node's Lagrange factor along each axis:
// the same values stencil<D> gives, with half the polynomial work.
template <int D>
__device__ __forceinline__ void lane_node(const GridParams &g, const float (&y)[D], int a0, int a1, int a2,
                                          int (&f)[D], float &weight) {
  const int an[3] = {a0, a1, a2};
  float wa[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float t = (y[a] - g.minv[a]) * g.scale;
    t = fminf(fmaxf(t, 0.0f), g.gmax);  // NaN -> 0: never leaves the grid
    const int fi = (int)t;
    const float dl = t - (float)fi;
    f[a] = fi;
    const int k = an[a];
    const float x = k == 0 ? 1.0f + dl : k == 1 ? dl : k == 2 ? 1.0f - dl : 2.0f - dl;
    const float l1 = lg1(x), l2 = lg2(x);
    wa[a] = (k == 1 || k == 2) ? l1 : l2;
  }
  if constexpr (D == 1) weight = wa[0];
  else if constexpr (D == 2) weight = wa[1] * wa[0];
  else weight = wa[2] * wa[1] * wa[0];
}

template <int D>
__device__ __forceinline__ void node_of_lane(const GridParams &g, const float (&y)[D], int node, int &idx,
                                             float &weight) {
  int f[D];
  const int a0 = node & 3, a1 = (node >> 2) & 3, a2 = node >> 4;
  lane_node<D>(g, y, a0, a1, a2, f, weight);
  const int ng = g.ng;
  if constexpr (D == 1) {
    idx = f[0] + a0;
  } else if constexpr (D == 2) {
    idx = (f[1] + a1) * ng + f[0] + a0;
  } else {
    idx = ((f[2] + a2) * ng + f[1] + a1) * ng + f[0] + a0;
  }
}


}  // namespace sgt
