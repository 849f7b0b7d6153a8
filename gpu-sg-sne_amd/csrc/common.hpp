// common.hpp -- shared definitions of the gfx950 SG-t-SNE engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace sgt {

// Error propagation: every host routine returns an int status (SGTSNE_*);
// the message is kept per thread for sgtsne_last_error().
void set_error(const char *fmt, ...);

#define SGT_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::sgt::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,              \
                       hipGetErrorString(e_));                                         \
      return SGTSNE_ERR_HIP;                                                           \
    }                                                                                  \
  } while (0)

#define SGT_CHECK(cond, code, ...)                                                     \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      ::sgt::set_error(__VA_ARGS__);                                                   \
      return (code);                                                                   \
    }                                                                                  \
  } while (0)

#define SGT_TRY(call)                                                                  \
  do {                                                                                 \
    int s_ = (call);                                                                   \
    if (s_ != 0) return s_;                                                            \
  } while (0)

// Launch-time check: the launch itself (bad configuration) is reported at once.
#define SGT_LAUNCH_CHECK() SGT_HIP(hipGetLastError())

constexpr int kWave = 64;  // CDNA wavefront width

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Per-point storage stride: Y, uY, gains, Fattr, Frep are AoS with d=1 -> 1,
// d=2 -> 2 (float2), d=3 -> 4 (float4, the 4th lane is padding) so that one
// point is one aligned vector load.
__host__ __device__ inline int point_stride(int d) { return d == 3 ? 4 : d; }

// ---- device helpers -------------------------------------------------------

template <int D>
struct VecT;
template <>
struct VecT<1> { using T = float; };
template <>
struct VecT<2> { using T = float2; };
template <>
struct VecT<3> { using T = float4; };

template <int D>
__device__ __forceinline__ void load_pt(const float *__restrict__ Y, int i, float (&y)[D]) {
  if constexpr (D == 1) {
    y[0] = Y[i];
  } else if constexpr (D == 2) {
    float2 v = reinterpret_cast<const float2 *>(Y)[i];
    y[0] = v.x; y[1] = v.y;
  } else {
    float4 v = reinterpret_cast<const float4 *>(Y)[i];
    y[0] = v.x; y[1] = v.y; y[2] = v.z;
  }
}

template <int D>
__device__ __forceinline__ void store_pt(float *__restrict__ Y, int i, const float (&y)[D]) {
  if constexpr (D == 1) {
    Y[i] = y[0];
  } else if constexpr (D == 2) {
    reinterpret_cast<float2 *>(Y)[i] = make_float2(y[0], y[1]);
  } else {
    reinterpret_cast<float4 *>(Y)[i] = make_float4(y[0], y[1], y[2], 0.f);
  }
}

// Wavefront reductions (64 lanes) through cross-lane shuffles.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Neumaier (compensated) accumulation: s + c carries the sum.
struct Neu {
  double s, c;
};
__device__ __forceinline__ void neu_add(Neu &a, double x) {
  double t = a.s + x;
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
}
__device__ __forceinline__ void neu_merge(Neu &a, const Neu &b) {
  neu_add(a, b.s);
  a.c += b.c;
}
__device__ __forceinline__ Neu wave_neu(Neu a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    Neu b;
    b.s = __shfl_xor(a.s, o, 64);
    b.c = __shfl_xor(a.c, o, 64);
    neu_merge(a, b);
  }
  return a;
}

}  // namespace sgt
