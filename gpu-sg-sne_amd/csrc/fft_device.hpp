// fft_device.hpp -- the LDS FFT passes of the non-periodic grid
// convolution, shared by the single-job launcher (kernels_fft.hip) and the
// batched engine (kernels_batch.hip).
//
// The reference evaluates the same Toeplitz (non-periodic) convolution as a
// sum of 2^d circulant / skew-circulant terms of size ng
// (HY/non_periodic_conv.cpp:26-708, CU/non_periodic_convF.cu:162-517).  For
// an FFT length L >= 2 ng - 1 a single circulant of size L with the symmetric
// embedding K_circ[i] = K(min(i, L - i)) gives exactly the same product on the
// ng outputs.  Here L is the smallest even 2^a 3^b 5^c >= 2 ng - 1 and every
// line transform is a Stockham autosort FFT held in LDS, in radix-4/2/3/5
// stages (the formulation of Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC 2008).  K_circ is even in
// every axis, so its spectrum is real and even: only its (L/2+1)^d
// half-spectrum is kept.  Two
// real grid channels share one complex line (their results are the real and
// imaginary parts of the inverse transform).  The two kernels K2 = (1+h^2r^2)^-2
// and K1 = (1+h^2r^2)^-1 are transformed together as K2 + i K1 (both spectra
// are real), and the unit-charge channel V0 is transformed alone and multiplied
// by K2^ + i K1^, which returns K2*V0 + i K1*V0: the potential and the direct
// normalisation term from one line.
//
// Layouts: grid V/Phi [m][ng^d] (x fastest, HY/gridding.cpp:22-24);
// complex work W per channel pair: 2-D [ng][L] (rows y < ng, all kx),
// 3-D [ng][L][L] (z < ng).
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace sgt {
namespace fft {

constexpr int kFftThreads = 256;
constexpr int kLdsElems = 2048;  // complex elements per ping-pong buffer (16 KiB)
constexpr int kMaxL = 1024;      // longest line transform
constexpr int kMaxPasses = 8;
constexpr int kMaxStages = 10;
constexpr int kPerThread = kLdsElems / kFftThreads;  // elements of a work-group's share per thread
// two ping-pong buffers and the twiddle table of the line length
constexpr size_t kLdsBytes = (2 * (size_t)kLdsElems + kMaxL) * sizeof(float2);
// LDS of one pass's work-group (nl lines of length L held Lp elements apart)
__host__ __device__ inline size_t pass_lds_bytes(int nl, int L, int Lp) {
  return (2 * (size_t)nl * Lp + L) * sizeof(float2);
}

// Even lengths whose only prime factors are 2, 3 and 5.
__host__ __device__ inline bool fft_smooth(int L) {
  if (L < 2 || (L & 1)) return false;
  while (L % 2 == 0) L /= 2;
  while (L % 3 == 0) L /= 3;
  while (L % 5 == 0) L /= 5;
  return L == 1;
}

// Smallest supported FFT length >= n.
__host__ __device__ inline int fft_length(int n) {
  int L = n < 2 ? 2 : n + (n & 1);
  while (!fft_smooth(L)) L += 2;
  return L;
}

// Twiddle tables, one per supported length in increasing order, each holding
// exp(-2 pi i t / L) for t < L; the table of L starts at fft_table_offset(L).
__host__ __device__ inline int fft_table_offset(int L) {
  int off = 0;
  for (int m = 2; m < L; m += 2)
    if (fft_smooth(m)) off += m;
  return off;
}
__host__ __device__ inline int fft_table_size() { return fft_table_offset(kMaxL + 2); }

// q = x / d for 0 <= x < 2^22, through a float reciprocal and one correction.
__device__ __forceinline__ int fdiv(int x, int d, float inv) {
  int q = (int)((float)x * inv);
  const int r = x - q * d;
  if (r < 0) --q;
  else if (r >= d) ++q;
  return q;
}

// IN_VFIX: the spread grid in 64-bit fixed point (batched engine); the pass
// converts it and clears every element it reads, so the grid is zero again
// for the next spread.
enum InKind { IN_VPAIR = 0, IN_W = 1, IN_KGEN = 2, IN_KT = 3, IN_VFIX = 4 };
// OUT_PHIQ: potentials node-major for the batched gather, channels 0-3 of a
// node as one float4 (Phi[4 node + c]) and channel 4 as a plane after them
// (Phi[4 G + node]).
enum OutKind { OUT_W = 0, OUT_PHIPAIR = 1, OUT_KT = 2, OUT_PHIQ = 3 };
enum MidKind { MID_FWD = 0, MID_INV = 1, MID_CONV = 2 };

// The pass shapes of one convolution (template arguments of the pass kernel).
enum PassKind {
  PK_KGEN = 0,     // kernel values -> half spectrum along the first axis
  PK_KT_T = 1,     // half spectrum along a strided axis
  PK_V_W = 2,      // grid channels -> W, forward along x
  PK_W_W_CONV = 3, // forward, multiply by the kernel spectrum, inverse (strided)
  PK_W_PHI = 4,    // inverse along x -> Phi
  PK_V_PHI = 5,    // 1-D: forward, multiply, inverse, -> Phi
  PK_W_W_FWD = 6,  // forward along a strided axis
  PK_W_W_INV = 7,  // inverse along a strided axis
};

struct PassArgs {
  int L, nl, Lp;         // FFT length, lines per work-group, LDS distance of lines
  int nst, rad[kMaxStages];  // Stockham stages and their radices
  int nlines, ninner;    // line = outer * ninner + inner
  int64_t i_os, i_is, i_es, pair_in;
  int n_in;
  int64_t o_os, o_is, o_es, pair_out;
  int n_out;
  float scale;    // scale of OUT_PHIPAIR
  float hsq;      // h^2 for IN_KGEN
  double fix_inv; // IN_VFIX: 1 / fixed-point scale
  int H;          // L / 2 + 1
  int kdim;       // spectrum index form for MID_CONV
  int kpow;       // IN_KGEN: 0 -> (K2, K1) pair, 2 -> K2 only
  int in_c[4][2], out_c[4][2], kcplx[4];  // per-pair channels (-1: none) and spectrum kind
  int npairs;     // channel pairs of the convolution
  int pairloop;   // a 3-D multiply pass: each work-group runs every pair of its lines
  const void *in;
  void *out;
  const float2 *Kh;
  const float2 *tw;  // twiddle table of length L
};

struct ConvPtrs {
  const void *V;  // float [d+1][ng^d] (or int64 fixed point for the batched engine)
  float *Phi;     // [d+1 (+1)][ng^d]
  float2 *W, *Kt1, *Kt2, *Kh;
  const float2 *tw;
};

__host__ __device__ inline int mirror(int e, int L) { return e <= (L >> 1) ? e : L - e; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// i * sg * z
__device__ __forceinline__ float2 mul_isg(float2 z, float sg) { return make_float2(-sg * z.y, sg * z.x); }

// R-point DFT in place; sg = -1 forward, +1 inverse (unnormalised).
template <int R>
__device__ __forceinline__ void butterfly(float2 (&v)[R], float sg) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const float2 t0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 t1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 t2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 t3 = mul_isg(make_float2(v[1].x - v[3].x, v[1].y - v[3].y), sg);
    v[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
    v[1] = make_float2(t1.x + t3.x, t1.y + t3.y);
    v[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
    v[3] = make_float2(t1.x - t3.x, t1.y - t3.y);
  } else if constexpr (R == 3) {
    const float c = -0.5f, s3 = 0.86602540378443864676f;  // cos, sin of 2 pi / 3
    const float2 a = make_float2(v[1].x + v[2].x, v[1].y + v[2].y);
    const float2 d = mul_isg(make_float2(v[1].x - v[2].x, v[1].y - v[2].y), sg * s3);
    const float2 m = make_float2(v[0].x + c * a.x, v[0].y + c * a.y);
    v[0] = make_float2(v[0].x + a.x, v[0].y + a.y);
    v[1] = make_float2(m.x + d.x, m.y + d.y);
    v[2] = make_float2(m.x - d.x, m.y - d.y);
  } else {  // R == 5
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;  // cos 2pi/5, cos 4pi/5
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;   // sin 2pi/5, sin 4pi/5
    const float2 a1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y);
    const float2 b1 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y);
    const float2 a2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y);
    const float2 b2 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
    const float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    const float2 n1 = mul_isg(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sg);
    const float2 n2 = mul_isg(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sg);
    v[0] = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = make_float2(m1.x + n1.x, m1.y + n1.y);
    v[4] = make_float2(m1.x - n1.x, m1.y - n1.y);
    v[2] = make_float2(m2.x + n2.x, m2.y + n2.y);
    v[3] = make_float2(m2.x - n2.x, m2.y - n2.y);
  }
}

// One radix-R Stockham stage over nl lines of length L; Ns is the length of
// the sub-transforms completed by the previous stages.  A thread owns one
// butterfly position j and applies it to every G-th line, so the position's
// index arithmetic and twiddles are computed once per stage, not per line.
template <int R>
__device__ __forceinline__ void stockham_stage(const float2 *x, float2 *y, int L, int Lp, int nl, int Ns,
                                               const float2 *__restrict__ tw, bool inv) {
  const int LR = L / R;
  const int stride = L / (Ns * R);
  const float invLR = 1.0f / (float)LR, invNs = 1.0f / (float)Ns;
  const float sg = inv ? 1.0f : -1.0f;
  const int G = LR >= kFftThreads ? 1 : kFftThreads / LR;  // line groups
  for (int s = threadIdx.x; s < G * LR; s += kFftThreads) {
    const int g = fdiv(s, LR, invLR);
    const int j = s - g * LR;
    const int q = fdiv(j, Ns, invNs);
    const int k = j - q * Ns;
    float2 w[R];
#pragma unroll
    for (int t = 1; t < R; ++t) {
      w[t] = tw[t * k * stride];
      if (inv) w[t].y = -w[t].y;
    }
    const int base = q * Ns * R + k;
    for (int l = g; l < nl; l += G) {
      const float2 *src = x + l * Lp;
      float2 *dst = y + l * Lp;
      float2 v[R];
#pragma unroll
      for (int t = 0; t < R; ++t) v[t] = src[j + t * LR];
#pragma unroll
      for (int t = 1; t < R; ++t) v[t] = cmul(v[t], w[t]);
      butterfly<R>(v, sg);
#pragma unroll
      for (int t = 0; t < R; ++t) dst[base + t * Ns] = v[t];
    }
  }
}

// Radix of stage s of a length-L transform in the order base_args uses
// (radix 4 first, then 2, 3, 5); 0 past the last stage.
constexpr int fixed_radix(int L, int s) {
  int m = L, st = 0;
  for (int r : {4, 2, 3, 5})
    while (m % r == 0) {
      if (st == s) return r;
      m /= r;
      ++st;
    }
  return 0;
}
constexpr int fixed_ns(int L, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= fixed_radix(L, i);
  return ns;
}

// stockham_stage for a length known at compile time: the same butterflies,
// twiddles and order of operations (so the same bits), with the index
// divisions by constants and the element offsets of a butterfly immediate.
template <int L, int R, int Ns>
__device__ __forceinline__ void stockham_stage_fixed(const float2 *x, float2 *y, int Lp, int nl,
                                                     const float2 *__restrict__ tw, bool inv) {
  constexpr int LR = L / R;
  constexpr int stride = L / (Ns * R);
  constexpr int G = LR >= kFftThreads ? 1 : kFftThreads / LR;
  const float sg = inv ? 1.0f : -1.0f;
  // the thread index through an empty asm statement: the per-thread index
  // work below then stays inside its stage instead of being hoisted out of
  // the callers' loops, where the stages of every length would all hold
  // their registers at once
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
#pragma unroll 1
  for (int s = tid; s < G * LR; s += kFftThreads) {
    const int g = s / LR;
    const int j = s - g * LR;
    const int q = j / Ns;
    const int k = j - q * Ns;
    float2 w[R];
#pragma unroll
    for (int t = 1; t < R; ++t) {
      w[t] = tw[t * k * stride];
      if (inv) w[t].y = -w[t].y;
    }
    const int base = q * Ns * R + k;
#pragma unroll 1
    for (int l = g; l < nl; l += G) {
      const float2 *src = x + l * Lp + j;
      float2 *dst = y + l * Lp + base;
      float2 v[R];
#pragma unroll
      for (int t = 0; t < R; ++t) v[t] = src[t * LR];
#pragma unroll
      for (int t = 1; t < R; ++t) v[t] = cmul(v[t], w[t]);
      butterfly<R>(v, sg);
#pragma unroll
      for (int t = 0; t < R; ++t) dst[t * Ns] = v[t];
    }
  }
}

template <int L, int S = 0>
__device__ __forceinline__ float2 *stockham_fixed(float2 *x, float2 *y, int Lp, int nl, bool inv,
                                                  const float2 *tw) {
  constexpr int R = fixed_radix(L, S);
  if constexpr (R == 0) {
    return x;
  } else {
    stockham_stage_fixed<L, R, fixed_ns(L, S)>(x, y, Lp, nl, tw, inv);
    __syncthreads();
    return stockham_fixed<L, S + 1>(y, x, Lp, nl, inv, tw);
  }
}

// Stockham autosort on nl lines (natural order in and out) with the
// twiddle table tw (LDS).  Returns the buffer holding the result.  The
// lengths of grid sides 16-72 (the batched engine's 3-D grids and the
// smaller 2-D ones) run as compile-time-length code, the others generically.
__device__ __forceinline__ float2 *stockham(float2 *x, float2 *y, const PassArgs &a, int nl, bool inv,
                                           const float2 *tw) {
  switch (a.L) {
    case 32: return stockham_fixed<32>(x, y, a.Lp, nl, inv, tw);
    case 60: return stockham_fixed<60>(x, y, a.Lp, nl, inv, tw);
    case 64: return stockham_fixed<64>(x, y, a.Lp, nl, inv, tw);
    case 72: return stockham_fixed<72>(x, y, a.Lp, nl, inv, tw);
    case 80: return stockham_fixed<80>(x, y, a.Lp, nl, inv, tw);
    case 90: return stockham_fixed<90>(x, y, a.Lp, nl, inv, tw);
    case 96: return stockham_fixed<96>(x, y, a.Lp, nl, inv, tw);
    case 100: return stockham_fixed<100>(x, y, a.Lp, nl, inv, tw);
    case 108: return stockham_fixed<108>(x, y, a.Lp, nl, inv, tw);
    case 120: return stockham_fixed<120>(x, y, a.Lp, nl, inv, tw);
    case 128: return stockham_fixed<128>(x, y, a.Lp, nl, inv, tw);
    case 144: return stockham_fixed<144>(x, y, a.Lp, nl, inv, tw);
    default: break;
  }
  int Ns = 1;
  for (int s = 0; s < a.nst; ++s) {
    const int R = a.rad[s];
    if (R == 4) stockham_stage<4>(x, y, a.L, a.Lp, nl, Ns, tw, inv);
    else if (R == 2) stockham_stage<2>(x, y, a.L, a.Lp, nl, Ns, tw, inv);
    else if (R == 3) stockham_stage<3>(x, y, a.L, a.Lp, nl, Ns, tw, inv);
    else stockham_stage<5>(x, y, a.L, a.Lp, nl, Ns, tw, inv);
    __syncthreads();
    float2 *tmp = x;
    x = y;
    y = tmp;
    Ns *= R;
  }
  return x;
}

// Visits every element (l, e) of a work-group's nl lines of length L once,
// for a pass along the contiguous axis: adjacent threads take adjacent e and
// a thread walks its lines.  Calls f(l, e, line, outer, inner, oi, oo) with
// line = line0 + l split as outer * ninner + inner and the element offsets
// oi = outer i_os + inner i_is + e i_es (oo likewise with the o_ strides),
// all stepped by addition from one split per thread.
template <class F>
__device__ __forceinline__ void for_elements(const PassArgs &a, int nl, int L, int line0, float invni, F &&f) {
  const int G = L >= kFftThreads ? 1 : kFftThreads / L;
  const float invL = 1.0f / (float)L;
  const int ninner = a.ninner;
  const int64_t si = (int64_t)G * a.i_is, so = (int64_t)G * a.o_is;        // G lines on
  const int64_t wi = a.i_os - (int64_t)ninner * a.i_is, wo = a.o_os - (int64_t)ninner * a.o_is;  // wrap
  for (int s = threadIdx.x; s < G * L; s += kFftThreads) {
    const int g = fdiv(s, L, invL);
    const int e = s - g * L;
    int line = line0 + g;
    int outer = fdiv(line, ninner, invni);
    int inner = line - outer * ninner;
    int64_t oi = outer * a.i_os + inner * a.i_is + e * a.i_es;
    int64_t oo = outer * a.o_os + inner * a.o_is + e * a.o_es;
    for (int l = g; l < nl; l += G) {
      f(l, e, line, outer, inner, oi, oo);
      line += G;
      inner += G;
      oi += si;
      oo += so;
      while (inner >= ninner) {
        inner -= ninner;
        ++outer;
        oi += wi;
        oo += wo;
      }
    }
  }
}

// Elements of one line a thread visits in a strided pass (lines across
// threads): at most ceil(L / (kFftThreads / nl)) over the supported lengths,
// checked at compile time below.
constexpr int kLineElems = 10;

// Line layout of a pass: `padded` passes (lines across threads) keep lines
// L + 1 elements apart in LDS, so the threads of a wave, which take adjacent
// lines at the same element, fall on different banks.
__host__ __device__ inline void set_lines(PassArgs &a, bool padded) {
  a.Lp = padded ? a.L + 1 : a.L;
  int nl = kLdsElems / a.Lp;
  if (nl < 1) nl = 1;
  if (nl > 64) nl = 64;
  a.nl = nl;
}

constexpr bool line_elems_fit() {
  for (int L = 2; L <= kMaxL; L += 2) {
    int m = L;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    if (m != 1) continue;
    int nl = kLdsElems / (L + 1);
    nl = nl < 1 ? 1 : nl > 64 ? 64 : nl;
    const int G = kFftThreads / nl;
    if ((L + G - 1) / G > kLineElems) return false;
  }
  return true;
}
static_assert(line_elems_fit(), "kLineElems below the elements per thread of a strided pass");

// One work-group's share of one pass: lines [blk * nl, blk * nl + nl) of
// channel pair p.  Every thread of the work-group must call it (barriers);
// the work-group has kFftThreads threads and nl * Lp <= kLdsElems.
// The global reads a work-group needs (its lines, the twiddle table and,
// for a convolution, the kernel spectrum) are all issued before the first
// barrier, so their latencies overlap instead of adding up stage by stage.
template <int IN, int OUT, int MID, bool ECONTIG>
__device__ inline void pass_body(const PassArgs &a, int blk, int p, float2 *lds) {
  const int L = a.L;
  const int Lp = a.Lp;
  const int nl = a.nl;
  // line indices stay below 2^22 (at most 1024^2 lines), where fdiv is exact
  const float invL = 1.0f / (float)L, invni = 1.0f / (float)a.ninner;
  float2 *b0 = lds;
  float2 *b1 = lds + nl * Lp;
  float2 *twl = lds + 2 * nl * Lp;
  const int line0 = blk * nl;

  for (int t = threadIdx.x; t < L; t += kFftThreads) twl[t] = a.tw[t];

  if constexpr (!ECONTIG) {
    // Lines across threads (the pass axis is strided in memory): thread
    // (g, l) owns elements e = g, g + G, ... of line l, for the load, the
    // spectrum multiply and the store alike, so each of those walks one
    // pointer by a fixed step.
    static_assert(IN == IN_W || IN == IN_KT, "strided passes read W or kernel half-spectra");
    static_assert(OUT == OUT_W || OUT == OUT_KT, "strided passes write W or kernel half-spectra");
    const int G = kFftThreads / nl;
    const int s = threadIdx.x;
    const bool act = s < G * nl;
    const int g = act ? fdiv(s, nl, 1.0f / (float)nl) : 0;
    const int l = s - g * nl;
    const int line = line0 + l;
    const bool lv = act && line < a.nlines;
    const int outer = lv ? fdiv(line, a.ninner, invni) : 0;
    const int inner = line - outer * a.ninner;
    // the 3-D multiply pass runs every channel pair of its lines, so the
    // kernel spectrum of those lines is read once for all of them (in 2-D
    // one pair per work-group keeps more work-groups for the few lines)
    const bool loop = MID == MID_CONV && a.pairloop;
    const int pa = loop ? 0 : p, pb = loop ? a.npairs : p + 1;
    float2 kreg[MID == MID_CONV ? kLineElems : 1];
    if (MID == MID_CONV && act) {
      // kernel spectrum at (mirror(e), mirror(outer), mirror(inner)):
      // index mirror(e) * ks + kb
      int ks = 1, kb = 0;
      if (a.kdim == 2) {
        ks = a.H;
        kb = mirror(inner, L);
      } else if (a.kdim == 3) {
        ks = a.H * a.H;
        kb = mirror(outer, L) * a.H + mirror(inner, L);
      }
#pragma unroll
      for (int k = 0; k < kLineElems; ++k) {
        const int e = g + k * G;
        kreg[k] = lv && e < L ? a.Kh[mirror(e, L) * ks + kb] : make_float2(0.f, 0.f);
      }
    }
    for (int pp = pa; pp < pb; ++pp) {
      if (act) {
        float2 *dl = b0 + l * Lp;
        if (IN == IN_W) {
          const float2 *src = static_cast<const float2 *>(a.in) + (int64_t)pp * a.pair_in + outer * a.i_os +
                              inner * a.i_is + g * a.i_es;
          const int64_t step = (int64_t)G * a.i_es;
          const int ne = lv ? a.n_in : 0;
          for (int e = g; e < L; e += G, src += step) dl[e] = e < ne ? *src : make_float2(0.f, 0.f);
        } else {  // IN_KT: (K2, K1) half-spectra, mirrored along the pass axis
          const float2 *src = static_cast<const float2 *>(a.in) + outer * a.i_os + inner * a.i_is;
          const int es = (int)a.i_es;
          const int ne = lv ? a.n_in : 0;
          for (int e = g; e < L; e += G) dl[e] = e < ne ? src[mirror(e, L) * es] : make_float2(0.f, 0.f);
        }
      }
      __syncthreads();

      float2 *r = stockham(b0, b1, a, nl, MID == MID_INV, twl);
      if (MID == MID_CONV) {
        // multiply by the kernel spectrum (real, or K2^ + i K1^ for the
        // pair that carries the charge alone), then inverse transform
        if (act) {
          float2 *rl = r + l * Lp;
          const bool cplx = a.kcplx[pp] != 0;
#pragma unroll
          for (int k = 0; k < kLineElems; ++k) {
            const int e = g + k * G;
            if (e < L) {
              const float kv = kreg[k].x, kvi = cplx ? kreg[k].y : 0.0f;
              const float2 v = rl[e];
              rl[e] = make_float2(v.x * kv - v.y * kvi, v.x * kvi + v.y * kv);
            }
          }
        }
        __syncthreads();
        float2 *other = (r == b0) ? b1 : b0;
        r = stockham(r, other, a, nl, true, twl);
      }

      if (lv) {
        const float2 *rl = r + l * Lp;
        float2 *dst = static_cast<float2 *>(a.out) + (OUT == OUT_W ? (int64_t)pp * a.pair_out : 0) +
                      outer * a.o_os + inner * a.o_is + g * a.o_es;
        const int64_t step = (int64_t)G * a.o_es;
        for (int e = g; e < a.n_out; e += G, dst += step) *dst = rl[e];
      }
      __syncthreads();  // LDS is reused by the next pair or the caller's next share
    }
    return;
  } else {
  const int total = nl * L;
  // kernel spectrum of each element this thread multiplies (real part, and
  // the imaginary part for the complex K2^ + i K1^ spectrum)
  float2 kreg[MID == MID_CONV ? kPerThread : 1];
  if (MID == MID_CONV) {
#pragma unroll
    for (int t = 0; t < kPerThread; ++t) {
      const int idx = threadIdx.x + t * kFftThreads;
      kreg[t] = make_float2(0.f, 0.f);
      if (idx < total) {
        const int l = fdiv(idx, L, invL), e = idx - l * L;
        const int line = line0 + l;
        if (line < a.nlines) {
          const int outer = fdiv(line, a.ninner, invni);
          const int inner = line - outer * a.ninner;
          const int me = mirror(e, L);
          int64_t ki;
          if (a.kdim == 1) ki = me;
          else if (a.kdim == 2) ki = (int64_t)me * a.H + mirror(inner, L);
          else ki = ((int64_t)me * a.H + mirror(outer, L)) * a.H + mirror(inner, L);
          const float2 k = a.Kh[ki];
          kreg[t] = make_float2(k.x, a.kcplx[p] ? k.y : 0.0f);
        }
      }
    }
  }

  // ---- load (zero padding beyond n_in) ----
  for_elements(a, nl, L, line0, invni, [&](int l, int e, int line, int outer, int inner, int64_t off, int64_t) {
    float2 v = make_float2(0.f, 0.f);
    if (line < a.nlines && e < a.n_in) {
      if (IN == IN_VPAIR) {
        const float *V = static_cast<const float *>(a.in);
        v.x = V[(int64_t)a.in_c[p][0] * a.pair_in + off];
        if (a.in_c[p][1] >= 0) v.y = V[(int64_t)a.in_c[p][1] * a.pair_in + off];
      } else if (IN == IN_VFIX) {
        long long *V = static_cast<long long *>(const_cast<void *>(a.in));
        long long *p0 = V + (int64_t)a.in_c[p][0] * a.pair_in + off;
        v.x = (float)((double)*p0 * a.fix_inv);
        *p0 = 0;
        if (a.in_c[p][1] >= 0) {
          long long *p1 = V + (int64_t)a.in_c[p][1] * a.pair_in + off;
          v.y = (float)((double)*p1 * a.fix_inv);
          *p1 = 0;
        }
      } else if (IN == IN_W) {
        const float2 *W = static_cast<const float2 *>(a.in);
        v = W[(int64_t)p * a.pair_in + off];
      } else if (IN == IN_KGEN) {
        const int me = mirror(e, L);
        const float r2 = (float)me * me + (float)inner * inner + (float)outer * outer;
        const float k1 = 1.0f / (1.0f + a.hsq * r2);
        v.x = k1 * k1;
        if (a.kpow == 0) v.y = k1;
      } else {  // IN_KT: (K2, K1) half-spectra, mirrored along the pass axis
        const float2 *K = static_cast<const float2 *>(a.in);
        const int me = mirror(e, L);
        v = K[outer * a.i_os + inner * a.i_is + me * a.i_es];
      }
    }
    b0[l * Lp + e] = v;
  });
  __syncthreads();

  float2 *r = stockham(b0, b1, a, nl, MID == MID_INV, twl);
  if (MID == MID_CONV) {
    // multiply by the kernel spectrum, then inverse transform
#pragma unroll
    for (int t = 0; t < kPerThread; ++t) {
      const int idx = threadIdx.x + t * kFftThreads;
      if (idx < total) {
        const int l = fdiv(idx, L, invL), e = idx - l * L;
        const float kv = kreg[t].x, kvi = kreg[t].y;
        const float2 v = r[l * Lp + e];
        r[l * Lp + e] = make_float2(v.x * kv - v.y * kvi, v.x * kvi + v.y * kv);
      }
    }
    __syncthreads();
    float2 *other = (r == b0) ? b1 : b0;
    r = stockham(r, other, a, nl, true, twl);
  }

  // ---- store ----
  for_elements(a, nl, L, line0, invni, [&](int l, int e, int line, int, int, int64_t, int64_t off) {
    if (line >= a.nlines || e >= a.n_out) return;
    const float2 v = r[l * Lp + e];
    if (OUT == OUT_W) {
      float2 *W = static_cast<float2 *>(a.out);
      W[(int64_t)p * a.pair_out + off] = v;
    } else if (OUT == OUT_PHIPAIR) {
      float *Phi = static_cast<float *>(a.out);
      Phi[(int64_t)a.out_c[p][0] * a.pair_out + off] = v.x * a.scale;
      if (a.out_c[p][1] >= 0) Phi[(int64_t)a.out_c[p][1] * a.pair_out + off] = v.y * a.scale;
    } else if (OUT == OUT_PHIQ) {
      float *Phi = static_cast<float *>(a.out);
      const int c0 = a.out_c[p][0], c1 = a.out_c[p][1];
      Phi[c0 < 4 ? off * 4 + c0 : 4 * a.pair_out + off] = v.x * a.scale;
      if (c1 >= 0) Phi[c1 < 4 ? off * 4 + c1 : 4 * a.pair_out + off] = v.y * a.scale;
    } else {  // the spectra of real even kernels are real: keep (Re, Im) = (K2^, K1^)
      float2 *K = static_cast<float2 *>(a.out);
      K[off] = v;
    }
  });
  __syncthreads();  // LDS is reused by the caller's next share
  }
}

// Work-groups of one pass (per channel pair).
__host__ __device__ inline int pass_blocks(const PassArgs &a) { return (a.nlines + a.nl - 1) / a.nl; }

// Pass shapes of a d-dimensional convolution, in launch order.
__host__ __device__ inline int pass_kinds(int d, int *kinds) {
  if (d == 1) {
    kinds[0] = PK_KGEN; kinds[1] = PK_V_PHI;
    return 2;
  }
  if (d == 2) {
    kinds[0] = PK_KGEN; kinds[1] = PK_KT_T; kinds[2] = PK_V_W; kinds[3] = PK_W_W_CONV; kinds[4] = PK_W_PHI;
    return 5;
  }
  kinds[0] = PK_KGEN; kinds[1] = PK_KT_T; kinds[2] = PK_KT_T; kinds[3] = PK_V_W; kinds[4] = PK_W_W_FWD;
  kinds[5] = PK_W_W_CONV; kinds[6] = PK_W_W_INV; kinds[7] = PK_W_PHI;
  return 8;
}

// Channel plan of one convolution.  zdirect: pair 0 = (V0, -) with the
// spectrum K2^ + i K1^ -> (Phi0, Psi0 in channel d+1); then the charge
// channels 1..d in pairs with K2^.  Otherwise: channels 0..d in pairs with K2^.
struct Plan {
  int pairs;
  int in_c[4][2], out_c[4][2], kcplx[4];
};

__host__ __device__ inline Plan make_plan(int d, bool zdirect) {
  Plan pl{};
  int p = 0;
  int c = 0;
  if (zdirect) {
    pl.in_c[0][0] = 0; pl.in_c[0][1] = -1;
    pl.out_c[0][0] = 0; pl.out_c[0][1] = d + 1;
    pl.kcplx[0] = 1;
    p = 1;
    c = 1;
  }
  for (; c <= d; c += 2, ++p) {
    pl.in_c[p][0] = c;
    pl.in_c[p][1] = c + 1 <= d ? c + 1 : -1;
    pl.out_c[p][0] = pl.in_c[p][0];
    pl.out_c[p][1] = pl.in_c[p][1];
    pl.kcplx[p] = 0;
  }
  for (; p < 4; ++p) {
    pl.in_c[p][0] = pl.in_c[p][1] = pl.out_c[p][0] = pl.out_c[p][1] = -1;
    pl.kcplx[p] = 0;
  }
  pl.pairs = zdirect ? 1 + (d + 1) / 2 : (d + 2) / 2;
  return pl;
}

// The fields every pass shares, written in place (a may live in global
// memory: no private copy of the argument block is made).
__host__ __device__ inline void base_args(PassArgs &a, const GridParams &g, const Plan &pl, const ConvPtrs &b) {
  a = PassArgs{};
  a.L = g.L;
  set_lines(a, false);
  // radix-4 stages first, then 2, 3, 5
  int m = g.L, ns = 0;
  while (m % 4 == 0 && ns < kMaxStages) { a.rad[ns++] = 4; m /= 4; }
  while (m % 2 == 0 && ns < kMaxStages) { a.rad[ns++] = 2; m /= 2; }
  while (m % 3 == 0 && ns < kMaxStages) { a.rad[ns++] = 3; m /= 3; }
  while (m % 5 == 0 && ns < kMaxStages) { a.rad[ns++] = 5; m /= 5; }
  a.nst = ns;
  a.H = g.L / 2 + 1;
  a.hsq = (float)(g.hgrid * g.hgrid);
  a.tw = b.tw + fft_table_offset(g.L);
  a.Kh = b.Kh;
  a.scale = 1.0f;
  a.fix_inv = 1.0;
  a.ninner = 1;
  a.kpow = 0;
  for (int p = 0; p < 4; ++p) {
    a.in_c[p][0] = pl.in_c[p][0];
    a.in_c[p][1] = pl.in_c[p][1];
    a.out_c[p][0] = pl.out_c[p][0];
    a.out_c[p][1] = pl.out_c[p][1];
    a.kcplx[p] = pl.kcplx[p];
  }
}

// Passes of a d-dimensional convolution.
__host__ __device__ inline int conv_npasses(int d) { return d == 1 ? 2 : d == 2 ? 5 : 8; }

// Arguments of pass i (in the order of pass_kinds) of one convolution,
// written in place, lines unpadded; returns the channel pairs the pass runs over.
__host__ __device__ inline int conv_pass_shape(const GridParams &g, bool zdirect, const ConvPtrs &b, int i,
                                               PassArgs &a) {
  const int d = g.d, ng = g.ng, L = g.L, H = L / 2 + 1;
  const Plan pl = make_plan(d, zdirect);
  const int64_t G = d == 1 ? ng : d == 2 ? (int64_t)ng * ng : (int64_t)ng * ng * ng;
  const int kpow = zdirect ? 0 : 2;
  base_args(a, g, pl, b);

  if (d == 1) {
    if (i == 0) {  // kernel spectra
      a.nlines = 1; a.ninner = 1; a.n_in = L; a.n_out = H; a.o_es = 1; a.kpow = kpow;
      a.out = b.Kh;
      return 1;
    }
    a.nlines = 1; a.ninner = 1;
    a.in = b.V; a.i_es = 1; a.n_in = ng; a.pair_in = G;
    a.out = b.Phi; a.o_es = 1; a.n_out = ng; a.pair_out = G;
    a.kdim = 1; a.scale = 1.0f / (float)L;
    return pl.pairs;
  }

  if (d == 2) {
    const int64_t wp = (int64_t)ng * L;
    switch (i) {
      case 0:
        a.nlines = H; a.ninner = H; a.n_in = L; a.o_is = H; a.o_es = 1; a.n_out = H; a.kpow = kpow;
        a.out = b.Kt1;
        return 1;
      case 1:
        a.nlines = H; a.ninner = H; a.in = b.Kt1; a.i_is = 1; a.i_es = H; a.n_in = L;
        a.out = b.Kh; a.o_is = 1; a.o_es = H; a.n_out = H;
        return 1;
      case 2:  // rows: V -> W
        a.nlines = ng; a.ninner = ng;
        a.in = b.V; a.i_is = ng; a.i_es = 1; a.n_in = ng; a.pair_in = G;
        a.out = b.W; a.o_is = L; a.o_es = 1; a.n_out = L; a.pair_out = wp;
        return pl.pairs;
      case 3:  // columns: fwd, x spectrum, inv
        a.nlines = L; a.ninner = L;
        a.in = b.W; a.i_is = 1; a.i_es = L; a.n_in = ng; a.pair_in = wp;
        a.out = b.W; a.o_is = 1; a.o_es = L; a.n_out = ng; a.pair_out = wp;
        a.kdim = 2;
        return pl.pairs;
      default:  // rows: inverse -> Phi
        a.nlines = ng; a.ninner = ng;
        a.in = b.W; a.i_is = L; a.i_es = 1; a.n_in = L; a.pair_in = wp;
        a.out = b.Phi; a.o_is = ng; a.o_es = 1; a.n_out = ng; a.pair_out = G;
        a.scale = 1.0f / ((float)L * (float)L);
        return pl.pairs;
    }
  }

  // d == 3
  const int64_t H2 = (int64_t)H * H, L2 = (int64_t)L * L;
  const int64_t wp = (int64_t)ng * L2;
  switch (i) {
    case 0:
      a.nlines = (int)H2; a.ninner = H; a.n_in = L; a.kpow = kpow;
      a.out = b.Kt1; a.o_os = H2; a.o_is = H; a.o_es = 1; a.n_out = H;
      return 1;
    case 1:
      a.nlines = (int)H2; a.ninner = H; a.in = b.Kt1; a.i_os = H2; a.i_is = 1; a.i_es = H; a.n_in = L;
      a.out = b.Kt2; a.o_os = H2; a.o_is = 1; a.o_es = H; a.n_out = H;
      return 1;
    case 2:
      a.nlines = (int)H2; a.ninner = H; a.in = b.Kt2; a.i_os = H; a.i_is = 1; a.i_es = H2; a.n_in = L;
      a.out = b.Kh; a.o_os = H; a.o_is = 1; a.o_es = H2; a.n_out = H;
      return 1;
    case 3:  // x: V -> W
      a.nlines = ng * ng; a.ninner = ng;
      a.in = b.V; a.i_os = (int64_t)ng * ng; a.i_is = ng; a.i_es = 1; a.n_in = ng; a.pair_in = G;
      a.out = b.W; a.o_os = L2; a.o_is = L; a.o_es = 1; a.n_out = L; a.pair_out = wp;
      return pl.pairs;
    case 4:  // y forward (all ky kept)
      a.nlines = ng * L; a.ninner = L;
      a.in = b.W; a.i_os = L2; a.i_is = 1; a.i_es = L; a.n_in = ng; a.pair_in = wp;
      a.out = b.W; a.o_os = L2; a.o_is = 1; a.o_es = L; a.n_out = L; a.pair_out = wp;
      return pl.pairs;
    case 5:  // z: fwd, x spectrum, inv (z < ng kept)
      a.nlines = (int)L2; a.ninner = L;
      a.in = b.W; a.i_os = L; a.i_is = 1; a.i_es = L2; a.n_in = ng; a.pair_in = wp;
      a.out = b.W; a.o_os = L; a.o_is = 1; a.o_es = L2; a.n_out = ng; a.pair_out = wp;
      a.kdim = 3;
      return pl.pairs;
    case 6:  // y inverse (y < ng kept)
      a.nlines = ng * L; a.ninner = L;
      a.in = b.W; a.i_os = L2; a.i_is = 1; a.i_es = L; a.n_in = L; a.pair_in = wp;
      a.out = b.W; a.o_os = L2; a.o_is = 1; a.o_es = L; a.n_out = ng; a.pair_out = wp;
      return pl.pairs;
    default:  // x inverse -> Phi
      a.nlines = ng * ng; a.ninner = ng;
      a.in = b.W; a.i_os = L2; a.i_is = L; a.i_es = 1; a.n_in = L; a.pair_in = wp;
      a.out = b.Phi; a.o_os = (int64_t)ng * ng; a.o_is = ng; a.o_es = 1; a.n_out = ng; a.pair_out = G;
      a.scale = 1.0f / ((float)L * (float)L * (float)L);
      return pl.pairs;
  }
}

// Whether a pass kind walks a strided axis (lines across threads).
__host__ __device__ inline bool pass_strided(int kind) {
  return kind == PK_KT_T || kind == PK_W_W_CONV || kind == PK_W_W_FWD || kind == PK_W_W_INV;
}

// Arguments of pass i (in the order of pass_kinds) of one convolution,
// written in place; returns the channel pairs the pass runs over.
__host__ __device__ inline int conv_pass(const GridParams &g, bool zdirect, const ConvPtrs &b, int i, PassArgs &a) {
  const int pairs = conv_pass_shape(g, zdirect, b, i, a);
  int kinds[kMaxPasses];
  pass_kinds(g.d, kinds);
  set_lines(a, pass_strided(kinds[i]));
  a.npairs = make_plan(g.d, zdirect).pairs;
  a.pairloop = kinds[i] == PK_W_W_CONV && g.d == 3;
  return a.pairloop ? 1 : pairs;  // its work-groups run every pair
}

// Arguments of every pass of one convolution (in the order of pass_kinds),
// and the channel pairs each pass runs over.  Returns the pass count.
__host__ __device__ inline int conv_setup(const GridParams &g, bool zdirect, const ConvPtrs &b, PassArgs *out,
                                          int *pairs_out) {
  const int np = conv_npasses(g.d);
  for (int i = 0; i < np; ++i) pairs_out[i] = conv_pass(g, zdirect, b, i, out[i]);
  return np;
}

}  // namespace fft
}  // namespace sgt
