// kernels_fft.hip -- non-periodic grid convolution of one job with the kernels
// K2 = (1 + h^2 r^2)^-2 and K1 = (1 + h^2 r^2)^-1 by zero-padded FFTs in LDS
// (the passes themselves are in fft_device.hpp, shared with the batched
// engine).
#include "common.hpp"
#include "kernels.hpp"
#include "fft_device.hpp"
#include "../../include/sgtsne.h"

#include <cmath>

namespace sgt {

namespace {

using namespace fft;

template <int IN, int OUT, int MID, bool EC>
__global__ __launch_bounds__(kFftThreads) void fft_pass_kernel(PassArgs a) {
  extern __shared__ float2 lds[];
  pass_body<IN, OUT, MID, EC>(a, blockIdx.x, blockIdx.y, lds);
}

// Fills every twiddle table: entry t of the table of length L is
// exp(-2 pi i t / L), evaluated in FP64 and rounded once.
__global__ void twiddle_kernel(float2 *tw, int total) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int off = 0, L = 2;
  for (; L <= kMaxL; L += 2) {
    if (!fft_smooth(L)) continue;
    if (g < off + L) break;
    off += L;
  }
  const int t = g - off;
  double sn, cs;
  sincospi(-2.0 * (double)t / (double)L, &sn, &cs);
  tw[g] = make_float2((float)cs, (float)sn);
}

template <int IN, int OUT, int MID, bool EC>
int run_pass(hipStream_t s, const PassArgs &a, int pairs) {
  const int blocks = pass_blocks(a);
  if (blocks == 0) return SGTSNE_OK;
  const size_t lds = pass_lds_bytes(a.nl, a.L, a.Lp);
  dim3 grid(blocks, pairs);
  hipLaunchKernelGGL((fft_pass_kernel<IN, OUT, MID, EC>), grid, dim3(kFftThreads), lds, s, a);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int run_kind(hipStream_t s, int kind, const PassArgs &a, int pairs) {
  switch (kind) {
    case PK_KGEN: return run_pass<IN_KGEN, OUT_KT, MID_FWD, true>(s, a, pairs);
    case PK_KT_T: return run_pass<IN_KT, OUT_KT, MID_FWD, false>(s, a, pairs);
    case PK_V_W: return run_pass<IN_VPAIR, OUT_W, MID_FWD, true>(s, a, pairs);
    case PK_W_W_CONV: return run_pass<IN_W, OUT_W, MID_CONV, false>(s, a, pairs);
    case PK_W_PHI: return run_pass<IN_W, OUT_PHIPAIR, MID_INV, true>(s, a, pairs);
    case PK_V_PHI: return run_pass<IN_VPAIR, OUT_PHIPAIR, MID_CONV, true>(s, a, pairs);
    case PK_W_W_FWD: return run_pass<IN_W, OUT_W, MID_FWD, false>(s, a, pairs);
    case PK_W_W_INV: return run_pass<IN_W, OUT_W, MID_INV, false>(s, a, pairs);
    default: SGT_CHECK(false, SGTSNE_ERR_ARG, "bad FFT pass kind");
  }
}

}  // namespace

size_t twiddle_count() { return (size_t)fft_table_size(); }

int fft_length_for(int ng) { return fft_length(2 * ng - 1); }

int init_twiddles(hipStream_t s, float2 *tw) {
  const int total = fft_table_size();
  twiddle_kernel<<<div_up(total, 256), 256, 0, s>>>(tw, total);
  SGT_LAUNCH_CHECK();
  return SGTSNE_OK;
}

int conv_workspace_bytes(int d, int ng, int L, size_t *w_bytes, size_t *k_bytes) {
  const int pairs = 3;  // at most (1 + ceil(d / 2)) for d <= 3
  const size_t H = (size_t)L / 2 + 1;
  size_t wl = d == 1 ? 0 : d == 2 ? (size_t)ng * L : (size_t)ng * L * L;
  *w_bytes = pairs * wl * sizeof(float2);
  size_t hd = d == 1 ? H : d == 2 ? H * H : H * H * H;
  *k_bytes = hd * sizeof(float2);
  return SGTSNE_OK;
}

int launch_conv(hipStream_t s, const GridParams &g, int zdirect, const float *V, float *Phi,
                const ConvBuffers &b) {
  SGT_CHECK(g.L <= kMaxL && g.L >= 2 * g.ng - 1 && fft_smooth(g.L), SGTSNE_ERR_ARG,
            "FFT length %d invalid for ng %d", g.L, g.ng);
  ConvPtrs ptr{V, Phi, b.W, b.Kt1, b.Kt2, b.Kh, b.tw};
  PassArgs args[kMaxPasses];
  int pairs[kMaxPasses], kinds[kMaxPasses];
  const int np = conv_setup(g, zdirect != 0, ptr, args, pairs);
  pass_kinds(g.d, kinds);
  for (int i = 0; i < np; ++i) SGT_TRY(run_kind(s, kinds[i], args[i], pairs[i]));
  return SGTSNE_OK;
}

}  // namespace sgt
