// noalloc.hip -- the grid convolution without any device allocation, for
// callers that cannot allocate (code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// sgtsne_grid_conv allocates its inputs, outputs and FFT workspace on every
// call.  sgtsne_grid_conv_noalloc takes device inputs and outputs from the
// caller and carves its workspace (complex lines W, the kernel half-spectra
// and their scratch, the twiddle tables) out of one file-scope device array
// of kScratchBytes, which every call on a device shares.  Host code never
// names that array: its address on each device is resolved once, on the
// first call with that device current, and kept per device.
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/sgtsne.h"

#include <mutex>

namespace sgt {

constexpr size_t kScratchBytes = 32ull << 20;  // 32 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace of one convolution: W, three kernel half-spectrum buffers and the
// twiddle tables, each 256-byte aligned (0 for an unsupported size).
size_t conv_scratch_bytes(int d, int ng) {
  if (d < 1 || d > 3 || ng < 4 || ng > 512) return 0;
  size_t wb = 0, kb = 0;
  conv_workspace_bytes(d, ng, fft_length_for(ng), &wb, &kb);
  return round256(wb) + 3 * round256(kb) + round256(twiddle_count() * sizeof(float2));
}

// The array's address on the current device (resolved on first use there).
int device_scratch(void **out) {
  int dev = -1;
  SGT_HIP(hipGetDevice(&dev));
  SGT_CHECK(dev >= 0 && dev < kMaxDevices, SGTSNE_ERR_NODEV, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    SGT_CHECK(e == hipSuccess && p != nullptr, SGTSNE_ERR_HIP, "scratch address on device %d: %s", dev,
              hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return SGTSNE_OK;
}

}  // namespace
}  // namespace sgt

using namespace sgt;

extern "C" {

int sgtsne_grid_conv_noalloc(int d, int ng, double h, int zdirect, const float *V, float *Phi, void *stream) {
  SGT_CHECK(V && Phi, SGTSNE_ERR_ARG, "NULL argument");
  SGT_CHECK(d >= 1 && d <= 3 && ng >= 4 && ng <= 512 && h > 0, SGTSNE_ERR_ARG,
            "bad grid_conv sizes (d=%d ng=%d)", d, ng);
  const size_t need = conv_scratch_bytes(d, ng);
  SGT_CHECK(need <= kScratchBytes, SGTSNE_ERR_ARG,
            "a d=%d grid of side %d needs %zu bytes of workspace; the allocation-free entry holds %zu", d, ng,
            need, kScratchBytes);
  void *base = nullptr;
  SGT_TRY(device_scratch(&base));
  GridParams g{};
  g.d = d;
  g.ng = ng;
  g.nGrid = ng - 2;
  g.L = fft_length_for(ng);
  g.hgrid = h;
  size_t wb = 0, kb = 0;
  conv_workspace_bytes(d, ng, g.L, &wb, &kb);
  char *p = static_cast<char *>(base);
  ConvBuffers cb{};
  cb.W = reinterpret_cast<float2 *>(p);
  p += round256(wb);
  cb.Kt1 = reinterpret_cast<float2 *>(p);
  p += round256(kb);
  cb.Kt2 = reinterpret_cast<float2 *>(p);
  p += round256(kb);
  cb.Kh = reinterpret_cast<float2 *>(p);
  p += round256(kb);
  cb.tw = reinterpret_cast<float2 *>(p);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // every buffer is written before it is read within one call (the twiddles
  // by their kernel, the spectra and lines by the passes), so nothing needs
  // clearing between calls
  SGT_TRY(init_twiddles(s, cb.tw));
  SGT_TRY(launch_conv(s, g, zdirect, V, Phi, cb));
  return SGTSNE_OK;
}

}  // extern "C"
