// gather_probe.hip -- the rate at which this GPU serves scattered 8-byte reads
// of a point array the size of the attractive term's Y (1.3M float2 points,
// 10.4 MB), for index streams of chosen locality.  It is the ceiling the
// attractive SpMV's neighbour gathers run against: each nonzero reads y_j at a
// column index, and the memory system serves those reads at a rate set by
// where the lines come from (L1, L2 or MALL/HBM), not by HBM bandwidth.
//
// Index streams (147 per row, the mean row length of P + P^T at k = 90):
// uniform over all points, or uniform within +-W rows of the row (clipped),
// optionally increasing within each row like a CSR row.  Two kernels read
// them: a flat one (U reads in flight per thread) and the CSR form of the
// attractive kernel (G lanes per row, masked U-entry steps, the pair term).
// Output: one line per (stream, kernel) with ms per pass and gathers/s.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define PROBE_CHECK(x)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                     \
    }                                                                                   \
  } while (0)

namespace {

__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// col[r L + k] for row r: uniform over [0, n) (w == 0) or over [r - w, r + w)
// clipped to [0, n); with sorted != 0 the k-th index of a row lies in the k-th
// of L equal slices of that range, so a row's indices increase.
__global__ void make_indices(int n, int L, int w, int sorted, int *col, float *val) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * L) return;
  const int r = (int)(i / L), k = (int)(i % L);
  const unsigned h = mix32((unsigned)i * 2654435761u + 12345u);
  int lo = 0, hi = n;
  if (w > 0) {
    lo = r - w < 0 ? 0 : r - w;
    hi = r + w > n ? n : r + w;
  }
  const int span = hi - lo;
  int c;
  if (sorted) {
    const int seg = span / L > 0 ? span / L : 1;
    c = lo + (int)(((long long)k * span) / L) + (int)(h % (unsigned)seg);
    if (c >= hi) c = hi - 1;
  } else {
    c = lo + (int)(h % (unsigned)span);
  }
  col[i] = c;
  val[i] = 1.0f / (float)L;
}

template <int U>
__global__ __launch_bounds__(256) void flat_gather(long long tot, const int *__restrict__ col,
                                                   const float2 *__restrict__ Y, float *out) {
  const long long S = (long long)gridDim.x * blockDim.x;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float acc = 0.f;
  for (; t + (U - 1) * S < tot; t += U * S) {
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(col + t + u * S);
    float2 y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) y[u] = Y[c[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += y[u].x * y[u].y;
  }
  for (; t < tot; t += S) {
    const float2 y = Y[col[t]];
    acc += y.x * y.y;
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int G, int U>
__global__ __launch_bounds__(256) void csr_gather(int n, int L, const int *__restrict__ col,
                                                  const float *__restrict__ val, const float2 *__restrict__ Y,
                                                  float2 *F) {
  const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  const int ngroups = gridDim.x * blockDim.x / G;
  for (int row = gid; row < n; row += ngroups) {
    const float2 yi = Y[row];
    float ax = 0.f, ay = 0.f;
    const int e = (row + 1) * L;
    for (int t = row * L + lane; t < e; t += U * G) {
      int c[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool in = t + u * G < e;
        c[u] = in ? __builtin_nontemporal_load(col + t + u * G) : -1;
        v[u] = in ? __builtin_nontemporal_load(val + t + u * G) : 0.f;
      }
      float2 yj[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c[u] >= 0) yj[u] = Y[c[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c[u] >= 0) {
          const float dx = yi.x - yj[u].x, dy = yi.y - yj[u].y;
          const float q = v[u] / (1.f + dx * dx + dy * dy);
          ax += q * dx;
          ay += q * dy;
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      ax += __shfl_xor(ax, o, G);
      ay += __shfl_xor(ay, o, G);
    }
    if (lane == 0) F[row] = make_float2(ax, ay);
  }
}

}  // namespace

int main() {
  const int n = 1300000, L = 147;
  const long long tot = (long long)n * L;
  int *col = nullptr;
  float *val = nullptr, *out = nullptr;
  float2 *Y = nullptr, *F = nullptr;
  const int flat_threads = 256 * 4096;
  PROBE_CHECK(hipMalloc(&col, tot * sizeof(int)));
  PROBE_CHECK(hipMalloc(&val, tot * sizeof(float)));
  PROBE_CHECK(hipMalloc(&Y, (size_t)n * sizeof(float2)));
  PROBE_CHECK(hipMalloc(&F, (size_t)n * sizeof(float2)));
  PROBE_CHECK(hipMalloc(&out, (size_t)flat_threads * sizeof(float)));
  std::vector<float2> hy(n);
  for (int i = 0; i < n; ++i) hy[i] = make_float2((float)(i % 977) * 1e-3f, (float)(i % 613) * 1e-3f);
  PROBE_CHECK(hipMemcpy(Y, hy.data(), (size_t)n * sizeof(float2), hipMemcpyHostToDevice));
  hipEvent_t a, b;
  PROBE_CHECK(hipEventCreate(&a));
  PROBE_CHECK(hipEventCreate(&b));
  struct Stream { int w, sorted; const char *name; };
  const Stream streams[] = {{0, 0, "uniform"},          {0, 1, "uniform, row-sorted"},
                            {65536, 0, "within +-65536"}, {8192, 0, "within +-8192"},
                            {8192, 1, "within +-8192, row-sorted"}, {1024, 1, "within +-1024, row-sorted"}};
  const int csr_blocks = 65536;  // 256-thread blocks; rows are grid-strided
  for (const Stream &st : streams) {
    make_indices<<<(unsigned)((tot + 255) / 256), 256>>>(n, L, st.w, st.sorted, col, val);
    PROBE_CHECK(hipGetLastError());
    PROBE_CHECK(hipDeviceSynchronize());
    auto timeit = [&](const char *kname, auto launch) {
      for (int r = 0; r < 3; ++r) launch();
      PROBE_CHECK(hipDeviceSynchronize());
      const int reps = 10;
      PROBE_CHECK(hipEventRecord(a));
      for (int r = 0; r < reps; ++r) launch();
      PROBE_CHECK(hipEventRecord(b));
      PROBE_CHECK(hipEventSynchronize(b));
      float ms = 0.f;
      PROBE_CHECK(hipEventElapsedTime(&ms, a, b));
      ms /= reps;
      std::printf("%-28s %-24s %8.3f ms  %6.1f G gathers/s\n", st.name, kname, ms, tot / (ms * 1e6));
      std::fflush(stdout);
    };
    timeit("flat, 8 in flight", [&] { flat_gather<8><<<4096, 256>>>(tot, col, Y, out); });
    timeit("flat, 16 in flight", [&] { flat_gather<16><<<4096, 256>>>(tot, col, Y, out); });
    timeit("CSR, 32 lanes x 16", [&] { csr_gather<32, 16><<<csr_blocks, 256>>>(n, L, col, val, Y, F); });
    timeit("CSR, 16 lanes x 16", [&] { csr_gather<16, 16><<<csr_blocks / 2, 256>>>(n, L, col, val, Y, F); });
    timeit("CSR, 32 lanes x 8", [&] { csr_gather<32, 8><<<csr_blocks / 2, 256>>>(n, L, col, val, Y, F); });
    PROBE_CHECK(hipGetLastError());
  }
  PROBE_CHECK(hipDeviceSynchronize());
  std::printf("done\n");
  return 0;
}
