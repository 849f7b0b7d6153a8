/*
 * sgtsne.hpp -- the reference's C++ entry points, header-only, over the C-ABI
 * of include/sgtsne.h.  A caller of the reference replaces
 *   #include "sgtsne.cuh"            (CU = code/cuda/SG-tSNE-CUDA/src)
 *   #include "sgtsne.hpp"            (HY = code/cuda/SG-tSNE-HYB/sgtsnepi/src)
 * with this header and links libsgtsne_hip.so.
 *
 * Provided, with the reference's names and argument meaning:
 *   tsneparams, sparse_matrix<T>     CU/types.hpp:27-72
 *   sgtsneCUDA<T>(P, params, y_in, timeInfo)        CU/sgtsne.cuh:13-18
 *   perplexityEqualization<T>(I, D, n, nn, u)        CU/perplexity_search.cuh:18-20
 *   sgtsne_hyb::sgtsne(P, params, y_in, timeInfo)    HY/sgtsne.hpp:68-72 (CSC, FP64, AoS)
 *   sgtsne_hyb::perplexityEqualization(I, D, n, nn, u) HY/sgtsne.hpp:85-90
 *
 * Differences a caller sees: P and the returned arrays live in host memory
 * (the CUDA path keeps P in managed memory); errors end the process with the
 * message on stderr, as the reference's gpuErrchk does (CU/common.cuh:15-22),
 * unless SGTSNE_THROW_ON_ERROR is defined, in which case std::runtime_error is
 * thrown.  Define SGTSNE_NO_REFERENCE_NAMES to keep the global names out.
 */
#ifndef SGTSNE_HPP_MI355X
#define SGTSNE_HPP_MI355X

#include "sgtsne.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace sgtsne_mi355x {

inline void fail(int status, const char *where) {
  std::string msg = std::string(where) + ": " + sgtsne_last_error() + " (status " + std::to_string(status) + ")";
#ifdef SGTSNE_THROW_ON_ERROR
  throw std::runtime_error(msg);
#else
  std::fprintf(stderr, "sgtsne error: %s\n", msg.c_str());
  std::exit(status == 0 ? 1 : (status < 0 ? -status : status));
#endif
}

inline void check(int status, const char *where) {
  if (status != SGTSNE_OK) fail(status, where);
}

}  // namespace sgtsne_mi355x

#ifndef SGTSNE_NO_REFERENCE_NAMES

typedef float coord;   // CU/types.hpp:18-20
typedef float matval;
typedef int matidx;

// CU/types.hpp:27-47: CSR (row = offsets [n+1], col = indices [nnz]).  The
// block, ELL and COO fields keep the reference's names so that callers that
// set or read them compile unchanged; this engine builds its own formats on
// the device from the CSR (sgtsneCUDA reads only m, n, nnz, row, col, val),
// so they stay zero / null unless the caller fills them.
template <typename dataPoint>
struct sparse_matrix {
  int m = 0;       // rows
  int n = 0;       // columns
  int nnz = 0;     // nonzeros
  int format = 0;  // 0 CSR, 1 BSR (never implemented by the reference), 2 COO, 3 hybrid ELL + COO
  matidx *row = nullptr;
  matidx *col = nullptr;
  dataPoint *val = nullptr;

  int nnzb = 0;
  int blockSize = 0;
  int blockRows = 0;

  unsigned int *ell_cols = nullptr, *coo_col_ids = nullptr, *coo_row_ids = nullptr;
  dataPoint *ell_data = nullptr, *coo_data = nullptr;
  int coo_size = 0, elements_in_rows = 0;
};

// CU/types.hpp:55-72, same fields and defaults.
typedef struct {
  int d = 2;
  float lambda = 1;
  float alpha = 12;
  int maxIter = 1000;
  int earlyIter = 250;
  int n = 0;
  float h = -1;
  bool dropLeaf = false;
  int np = 0;
  int format = 2;
  int bs = 0;
  int ComputeError = -1;
  int ng = 50;
  int sim = 0;
} tsneparams;

namespace sgtsne_mi355x {
inline sgtsne_params to_c(const tsneparams &p) {
  sgtsne_params c;
  sgtsne_params_default(&c);
  c.d = p.d;
  c.lambda = p.lambda;
  c.alpha = p.alpha;
  c.max_iter = p.maxIter;
  c.early_iter = p.earlyIter;
  c.n = p.n;
  c.h = p.h;
  c.drop_leaf = p.dropLeaf ? 1 : 0;
  c.np = p.np;
  c.format = p.format;
  c.bs = p.bs;
  c.compute_error = p.ComputeError;
  c.ng = p.ng;
  c.layout = SGTSNE_LAYOUT_SOA;  // the CUDA path keeps Y as [d][N]
  return c;
}
}  // namespace sgtsne_mi355x

// CU/sgtsne.cu:26-128: P is the row-stochastic kNN CSR; y_in is [d][N] or
// NULL; returns a malloc'ed [d][N] embedding (the caller frees it).
// timeInfo receives 7 doubles per iteration and the loop time at 7*maxIter.
template <class dataPoint>
dataPoint *sgtsneCUDA(sparse_matrix<dataPoint> *P, tsneparams params, dataPoint *y_in, double *timeInfo) {
  static_assert(sizeof(dataPoint) == sizeof(float), "only the float instantiation exists (CU/sgtsne.cu:130-131)");
  sgtsne_params c = sgtsne_mi355x::to_c(params);
  const int n = P->n;
  if (c.n == 0) c.n = n;
  dataPoint *y = static_cast<dataPoint *>(std::malloc(sizeof(dataPoint) * (size_t)n * params.d));
  if (!y) sgtsne_mi355x::fail(SGTSNE_ERR_NOMEM, "sgtsneCUDA");
  const int st = sgtsne_run(n, (int64_t)P->nnz, P->row, P->col, reinterpret_cast<const float *>(P->val), &c,
                            reinterpret_cast<const float *>(y_in), reinterpret_cast<float *>(y), timeInfo);
  sgtsne_mi355x::check(st, "sgtsneCUDA");
  return y;
}

// CU/perplexity_search.cu:114-141: I and D are [n][nn+1] with the point itself
// first; returns the row-stochastic kNN CSR (host arrays, malloc'ed).
template <class dataPoint>
sparse_matrix<dataPoint> perplexityEqualization(int *I, dataPoint *D, int n, int nn, dataPoint u) {
  std::vector<double> Dd((size_t)n * (nn + 1));
  for (size_t t = 0; t < Dd.size(); t++) Dd[t] = (double)D[t];
  sparse_matrix<dataPoint> P;
  P.m = P.n = n;
  P.nnz = n * nn;
  P.row = static_cast<matidx *>(std::malloc(sizeof(matidx) * ((size_t)n + 1)));
  P.col = static_cast<matidx *>(std::malloc(sizeof(matidx) * (size_t)n * nn));
  std::vector<float> v((size_t)n * nn);
  P.val = static_cast<dataPoint *>(std::malloc(sizeof(dataPoint) * (size_t)n * nn));
  if (!P.row || !P.col || !P.val) sgtsne_mi355x::fail(SGTSNE_ERR_NOMEM, "perplexityEqualization");
  sgtsne_mi355x::check(sgtsne_perplexity_equalize(I, Dd.data(), n, nn, (double)u, P.row, P.col, v.data()),
                       "perplexityEqualization");
  for (size_t t = 0; t < v.size(); t++) P.val[t] = (dataPoint)v[t];
  return P;
}

// The HYB / SG-t-SNE-Pi interface: FP64, AoS [N][d], CSC with uint32 indices.
namespace sgtsne_hyb {

// HY/sgtsne.hpp:25-37
typedef struct {
  int d = 2;
  double lambda = 1;
  double alpha = 12;
  int maxIter = 1000;
  int earlyIter = 250;
  int n = 0;
  double h = -1;
  bool dropLeaf = false;
  int np = 0;
} tsneparams;

// HY/sgtsne.hpp:44-54 (CSC: col = offsets [n+1], row = indices [nnz])
typedef struct {
  int m;
  int n;
  int nnz;
  uint32_t *row;
  uint32_t *col;
  double *val;
} sparse_matrix;

// HY/sparsematrix.cpp:27-67 (makeStochastic): every column j whose sum is
// nonzero (or whose first stored value is nonzero) and differs from 1 by more
// than 1e-12 is divided by its sum, in FP64, in place.  Returns the number of
// columns that were already stochastic.
inline uint32_t makeStochastic(sparse_matrix P) {
  uint32_t nStoch = 0;
  for (int j = 0; j < P.n; j++) {
    const uint32_t t0 = P.col[j], t1 = P.col[j + 1];
    double sum = 0;
    for (uint32_t t = t0; t < t1; t++) sum += P.val[t];
    const bool first_nonzero = t0 < t1 && P.val[t0] != 0.0;
    if (sum != 0.0 || first_nonzero) {
      if (std::abs(sum - 1) > 1e-12)
        for (uint32_t t = t0; t < t1; t++) P.val[t] /= sum;
      else
        nStoch++;
    }
  }
  return nStoch;
}

// HY/sgtsne.cpp:37-197.  P is first made column stochastic in place, as the
// reference does (HY/sgtsne.cpp:60); the column-stochastic CSC of P is the
// row-stochastic CSR of P^T, and P + P^T is the same matrix, so it is passed
// as CSR unchanged.  The engine computes in FP32 (P and Y are rounded once).
// *timeInfo, when given, receives HYB's layout: 8 doubles per iteration in
// seconds (0 attractive, 1 spread, 2 grid convolution, 3 gather, 4 forces and
// Z, 5 centring and bounding box, 6-7 zero: no CPU/GPU overlap to wait for)
// and the loop time in seconds at 8*maxIter (HY/gradient_descendHYB.cpp:
// 121-142,210,248).
inline double *sgtsne(sparse_matrix P, tsneparams params, double *y_in = nullptr, double **timeInfo = nullptr) {
  sgtsne_params c;
  sgtsne_params_default(&c);
  c.d = params.d;
  c.lambda = params.lambda;
  c.alpha = params.alpha;
  c.max_iter = params.maxIter;
  c.early_iter = params.earlyIter;
  c.n = P.n;
  c.h = params.h;
  c.drop_leaf = params.dropLeaf ? 1 : 0;
  c.np = params.np;
  c.layout = SGTSNE_LAYOUT_AOS;
  const int n = P.n;
  makeStochastic(P);
  std::vector<int32_t> rp((size_t)n + 1), ci((size_t)P.nnz);
  std::vector<float> v((size_t)P.nnz);
  for (int i = 0; i <= n; i++) rp[i] = (int32_t)P.col[i];
  for (int t = 0; t < P.nnz; t++) {
    ci[t] = (int32_t)P.row[t];
    v[t] = (float)P.val[t];
  }
  std::vector<float> yi, yo((size_t)n * params.d);
  if (y_in) yi.assign(y_in, y_in + (size_t)n * params.d);
  const int iters = params.maxIter > 0 ? params.maxIter : 0;
  std::vector<double> t7;
  if (timeInfo && *timeInfo) t7.assign((size_t)7 * iters + 1, 0.0);
  sgtsne_mi355x::check(sgtsne_run(n, P.nnz, rp.data(), ci.data(), v.data(), &c, y_in ? yi.data() : nullptr,
                                  yo.data(), t7.empty() ? nullptr : t7.data()),
                       "sgtsne");
  if (!t7.empty()) {
    // the C-ABI's CUDA layout (ms: 0 attractive, 1 spread, 2 convolution,
    // 3 gather, 4 forces and Z, 5 repulsion total, 6 preprocessing)
    double *ti = *timeInfo;
    for (int it = 0; it < iters; it++) {
      const double *s7 = &t7[(size_t)7 * it];
      double *s8 = ti + (size_t)8 * it;
      s8[0] = s7[0] / 1e3;
      s8[1] = s7[1] / 1e3;
      s8[2] = s7[2] / 1e3;
      s8[3] = s7[3] / 1e3;
      s8[4] = s7[4] / 1e3;
      s8[5] = s7[6] / 1e3;
      s8[6] = 0.0;
      s8[7] = 0.0;
    }
    ti[(size_t)8 * iters] = t7[(size_t)7 * iters] / 1e3;
  }
  double *y = static_cast<double *>(std::malloc(sizeof(double) * yo.size()));
  if (!y) sgtsne_mi355x::fail(SGTSNE_ERR_NOMEM, "sgtsne");
  for (size_t t = 0; t < yo.size(); t++) y[t] = yo[t];
  return y;
}

// HY/sgtsne.cpp:257-297: returns the column-stochastic CSC kNN graph.
inline sparse_matrix perplexityEqualization(int *I, double *D, int n, int nn, double u) {
  std::vector<int32_t> rp((size_t)n + 1), ci((size_t)n * nn);
  std::vector<float> v((size_t)n * nn);
  sgtsne_mi355x::check(sgtsne_perplexity_equalize(I, D, n, nn, u, rp.data(), ci.data(), v.data()),
                       "perplexityEqualization");
  sparse_matrix P;
  P.m = P.n = n;
  P.nnz = n * nn;
  P.col = static_cast<uint32_t *>(std::malloc(sizeof(uint32_t) * ((size_t)n + 1)));
  P.row = static_cast<uint32_t *>(std::malloc(sizeof(uint32_t) * (size_t)n * nn));
  P.val = static_cast<double *>(std::malloc(sizeof(double) * (size_t)n * nn));
  if (!P.col || !P.row || !P.val) sgtsne_mi355x::fail(SGTSNE_ERR_NOMEM, "perplexityEqualization");
  for (int i = 0; i <= n; i++) P.col[i] = (uint32_t)rp[i];
  for (size_t t = 0; t < ci.size(); t++) {
    P.row[t] = (uint32_t)ci[t];
    P.val[t] = v[t];
  }
  return P;
}

}  // namespace sgtsne_hyb

#endif  // SGTSNE_NO_REFERENCE_NAMES

#endif  // SGTSNE_HPP_MI355X
