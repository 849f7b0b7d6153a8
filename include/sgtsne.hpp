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
// ELL / COO fields of the reference are built on the device here.
template <typename dataPoint>
struct sparse_matrix {
  int m = 0;
  int n = 0;
  int nnz = 0;
  int format = 0;
  matidx *row = nullptr;
  matidx *col = nullptr;
  dataPoint *val = nullptr;
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

// HY/sgtsne.cpp:37-197.  The column-stochastic CSC of P is the row-stochastic
// CSR of P^T, and P + P^T is the same matrix: it is passed as CSR unchanged.
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
  std::vector<int32_t> rp((size_t)n + 1), ci((size_t)P.nnz);
  std::vector<float> v((size_t)P.nnz);
  for (int i = 0; i <= n; i++) rp[i] = (int32_t)P.col[i];
  for (int t = 0; t < P.nnz; t++) {
    ci[t] = (int32_t)P.row[t];
    v[t] = (float)P.val[t];
  }
  std::vector<float> yi, yo((size_t)n * params.d);
  if (y_in) yi.assign(y_in, y_in + (size_t)n * params.d);
  double *ti = (timeInfo && *timeInfo) ? *timeInfo : nullptr;
  sgtsne_mi355x::check(sgtsne_run(n, P.nnz, rp.data(), ci.data(), v.data(), &c, y_in ? yi.data() : nullptr,
                                  yo.data(), ti),
                       "sgtsne");
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
