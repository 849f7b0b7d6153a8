/*
 * sgtsne_oracle.c -- FP64 CPU restatement of the SG-t-SNE gradient path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU
 * baseline.  Never linked into, loaded by, or called from the product path.
 * Each function cites the reference file:line whose algorithm it restates
 * (HY = code/cuda/SG-tSNE-HYB/sgtsnepi/src, CU = code/cuda/SG-tSNE-CUDA/src).
 *
 * Parity pins (tests/test_oracle.py): closed forms for the Lagrange weights
 * and the perplexity bisection, brute force for kNN / symmetrisation /
 * attractive term, the direct Toeplitz sum for the FFT convolution, numpy.fft
 * for or_fft, the exact O(N^2) repulsion with the reference's own acceptance
 * criteria (HY/test_modules.cpp:122-156), the S2G/G2S adjoint identity and
 * the Z identity.
 */
#include "sgtsne_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void or_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Grid-size table: HY/qq.cpp:108-130.  The list of FFT-friendly sizes is a
 * data table that must be identical for parity of the interpolation error. */
static const int kGridSizes[] = {
    8,   9,   10,  11,  12,  13,  14,  15,  16,  20,  25,  26,  28,  32,  33,  35,  36,
    39,  40,  42,  44,  45,  48,  49,  50,  52,  54,  55,  56,  60,  63,  64,  65,  66,
    70,  72,  75,  77,  78,  80,  84,  88,  90,  91,  96,  98,  99,  100, 104, 105, 108,
    110, 112, 117, 120, 125, 126, 130, 132, 135, 140, 144, 147, 150, 154, 156, 160, 165,
    168, 175, 176, 180, 182, 189, 192, 195, 196, 198, 200, 208, 210, 216, 220, 224, 225,
    231, 234, 240, 245, 250, 252, 260, 264, 270, 273, 275, 280, 288, 294, 297, 300, 308,
    312, 315, 320, 325, 330, 336, 343, 350, 351, 352, 360, 364, 375, 378, 385, 390, 392,
    396, 400, 416, 420, 432, 440, 441, 448, 450, 455, 462, 468, 480, 490, 495, 500, 504,
    512};

int or_best_grid_size(int nGrid) {
  const int cnt = (int)(sizeof(kGridSizes) / sizeof(kGridSizes[0]));
  for (int i = 0; i < cnt; i++)
    if (nGrid + 2 <= kGridSizes[i]) return kGridSizes[i] - 2;
  return kGridSizes[cnt - 1] - 2;
}

/* ------------------------------------------------------------------------ */
/* Exact kNN.  Ordering key: (squared distance, index).                      */
static inline int key_less(double da, int ia, double db, int ib) {
  return da < db || (da == db && ia < ib);
}

/* One exact row: the k nearest (squared distance, index) keys of point i
 * among all others, ascending, written after the point itself. */
static void knn_row(const float *X, int n, int dim, int k, int i, double *hd, int *hi, int32_t *Ii, double *Di) {
  const float *xi = X + (size_t)i * dim;
  int cnt = 0; /* max-heap of the k best (worst at root) */
  for (int j = 0; j < n; j++) {
    if (j == i) continue;
    const float *xj = X + (size_t)j * dim;
    double acc = 0.0;
    for (int l = 0; l < dim; l++) {
      double df = (double)xi[l] - (double)xj[l];
      acc = fma(df, df, acc);
    }
    if (cnt < k) {
      int c = cnt++;
      hd[c] = acc; hi[c] = j;
      while (c > 0) {
        int p = (c - 1) / 2;
        if (key_less(hd[p], hi[p], hd[c], hi[c])) {
          double td = hd[p]; hd[p] = hd[c]; hd[c] = td;
          int ti = hi[p]; hi[p] = hi[c]; hi[c] = ti;
          c = p;
        } else break;
      }
    } else if (k > 0 && key_less(acc, j, hd[0], hi[0])) {
      hd[0] = acc; hi[0] = j;
      int c = 0;
      for (;;) {
        int l = 2 * c + 1, r = l + 1, m = c;
        if (l < k && key_less(hd[m], hi[m], hd[l], hi[l])) m = l;
        if (r < k && key_less(hd[m], hi[m], hd[r], hi[r])) m = r;
        if (m == c) break;
        double td = hd[m]; hd[m] = hd[c]; hd[c] = td;
        int ti = hi[m]; hi[m] = hi[c]; hi[c] = ti;
        c = m;
      }
    }
  }
  /* heap-sort ascending into the output row (self first) */
  Ii[0] = i; Di[0] = 0.0;
  for (int e = cnt - 1; e >= 0; e--) {
    Ii[1 + e] = hi[0]; Di[1 + e] = hd[0];
    hd[0] = hd[e]; hi[0] = hi[e];
    int c = 0, sz = e;
    for (;;) {
      int l = 2 * c + 1, r = l + 1, m = c;
      if (l < sz && key_less(hd[m], hi[m], hd[l], hi[l])) m = l;
      if (r < sz && key_less(hd[m], hi[m], hd[r], hi[r])) m = r;
      if (m == c) break;
      double td = hd[m]; hd[m] = hd[c]; hd[c] = td;
      int ti = hi[m]; hi[m] = hi[c]; hi[c] = ti;
      c = m;
    }
  }
}

void or_knn_exact(const float *X, int n, int dim, int k, int32_t *I, double *D) {
#pragma omp parallel
  {
    double *hd = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    int *hi = (int *)malloc(sizeof(int) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < n; i++)
      knn_row(X, n, dim, k, i, hd, hi, I + (size_t)i * (k + 1), D + (size_t)i * (k + 1));
    free(hd); free(hi);
  }
}

void or_knn_rows(const float *X, int n, int dim, int k, const int32_t *rows, int nrows, int32_t *I, double *D) {
#pragma omp parallel
  {
    double *hd = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    int *hi = (int *)malloc(sizeof(int) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < nrows; r++)
      knn_row(X, n, dim, k, rows[r], hd, hi, I + (size_t)r * (k + 1), D + (size_t)r * (k + 1));
    free(hd); free(hi);
  }
}

/* ------------------------------------------------------------------------ */
/* Perplexity bisection: HY/sgtsne.cpp:201-254.                              */
int or_equalize_vertex(double *val, const double *dist, double u, int k, double *beta_out) {
  int found = 0;
  double beta = 1.0;
  double min_beta = -DBL_MAX, max_beta = DBL_MAX;
  const double tol = 1e-5;
  int iter = 0;
  double sum_P = 0.0;
  while (!found && iter < 200) {
    for (int m = 0; m < k; m++) val[m] = exp(-beta * dist[m + 1]);
    sum_P = DBL_MIN;
    for (int m = 0; m < k; m++) sum_P += val[m];
    double H = 0.0;
    for (int m = 0; m < k; m++) H += beta * (dist[m + 1] * val[m]);
    H = (H / sum_P) + log(sum_P);
    double Hdiff = H - log(u);
    if (Hdiff < tol && -Hdiff < tol) {
      found = 1;
    } else if (Hdiff > 0) {
      min_beta = beta;
      if (max_beta == DBL_MAX || max_beta == -DBL_MAX) beta *= 2.0;
      else beta = (beta + max_beta) / 2.0;
    } else {
      max_beta = beta;
      if (min_beta == -DBL_MAX || min_beta == DBL_MAX) beta /= 2.0;
      else beta = (beta + min_beta) / 2.0;
    }
    iter++;
  }
  for (int m = 0; m < k; m++) val[m] /= sum_P;
  if (beta_out) *beta_out = beta;
  return iter;
}

void or_perplexity_equalize(const int32_t *I, const double *D, int n, int k, double u,
                            int64_t *rowptr, int32_t *col, double *val) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++)
    or_equalize_vertex(val + (size_t)i * k, D + (size_t)i * (k + 1), u, k, NULL);
  for (int i = 0; i < n; i++) {
    rowptr[i] = (int64_t)i * k;
    for (int e = 0; e < k; e++) col[(size_t)i * k + e] = I[(size_t)i * (k + 1) + e + 1];
  }
  rowptr[n] = (int64_t)n * k;
}

/* HY/sparsematrix.cpp:27-67 */
int or_make_stochastic(int n, const int64_t *rowptr, double *val) {
  int nst = 0;
#pragma omp parallel for reduction(+ : nst)
  for (int j = 0; j < n; j++) {
    double sum = 0.0;
    for (int64_t t = rowptr[j]; t < rowptr[j + 1]; t++) sum += val[t];
    if (sum != 0.0 || (rowptr[j + 1] > rowptr[j] && val[rowptr[j]] != 0.0)) {
      if (fabs(sum - 1.0) > 1e-12)
        for (int64_t t = rowptr[j]; t < rowptr[j + 1]; t++) val[t] /= sum;
      else
        nst++;
    }
  }
  return nst;
}

/* HY/graph_rescaling.cpp:19-104 (dist = false branch) */
void or_lambda_rescale(int n, const int64_t *rowptr, double *val, double lambda, int drop_leaf) {
  const double tolBinary = 1e-5;
  const int maxIter = 100;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    double fval = 1.0 - lambda;
    double sig2 = 1.0, a = -1e3, c = INFINITY;
    int iter = 0;
    double sum_i = 0.0;
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) val[j] = -log(val[j]);
    while (fabs(fval) > tolBinary && iter < maxIter) {
      iter++;
      sum_i = 0.0;
      if (fval > 0) {
        a = sig2;
        sig2 = isinf(c) ? 2 * a : 0.5 * (a + c);
      } else {
        c = sig2;
        sig2 = 0.5 * (a + c);
      }
      for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) sum_i += exp(-val[j] * sig2);
      fval = sum_i - lambda;
    }
    sum_i = 0.0;
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
      val[j] = exp(-val[j] * sig2);
      sum_i += val[j];
    }
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) val[j] /= sum_i;
    if (drop_leaf && rowptr[i + 1] - rowptr[i] == 1) val[rowptr[i]] = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Symmetrisation: P + P^T (HY/sparsematrix.cpp:69-164), then divide by the
 * total sum (HY/sgtsne.cpp:78-85).                                          */
typedef struct { int32_t c; double v; } colval_t;
static int cmp_colval(const void *a, const void *b) {
  int32_t ca = ((const colval_t *)a)->c, cb = ((const colval_t *)b)->c;
  return (ca > cb) - (ca < cb);
}

int64_t or_symmetrize(int n, const int64_t *rowptr, const int32_t *col, const double *val,
                      int64_t *rowptr_out, int32_t *col_out, double *val_out) {
  /* gather both directions per row */
  int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  for (int i = 0; i < n; i++)
    for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
      cnt[i + 1]++;
      cnt[col[t] + 1]++;
    }
  for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  int64_t tot = cnt[n];
  colval_t *buf = (colval_t *)malloc(sizeof(colval_t) * (size_t)(tot > 0 ? tot : 1));
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) pos[i] = cnt[i];
  for (int i = 0; i < n; i++)
    for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
      colval_t a = {col[t], val[t]}, b = {i, val[t]};
      buf[pos[i]++] = a;
      buf[pos[col[t]]++] = b;
    }
  /* sort each row, merge duplicates */
  int64_t *rl = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    colval_t *r = buf + cnt[i];
    int64_t len = cnt[i + 1] - cnt[i];
    qsort(r, (size_t)len, sizeof(colval_t), cmp_colval);
    int64_t w = 0;
    for (int64_t e = 0; e < len; e++) {
      if (w > 0 && r[w - 1].c == r[e].c) r[w - 1].v += r[e].v;
      else r[w++] = r[e];
    }
    rl[i + 1] = w;
  }
  for (int i = 0; i < n; i++) rl[i + 1] += rl[i];
  int64_t nnz = rl[n];
  if (col_out != NULL) {
    double sum = 0.0;
    for (int i = 0; i < n; i++)
      for (int64_t e = 0; e < rl[i + 1] - rl[i]; e++) sum += buf[cnt[i] + e].v;
    for (int i = 0; i <= n; i++) rowptr_out[i] = rl[i];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
      for (int64_t e = 0; e < rl[i + 1] - rl[i]; e++) {
        col_out[rl[i] + e] = buf[cnt[i] + e].c;
        val_out[rl[i] + e] = buf[cnt[i] + e].v / sum;
      }
  }
  free(cnt); free(buf); free(pos); free(rl);
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* Attractive term: HY/pq.cpp:15-57 (row-parallel form; the reference's CSB
 * kernel is tested equal to it at 1e-10 in HY/test_modules.cpp:76-120).     */
void or_attractive(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
                   const double *Y, double *Fattr) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    double acc[3] = {0, 0, 0};
    const double *yi = Y + (size_t)i * d;
    for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
      const double *yj = Y + (size_t)col[t] * d;
      double dist = 0.0;
      for (int a = 0; a < d; a++) dist += (yi[a] - yj[a]) * (yi[a] - yj[a]);
      double pq = val[t] / (1.0 + dist);
      for (int a = 0; a < d; a++) acc[a] += pq * (yi[a] - yj[a]);
    }
    for (int a = 0; a < d; a++) Fattr[(size_t)i * d + a] = acc[a];
  }
}

/* Exact repulsion: HY/qq.cpp:25-70 */
double or_repulsive_exact(int n, int d, const double *Y, double *Frep) {
  double *zv = (double *)calloc((size_t)n, sizeof(double));
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    double f[3] = {0, 0, 0}, z = 0.0;
    const double *yi = Y + (size_t)i * d;
    for (int j = 0; j < n; j++) {
      if (i == j) continue;
      const double *yj = Y + (size_t)j * d;
      double dist = 0.0;
      for (int a = 0; a < d; a++) dist += (yj[a] - yi[a]) * (yj[a] - yi[a]);
      double w = 1.0 / (1.0 + dist);
      for (int a = 0; a < d; a++) f[a] += (yi[a] - yj[a]) * w * w;
      z += w;
    }
    for (int a = 0; a < d; a++) Frep[(size_t)i * d + a] = f[a];
    zv[i] = z;
  }
  double Z = 0.0;
  for (int i = 0; i < n; i++) Z += zv[i];
  for (size_t t = 0; t < (size_t)n * d; t++) Frep[t] /= Z;
  free(zv);
  return Z;
}

void or_phi_exact(int n, int d, const double *Y, double *Phi) {
  const int m = d + 1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    double acc[4] = {0, 0, 0, 0};
    const double *yi = Y + (size_t)i * d;
    for (int j = 0; j < n; j++) {
      const double *yj = Y + (size_t)j * d;
      double dist = 0.0;
      for (int a = 0; a < d; a++) dist += (yj[a] - yi[a]) * (yj[a] - yi[a]);
      double w = 1.0 / (1.0 + dist);
      w = w * w;
      acc[0] += w;
      for (int a = 0; a < d; a++) acc[a + 1] += w * yj[a];
    }
    for (int c = 0; c < m; c++) Phi[(size_t)i * m + c] = acc[c];
  }
}

/* ------------------------------------------------------------------------ */
/* Lagrange cubic weights: HY/gridding.cpp:28-40 (LAGRANGE_INTERPOLATION).   */
static inline double lg1(double t) { return 0.5 * t * t * t - 1.0 * t * t - 0.5 * t + 1.0; }
static inline double lg2(double t) {
  const double cc = 1.0 / 6.0;
  return -cc * t * t * t + 1.0 * t * t - 11.0 * cc * t + 1.0;
}

void or_lagrange_weights(double delta, double *w) {
  w[0] = lg2(1.0 + delta);
  w[1] = lg1(delta);
  w[2] = lg1(1.0 - delta);
  w[3] = lg2(2.0 - delta);
}

static inline int stencil_setup(int d, const double *y, int *f, double w[3][4]) {
  for (int a = 0; a < d; a++) {
    f[a] = (int)floor(y[a]);
    or_lagrange_weights(y[a] - (double)f[a], w[a]);
  }
  return 0;
}

static void s2g_range(int i0, int i1, int d, int ng, int nVec, const double *y, const double *q,
                      double *V) {
  const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
  for (int i = i0; i < i1; i++) {
    int f[3] = {0, 0, 0};
    double w[3][4];
    stencil_setup(d, y + (size_t)i * d, f, w);
    const double *qi = q + (size_t)i * nVec;
    if (d == 1) {
      for (int j = 0; j < nVec; j++)
        for (int a = 0; a < 4; a++) V[(size_t)j * G + f[0] + a] += qi[j] * w[0][a];
    } else if (d == 2) {
      for (int j = 0; j < nVec; j++)
        for (int b = 0; b < 4; b++) {
          double qv2 = qi[j] * w[1][b];
          for (int a = 0; a < 4; a++)
            V[(size_t)j * G + (size_t)(f[1] + b) * ng + f[0] + a] += qv2 * w[0][a];
        }
    } else {
      for (int j = 0; j < nVec; j++)
        for (int c = 0; c < 4; c++)
          for (int b = 0; b < 4; b++) {
            double qv = qi[j] * w[1][b] * w[2][c];
            for (int a = 0; a < 4; a++)
              V[(size_t)j * G + ((size_t)(f[2] + c) * ng + (f[1] + b)) * ng + f[0] + a] +=
                  qv * w[0][a];
          }
    }
  }
}

/* Spread with per-thread private grids and a reduction, as the reference
 * does with its np workers (HY/nuconv.cpp:55-86, HY/gridding.cpp:164-211). */
void or_s2g(int n, int d, int ng, int nVec, const double *y, const double *q, double *V) {
  const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
  const size_t szV = G * (size_t)nVec;
  int nt = or_get_threads();
  if (nt <= 1 || n < 4096) {
    s2g_range(0, n, d, ng, nVec, y, q, V);
    return;
  }
  double *priv = (double *)calloc(szV * (size_t)nt, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
    int t = 0, tn = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    tn = omp_get_num_threads();
#endif
    int i0 = (int)((int64_t)n * t / tn), i1 = (int)((int64_t)n * (t + 1) / tn);
    s2g_range(i0, i1, d, ng, nVec, y, q, priv + szV * (size_t)t);
#pragma omp barrier
#pragma omp for schedule(static)
    for (size_t e = 0; e < szV; e++) {
      double s = 0.0;
      for (int p = 0; p < tn; p++) s += priv[szV * (size_t)p + e];
      V[e] += s;
    }
  }
  free(priv);
}

void or_g2s(int n, int d, int ng, int nVec, const double *y, const double *V, double *Phi) {
  const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    int f[3] = {0, 0, 0};
    double w[3][4];
    stencil_setup(d, y + (size_t)i * d, f, w);
    for (int j = 0; j < nVec; j++) {
      double acc = 0.0;
      const double *Vj = V + (size_t)j * G;
      if (d == 1) {
        for (int a = 0; a < 4; a++) acc += Vj[f[0] + a] * w[0][a];
      } else if (d == 2) {
        for (int b = 0; b < 4; b++)
          for (int a = 0; a < 4; a++)
            acc += Vj[(size_t)(f[1] + b) * ng + f[0] + a] * w[1][b] * w[0][a];
      } else {
        for (int c = 0; c < 4; c++)
          for (int b = 0; b < 4; b++) {
            double wbc = w[1][b] * w[2][c];
            for (int a = 0; a < 4; a++)
              acc += Vj[((size_t)(f[2] + c) * ng + (f[1] + b)) * ng + f[0] + a] * wbc * w[0][a];
          }
      }
      Phi[(size_t)i * nVec + j] = acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (recursive decimation in time, generic radix
 * butterflies).  FFTW sign convention, unnormalised.                        */
typedef struct { double re, im; } cpx;

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

static void fft_rec(int n, const cpx *in, int is, cpx *out, const cpx *tw, int ts, cpx *tmp) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = smallest_factor(n);
  int m = n / p;
  for (int r = 0; r < p; r++) fft_rec(m, in + (size_t)r * is, is * p, out + (size_t)r * m, tw, ts * p, tmp);
  const int N = n * ts; /* twiddle table length */
  for (int k = 0; k < n; k++) {
    int kk = k % m;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < p; r++) {
      cpx a = out[(size_t)r * m + kk];
      long e = ((long)r * k) % n;
      cpx w = tw[(size_t)(e * ts) % N];
      sr += a.re * w.re - a.im * w.im;
      si += a.re * w.im + a.im * w.re;
    }
    tmp[k].re = sr; tmp[k].im = si;
  }
  memcpy(out, tmp, sizeof(cpx) * (size_t)n);
}

static cpx *make_twiddles(int n, int sign) {
  cpx *tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int j = 0; j < n; j++) {
    double a = 2.0 * M_PI * (double)j / (double)n;
    tw[j].re = cos(a);
    tw[j].im = sign * sin(a);
  }
  return tw;
}

void or_fft(int n, double *data, int sign) {
  cpx *tw = make_twiddles(n, sign < 0 ? -1 : 1);
  cpx *in = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  memcpy(in, data, sizeof(cpx) * (size_t)n);
  fft_rec(n, in, 1, (cpx *)data, tw, 1, tmp);
  free(tw); free(in); free(tmp);
}

/* FFT along every axis of a d-dimensional ng^d complex array, for `howmany`
 * consecutive arrays. */
static void fft_nd(int d, int ng, int howmany, cpx *A, int sign) {
  cpx *tw = make_twiddles(ng, sign);
  const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
  for (int ax = 0; ax < d; ax++) {
    size_t stride = ax == 0 ? 1 : ax == 1 ? (size_t)ng : (size_t)ng * ng;
    size_t lines = G / (size_t)ng;
#pragma omp parallel
    {
      cpx *in = (cpx *)malloc(sizeof(cpx) * (size_t)ng);
      cpx *out = (cpx *)malloc(sizeof(cpx) * (size_t)ng);
      cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)ng);
#pragma omp for schedule(static)
      for (size_t L = 0; L < lines * (size_t)howmany; L++) {
        size_t arr = L / lines, l = L % lines;
        /* base offset of line l along axis ax */
        size_t lo = l % stride, hi = l / stride;
        size_t base = arr * G + hi * stride * (size_t)ng + lo;
        for (int e = 0; e < ng; e++) in[e] = A[base + (size_t)e * stride];
        fft_rec(ng, in, 1, out, tw, 1, tmp);
        for (int e = 0; e < ng; e++) A[base + (size_t)e * stride] = out[e];
      }
      free(in); free(out); free(tmp);
    }
  }
  free(tw);
}

/* kernel{1,2,3}d: HY/non_periodic_conv.cpp:10-20 (kpow = 2); kpow = 1 is
 * the plain Cauchy kernel 1/(1 + h^2 r^2) used for the direct Z. */
static inline double kern(double hsq, double r2, int kpow) {
  double t = 1.0 / (1.0 + hsq * r2);
  return kpow == 2 ? t * t : t;
}

/* HY/non_periodic_conv.cpp:26-708: sum over the 2^d even/odd terms of
 * circulant (+) and skew-circulant (-, twiddled by exp(-i pi k / n)) parts. */
void or_conv_nopad_fft_k(int d, int ng, int nVec, double h, int kpow, const double *V, double *Phi) {
  const double hsq = h * h;
  const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
  cpx *K = (cpx *)malloc(sizeof(cpx) * G);
  cpx *X = (cpx *)malloc(sizeof(cpx) * G * (size_t)nVec);
  cpx *w = (cpx *)malloc(sizeof(cpx) * (size_t)ng);
  for (int i = 0; i < ng; i++) {
    double a = -2.0 * M_PI * (double)i / (2.0 * ng);
    w[i].re = cos(a); w[i].im = sin(a);
  }
  memset(Phi, 0, sizeof(double) * G * (size_t)nVec);
  const int nterms = 1 << d;
  for (int par = 0; par < nterms; par++) {
    /* kernel */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < G; idx++) {
      int ii[3] = {(int)(idx % ng), d > 1 ? (int)((idx / ng) % ng) : 0, d > 2 ? (int)(idx / ((size_t)ng * ng)) : 0};
      double acc = 0.0;
      for (int sub = 0; sub < nterms; sub++) {
        double sgn = 1.0, r2 = 0.0;
        int ok = 1;
        for (int a = 0; a < d; a++) {
          int c = ii[a];
          if (sub & (1 << a)) {
            if (c == 0) { ok = 0; break; }
            c = ng - c;
            if (par & (1 << a)) sgn = -sgn;
          }
          r2 += (double)c * c;
        }
        if (ok) acc += sgn * kern(hsq, r2, kpow);
      }
      cpx t = {acc, 0.0};
      for (int a = 0; a < d; a++)
        if (par & (1 << a)) {
          cpx ww = w[ii[a]];
          cpx u = {t.re * ww.re - t.im * ww.im, t.re * ww.im + t.im * ww.re};
          t = u;
        }
      K[idx] = t;
    }
    /* right-hand sides */
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < G; idx++) {
      int ii[3] = {(int)(idx % ng), d > 1 ? (int)((idx / ng) % ng) : 0, d > 2 ? (int)(idx / ((size_t)ng * ng)) : 0};
      cpx t = {1.0, 0.0};
      for (int a = 0; a < d; a++)
        if (par & (1 << a)) {
          cpx ww = w[ii[a]];
          cpx u = {t.re * ww.re - t.im * ww.im, t.re * ww.im + t.im * ww.re};
          t = u;
        }
      for (int c = 0; c < nVec; c++) {
        double v = V[(size_t)c * G + idx];
        X[(size_t)c * G + idx].re = v * t.re;
        X[(size_t)c * G + idx].im = v * t.im;
      }
    }
    fft_nd(d, ng, 1, K, -1);
    fft_nd(d, ng, nVec, X, -1);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < G * (size_t)nVec; e++) {
      cpx a = X[e], b = K[e % G];
      X[e].re = a.re * b.re - a.im * b.im;
      X[e].im = a.re * b.im + a.im * b.re;
    }
    fft_nd(d, ng, nVec, X, +1);
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < G; idx++) {
      int ii[3] = {(int)(idx % ng), d > 1 ? (int)((idx / ng) % ng) : 0, d > 2 ? (int)(idx / ((size_t)ng * ng)) : 0};
      cpx t = {1.0, 0.0};
      for (int a = 0; a < d; a++)
        if (par & (1 << a)) {
          cpx ww = {w[ii[a]].re, -w[ii[a]].im}; /* conj */
          cpx u = {t.re * ww.re - t.im * ww.im, t.re * ww.im + t.im * ww.re};
          t = u;
        }
      for (int c = 0; c < nVec; c++) {
        cpx a = X[(size_t)c * G + idx];
        Phi[(size_t)c * G + idx] += a.re * t.re - a.im * t.im;
      }
    }
  }
  const double scale = 1.0 / ((double)nterms * (double)G);
  for (size_t e = 0; e < G * (size_t)nVec; e++) Phi[e] *= scale;
  free(K); free(X); free(w);
}

void or_conv_nopad_fft(int d, int ng, int nVec, double h, const double *V, double *Phi) {
  or_conv_nopad_fft_k(d, ng, nVec, h, 2, V, Phi);
}

void or_conv_nopad_direct(int d, int ng, int nVec, double h, int kpow, const double *V, double *Phi) {
  const double hsq = h * h;
  const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
#pragma omp parallel for schedule(static)
  for (size_t o = 0; o < G; o++) {
    int oi[3] = {(int)(o % ng), d > 1 ? (int)((o / ng) % ng) : 0, d > 2 ? (int)(o / ((size_t)ng * ng)) : 0};
    double acc[4] = {0, 0, 0, 0};
    for (size_t s = 0; s < G; s++) {
      int si[3] = {(int)(s % ng), d > 1 ? (int)((s / ng) % ng) : 0, d > 2 ? (int)(s / ((size_t)ng * ng)) : 0};
      double r2 = 0.0;
      for (int a = 0; a < d; a++) r2 += (double)(oi[a] - si[a]) * (oi[a] - si[a]);
      double k = 1.0 / (1.0 + hsq * r2);
      if (kpow == 2) k = k * k;
      for (int c = 0; c < nVec && c < 4; c++) acc[c] += k * V[(size_t)c * G + s];
    }
    for (int c = 0; c < nVec && c < 4; c++) Phi[(size_t)c * G + o] = acc[c];
  }
}

/* ------------------------------------------------------------------------ */
/* Interpolated repulsion: HY/qq.cpp:132-244 + HY/nuconv.cpp:22-153.        */
double or_repulsive_interp(int n, int d, double *y, double h, double *Frep, int *ng_out) {
  return or_repulsive_interp_z(n, d, y, h, 0, Frep, ng_out);
}

double or_repulsive_interp_z(int n, int d, double *y, double h, int zmode, double *Frep, int *ng_out) {
  return or_repulsive_interp_g(n, d, y, h, zmode, 0, Frep, ng_out);
}

double or_repulsive_interp_g(int n, int d, double *y, double h, int zmode, int fixed_ng, double *Frep,
                             int *ng_out) {
  const int m = d + 1;
  double miny[3] = {INFINITY, INFINITY, INFINITY};
  for (int i = 0; i < n; i++)
    for (int a = 0; a < d; a++)
      if (miny[a] > y[(size_t)i * d + a]) miny[a] = y[(size_t)i * d + a];
  for (size_t t = 0; t < (size_t)n * d; t++) y[t] -= miny[t % d];
  double maxy = 0.0;
  for (size_t t = 0; t < (size_t)n * d; t++) maxy = maxy < y[t] ? y[t] : maxy;
  int nGrid;
  if (fixed_ng > 0) {
    /* fixed grid: n1 = getBestGridSize(params.ng), CU/gradient_descend.cu:139-140 */
    nGrid = or_best_grid_size(fixed_ng);
  } else {
    nGrid = (int)ceil(maxy / h);
    if (nGrid < 14) nGrid = 14;
    nGrid = or_best_grid_size(nGrid);
  }
  const int ng = nGrid + 2;
  if (ng_out) *ng_out = ng;

  double *yt = (double *)malloc(sizeof(double) * (size_t)n * d);
  double *VScat = (double *)malloc(sizeof(double) * (size_t)n * m);
  double *PhiScat = (double *)calloc((size_t)n * m, sizeof(double));
  memcpy(yt, y, sizeof(double) * (size_t)n * d);
  for (int i = 0; i < n; i++) {
    VScat[(size_t)i * m] = 1.0;
    for (int a = 0; a < d; a++) VScat[(size_t)i * m + a + 1] = yt[(size_t)i * d + a];
  }
  /* nuconv: scale into [0, nGrid-1) */
  double mx = 0.0;
  for (size_t t = 0; t < (size_t)n * d; t++) mx = mx < yt[t] ? yt[t] : mx;
  for (size_t t = 0; t < (size_t)n * d; t++) {
    yt[t] /= mx;
    if (yt[t] == 1.0) yt[t] -= DBL_EPSILON;
    yt[t] *= (double)(nGrid - 1);
  }
  int bad = 0;
  for (size_t t = 0; t < (size_t)n * d; t++)
    if (yt[t] >= nGrid - 1 || yt[t] < 0 || yt[t] != yt[t]) bad = 1;
  double Z = NAN;
  if (!bad) {
    double hh = mx / ((double)(nGrid - 1) - DBL_EPSILON);
    const size_t G = (size_t)(d == 1 ? ng : d == 2 ? ng * ng : ng * ng * ng);
    double *VGrid = (double *)calloc(G * (size_t)m, sizeof(double));
    double *PhiGrid = (double *)calloc(G * (size_t)m, sizeof(double));
    or_s2g(n, d, ng, m, yt, VScat, VGrid);
    or_conv_nopad_fft(d, ng, m, hh, VGrid, PhiGrid);
    or_g2s(n, d, ng, m, yt, PhiGrid, PhiScat);
    if (zmode == 1) {
      /* direct Z: the unit charges convolved with 1/(1 + r^2), gathered */
      double *Psi = (double *)calloc(G, sizeof(double));
      double *psi = (double *)malloc(sizeof(double) * (size_t)n);
      or_conv_nopad_fft_k(d, ng, 1, hh, 1, VGrid, Psi);
      or_g2s(n, d, ng, 1, yt, Psi, psi);
      Z = 0.0;
      for (int i = 0; i < n; i++) Z += psi[i];
      Z = Z - n;
      free(Psi); free(psi);
    } else {
      /* zetaAndForce: HY/qq.cpp:75-106 */
      Z = 0.0;
      for (int i = 0; i < n; i++) {
        double Ysq = 0.0;
        for (int a = 0; a < d; a++) {
          double ya = y[(size_t)i * d + a];
          Ysq += ya * ya;
          Z -= 2.0 * (ya * PhiScat[(size_t)i * m + a + 1]);
        }
        Z += (1.0 + 2.0 * Ysq) * PhiScat[(size_t)i * m];
      }
      Z = Z - n;
    }
    free(VGrid); free(PhiGrid);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
      for (int a = 0; a < d; a++)
        Frep[(size_t)i * d + a] =
            (y[(size_t)i * d + a] * PhiScat[(size_t)i * m] - PhiScat[(size_t)i * m + a + 1]) / Z;
  }
  free(yt); free(VScat); free(PhiScat);
  return Z;
}

/* ------------------------------------------------------------------------ */
static inline double sgn(double x) { return x == 0.0 ? 0.0 : (x < 0.0 ? -1.0 : 1.0); }

/* HY/gradient_descendHYB.cpp:29-43 (compute_dy) + :46-78 (update_positions) */
void or_update(int n, int d, const double *Fattr, const double *Frep, double alpha,
               double momentum, double eta, double *Y, double *uY, double *gains) {
  const size_t nd = (size_t)n * d;
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < nd; t++) {
    double dy = alpha * Fattr[t] - Frep[t];
    gains[t] = (sgn(dy) != sgn(uY[t])) ? (gains[t] + .2) : (gains[t] * .8);
    if (gains[t] < .01) gains[t] = .01;
    uY[t] = momentum * uY[t] - eta * gains[t] * dy;
    Y[t] = Y[t] + uY[t];
  }
  double mean[3] = {0, 0, 0};
  for (int a = 0; a < d; a++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += Y[(size_t)i * d + a];
    mean[a] = s / n;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int a = 0; a < d; a++) Y[(size_t)i * d + a] -= mean[a];
}

/* CU/pq.cu:61-125 */
double or_kl(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
             const double *Y, double Z, double alpha) {
  double C = 0.0;
#pragma omp parallel for reduction(+ : C) schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    const double *yi = Y + (size_t)i * d;
    for (int64_t t = rowptr[i]; t < rowptr[i + 1]; t++) {
      const double *yj = Y + (size_t)col[t] * d;
      double dist = 0.0;
      for (int a = 0; a < d; a++) dist += (yi[a] - yj[a]) * (yi[a] - yj[a]);
      double p = val[t] * alpha;
      double q = 1.0 / (Z * (1.0 + dist));
      C += p * log((p + (double)FLT_MIN) / (q + (double)FLT_MIN));
    }
  }
  return C;
}

/* HY/gradient_descendHYB.cpp:160-244 with the CPU repulsion of HY/qq.cpp */
void or_run(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
            double alpha0, int early_iter, double h, int it0, int it1,
            double *Y, double *uY, double *gains, double *zeta_out, int *ng_out) {
  or_run_z(n, d, rowptr, col, val, alpha0, early_iter, h, 0, it0, it1, Y, uY, gains, zeta_out, ng_out);
}

void or_run_z(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
              double alpha0, int early_iter, double h, int zmode, int it0, int it1,
              double *Y, double *uY, double *gains, double *zeta_out, int *ng_out) {
  or_run_g(n, d, rowptr, col, val, alpha0, early_iter, h, zmode, 0, it0, it1, Y, uY, gains, zeta_out, ng_out);
}

void or_run_g(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
              double alpha0, int early_iter, double h, int zmode, int fixed_ng, int it0, int it1,
              double *Y, double *uY, double *gains, double *zeta_out, int *ng_out) {
  /* kl_minimization, HY/gradient_descendHYB.cpp:160-225 / CU/gradient_descend.cu:158-236:
   * iteration `it` runs with the alpha and momentum in force, and the switch
   * happens after the update of iteration stop_lying_iter (resp. 250), so
   * iterations 0..early_iter use alpha0 and 0..250 use momentum 0.5 */
  const int mom_switch_iter = 250;
  const double eta = 200.0;
  double *Fattr = (double *)malloc(sizeof(double) * (size_t)n * d);
  double *Frep = (double *)malloc(sizeof(double) * (size_t)n * d);
  for (int it = it0; it < it1; it++) {
    double alpha = it <= early_iter ? alpha0 : 1.0;
    double momentum = it <= mom_switch_iter ? 0.5 : 0.8;
    int ng = 0;
    double Z = or_repulsive_interp_g(n, d, Y, h, zmode, fixed_ng, Frep, &ng);
    or_attractive(n, d, rowptr, col, val, Y, Fattr);
    or_update(n, d, Fattr, Frep, alpha, momentum, eta, Y, uY, gains);
    if (zeta_out) zeta_out[it - it0] = Z;
    if (ng_out) ng_out[it - it0] = ng;
  }
  free(Fattr); free(Frep);
}
