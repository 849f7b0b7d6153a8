/*
 * sgtsne_oracle.h -- FP64 CPU restatement of the SG-t-SNE gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (the HIP library) never does.
 *
 * Every function restates, in plain C99 + OpenMP, the algorithm of the
 * reference file:line it cites (paths relative to the reference root):
 *   HY = code/cuda/SG-tSNE-HYB/sgtsnepi/src   (FP64 CPU path, AoS Y)
 *   CU = code/cuda/SG-tSNE-CUDA/src           (FP32 GPU path, SoA Y)
 *
 * Conventions: Y is AoS [n][d] (HYB layout); sparse matrices are CSR with
 * int64 row pointers, int32 column indices and double values; the kNN arrays
 * are row-major [n][k+1] with the point itself first.
 */
#ifndef SGTSNE_ORACLE_H
#define SGTSNE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid-size table lookup: HY/qq.cpp:108-130 (= CU/utils.cu:194-215). */
int or_best_grid_size(int nGrid);

/* Exact all-kNN (k+1 neighbours, self first).  Squared distances are FP64
 * sums in dimension order, each term added with fma(diff, diff, acc); ties are
 * broken by the lower index.  Replaces the approximate FLANN search of
 * CU/allknn.cuh:75-92 / HY/demo_perplexity_equalization.cpp:24-45. */
void or_knn_exact(const float *X, int n, int dim, int k, int32_t *I, double *D);

/* The same rows for the query points rows[0..nrows) only (against all n
 * points); row r of I and D belongs to point rows[r]. */
void or_knn_rows(const float *X, int n, int dim, int k, const int32_t *rows, int nrows, int32_t *I, double *D);

/* Perplexity bisection of one row: HY/sgtsne.cpp:201-254 (equalizeVertex).
 * dist has k+1 entries (self first, skipped).  Writes k row-normalised
 * values; returns the number of bisection iterations, beta in *beta. */
int or_equalize_vertex(double *val, const double *dist, double u, int k, double *beta);

/* All rows + kNN->CSR: HY/sgtsne.cpp:257-297 (perplexityEqualization),
 * emitted in CSR (row i holds the k non-self neighbours of i). */
void or_perplexity_equalize(const int32_t *I, const double *D, int n, int k, double u,
                            int64_t *rowptr, int32_t *col, double *val);

/* Row-stochastic fix-up: HY/sparsematrix.cpp:27-67 (makeStochastic). Returns
 * the number of rows already stochastic. */
int or_make_stochastic(int n, const int64_t *rowptr, double *val);

/* Lambda rescaling (bisection per row): HY/graph_rescaling.cpp:19-104. */
void or_lambda_rescale(int n, const int64_t *rowptr, double *val, double lambda, int drop_leaf);

/* Symmetrisation P + P^T (HY/sparsematrix.cpp:69-164) followed by total-sum
 * normalisation (HY/sgtsne.cpp:78-85).  Output rows are sorted by column.
 * Two calls: first with col_out == NULL returns nnz_sym; second fills. */
int64_t or_symmetrize(int n, const int64_t *rowptr, const int32_t *col, const double *val,
                      int64_t *rowptr_out, int32_t *col_out, double *val_out);

/* Attractive term: HY/pq.cpp:15-57.
 * Fattr_i = sum_j P_ij (y_i - y_j) / (1 + |y_i - y_j|^2). */
void or_attractive(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
                   const double *Y, double *Fattr);

/* Exact O(N^2) repulsion: HY/qq.cpp:25-70.  Frep normalised by Z, returns Z. */
double or_repulsive_exact(int n, int d, const double *Y, double *Frep);

/* Exact O(N^2) sum of the squared kernel against charges (1, y): the value
 * that the interpolation approximates, PhiScat [n][d+1]. */
void or_phi_exact(int n, int d, const double *Y, double *Phi);

/* Lagrange weights g1/g2: HY/gridding.cpp:28-40.  w[4] for offset delta. */
void or_lagrange_weights(double delta, double *w);

/* Spread (S2G) on an ng^d grid with nVec channels: HY/gridding.cpp:57-101,
 * 164-211, 260-318 (single-worker form).  y are grid coordinates [n][d],
 * q charges [n][nVec], V [nVec][ng^d] (x fastest), accumulated (+=). */
void or_s2g(int n, int d, int ng, int nVec, const double *y, const double *q, double *V);

/* Gather (G2S): HY/gridding.cpp:455-624.  Phi [n][nVec]. */
void or_g2s(int n, int d, int ng, int nVec, const double *y, const double *V, double *Phi);

/* Non-periodic convolution with kernel (1 + h^2 r^2)^-2 on an ng^d grid,
 * through the 2^d even/odd circulant terms and an own mixed-radix FFT:
 * HY/non_periodic_conv.cpp:26-708 (conv{1,2,3}dnopad). */
void or_conv_nopad_fft(int d, int ng, int nVec, double h, const double *V, double *Phi);

/* Same with kernel power kpow: 2 = (1 + h^2 r^2)^-2, 1 = (1 + h^2 r^2)^-1. */
void or_conv_nopad_fft_k(int d, int ng, int nVec, double h, int kpow, const double *V, double *Phi);

/* The same convolution as a direct Toeplitz sum (O(ng^{2d})). */
void or_conv_nopad_direct(int d, int ng, int nVec, double h, int kpow, const double *V, double *Phi);

/* Complex DFT (FFTW sign convention, unnormalised), mixed radix, any n.
 * data is interleaved re/im, length n, stride 1.  sign = -1 forward. */
void or_fft(int n, double *data, int sign);

/* Interpolated repulsion: HY/qq.cpp:132-244 with HY/nuconv.cpp:22-153 and
 * zetaAndForce HY/qq.cpp:75-106.  y is shifted to min 0 in place, exactly as
 * the reference does.  Returns Z; *ng_out receives the grid side nGrid+2. */
double or_repulsive_interp(int n, int d, double *y, double h, double *Frep, int *ng_out);

/* As above with the choice of Z: zmode 0 = the reference's algebraic Z from
 * the squared kernel (HY/qq.cpp:75-106); zmode 1 = direct Z, the unit charges
 * convolved with 1/(1 + r^2) on the same grid and gathered, minus N.  Frep is
 * the same in both modes (only its normalisation differs). */
double or_repulsive_interp_z(int n, int d, double *y, double h, int zmode, double *Frep, int *ng_out);

/* As or_repulsive_interp_z; fixed_ng > 0 selects the CUDA path's fixed grid
 * (nGrid = or_best_grid_size(fixed_ng), CU/gradient_descend.cu:139-140 with
 * the scaling of CU/nuconv.cu:55-75) instead of the h rule. */
double or_repulsive_interp_g(int n, int d, double *y, double h, int zmode, int fixed_ng, double *Frep,
                             int *ng_out);

/* Gradient + update with mean-centring: HY/gradient_descendHYB.cpp:29-80. */
void or_update(int n, int d, const double *Fattr, const double *Frep, double alpha,
               double momentum, double eta, double *Y, double *uY, double *gains);

/* KL cost: CU/pq.cu:61-125 (ComputeKLkernel + reduce). */
double or_kl(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
             const double *Y, double Z, double alpha);

/* Full gradient-descent loop: HY/gradient_descendHYB.cpp:160-244 with the
 * CPU repulsion (HY/qq.cpp:132).  Runs iterations [it0, it1) of the schedule
 * (alpha = alpha0 until earlyIter, momentum .5 -> .8 after iteration 250,
 * eta = 200), updating Y, uY, gains in place.  zeta_out (nullable) receives
 * Z of every iteration run; ng_out (nullable) the grid side. */
void or_run(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
            double alpha0, int early_iter, double h, int it0, int it1,
            double *Y, double *uY, double *gains, double *zeta_out, int *ng_out);

void or_run_z(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
              double alpha0, int early_iter, double h, int zmode, int it0, int it1,
              double *Y, double *uY, double *gains, double *zeta_out, int *ng_out);

/* As or_run_z with the grid rule of or_repulsive_interp_g. */
void or_run_g(int n, int d, const int64_t *rowptr, const int32_t *col, const double *val,
              double alpha0, int early_iter, double h, int zmode, int fixed_ng, int it0, int it1,
              double *Y, double *uY, double *gains, double *zeta_out, int *ng_out);

/* Threads used by OpenMP regions (0 = leave the runtime default). */
void or_set_threads(int nthreads);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
