// kernels.hpp -- launchers of the gfx950 kernels (host-callable).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sgt {

// Scalars the iteration keeps on the device (reductions and Z).
struct DevScalars {
  double sum[4];    // sum of Y per dimension (after the update)
  double minv[4];   // min of the centred Y per dimension
  double maxv[4];   // max of the centred Y per dimension
  double mean[4];   // mean subtracted by the centring kernel
  double Z;         // normalisation term (sum z_i - N)
  double kl;        // KL divergence (tsneCost)
  double zexact;    // exact-repulsion Z (validation)
  double pad;
};

// Partial-reduction workspace: one slot per block plus a ticket counter.
struct RedWork {
  double *part;       // [max_blocks * 8]
  unsigned *ticket;   // single counter, reset by the last block
  int max_blocks;
};

// Grid mapping of one iteration (computed on the host from min/max).
struct GridParams {
  int d;
  int nGrid;        // interpolation intervals + 1 (reference nGridDim)
  int ng;           // grid side = nGrid + 2
  int L;            // FFT length (even 2^a 3^b 5^c >= 2 ng - 1)
  float minv[3];    // min of Y per dimension (grid origin)
  float scale;      // (nGrid - 1) / maxy
  float gmax;       // largest float below nGrid - 1 (clamp)
  float center[3];  // charge shift (bounding-box centre)
  double hgrid;     // grid spacing in embedding units
};

// ---- attractive term (kernels_spmv.hip) ----
int launch_attr_csr_seg(hipStream_t s, int d, int group, int n, const int32_t *rowptr,
                        const int32_t *col, const float *val, const float *Y, float *F);
int launch_attr_csr_seg2(hipStream_t s, int d, int group, int unroll, int n, const int32_t *rowptr,
                         const int32_t *col, const float *val, const float *Y, float *F);
int launch_attr_csr_blk(hipStream_t s, int d, int group, int unroll, int rows_per_block, int n,
                        const int32_t *rowptr, const int32_t *col, const float *val, const float *Y, float *F);
// windowed blocked form: win points of Y around each row range staged in LDS
// chunk-cached blocked form: per row range, nslots chunks of 2^chs points of Y
// staged in LDS (slots[range * nslots + k] = chunk id or -1)
int launch_attr_csr_cc(hipStream_t s, int d, int group, int unroll, int rows_per_block, int chs, int nslots,
                       const int32_t *slots, int n, const int32_t *rowptr, const int32_t *col, const float *val,
                       const float *Y, float *F);
// builds slots: counts is scratch of div_up(n, rows_per_block) * div_up(n, 2^chs) ints
int launch_chunk_cache(hipStream_t s, int n, int rows_per_block, int chs, int nslots, const int32_t *rowptr,
                       const int32_t *col, int32_t *counts, int32_t *slots);
int launch_attr_csr_win(hipStream_t s, int d, int group, int unroll, int rows_per_block, int win, int n,
                        const int32_t *rowptr, const int32_t *col, const float *val, const float *Y, float *F);
// column-localised form: per row block, Y at its distinct columns staged in LDS
int launch_attr_csr_loc(hipStream_t s, int d, int group, int nblk, int cap, const int32_t *brow, const int32_t *bu,
                        const int32_t *ucol, const int32_t *rowptr, const uint16_t *lidx, const float *val,
                        const float *Y, float *F);
int launch_attr_csr_row(hipStream_t s, int d, int n, const int32_t *rowptr, const int32_t *col,
                        const float *val, const float *Y, float *F);
int launch_attr_coo(hipStream_t s, int d, int64_t nnz, const int32_t *row, const int32_t *col,
                    const float *val, const float *Y, float *F);
int launch_attr_ell(hipStream_t s, int d, int n, int width, const int32_t *ecol, const float *eval,
                    const float *Y, float *F);
int launch_kl(hipStream_t s, int d, int n, const int32_t *rowptr, const int32_t *col,
              const float *val, const float *Y, double alpha, const DevScalars *sc, RedWork rw,
              DevScalars *out);

// ---- repulsion (kernels_grid.hip) ----
int launch_fill(hipStream_t s, float *p, int64_t count, float v);
int launch_stream_copy(hipStream_t s, const float4 *in, float4 *out, int64_t count);
int launch_center_minmax(hipStream_t s, int d, int n, float *Y, int apply_mean, RedWork rw,
                         DevScalars *sc);
// S2G into V [d+1][ng^d] (V is overwritten).  scratch: float workspace for the
// per-block private grids of the LDS path; *needed receives the floats it
// wants (when larger than scratch_floats the global-atomic path runs).
int launch_s2g(hipStream_t s, const GridParams &g, int n, const float *Y, float *V, float *scratch,
               int64_t scratch_floats, int64_t *needed, int mode = 0);
int64_t s2g_scratch_need(const GridParams &g, int n, int mode = 0);
int launch_g2s_force(hipStream_t s, const GridParams &g, int n, const float *Y, const float *Phi,
                     float *Frep, int zdirect, RedWork rw, DevScalars *sc);
int launch_finalize_kl(hipStream_t s, int nblocks, RedWork rw, DevScalars *sc);
int launch_update(hipStream_t s, int d, int n, float *Y, float *uY, float *gains,
                  const float *Fattr, const float *Frep, float alpha, float momentum, float eta,
                  RedWork rw, DevScalars *sc);
int launch_rep_exact(hipStream_t s, int d, int n, const float *Y, float *Frep, RedWork rw,
                     DevScalars *sc);

// ---- FFT convolution (kernels_fft.hip) ----
struct ConvBuffers {
  float2 *W;      // complex work [pairs][...]
  float2 *Kt1;    // (K2, K1) half-spectrum scratch
  float2 *Kt2;
  float2 *Kh;     // (K2^, K1^) half-spectra of the kernels [(L/2+1)^d]
  float2 *tw;     // twiddle tables of every supported length (twiddle_count())
};
int conv_workspace_bytes(int d, int ng, int L, size_t *w_bytes, size_t *k_bytes);
// V [d+1][ng^d] -> Phi [d+1 (+1 with zdirect: K1 * V0 in channel d+1)][ng^d]
int launch_conv(hipStream_t s, const GridParams &g, int zdirect, const float *V, float *Phi,
                const ConvBuffers &b);
// twiddle tables of every supported FFT length (twiddle_count() float2)
size_t twiddle_count();
int init_twiddles(hipStream_t s, float2 *tw);
// FFT length of a grid of side ng: smallest even 2^a 3^b 5^c >= 2 ng - 1
int fft_length_for(int ng);

// ---- P construction (kernels_prep.hip) ----
void retain_pool_memory();
int launch_knn_exact(hipStream_t s, const double *X, int n, int dim, int k, int32_t *I, double *D);
int launch_knn(hipStream_t s, const double *X, int n, int dim, int k, int32_t *I, double *D, int method,
               int *fallback_rows);
// launch_knn in two halves for callers that must not wait in between:
// knn_begin queues everything up to the certificate and a copy of the count
// of uncertified rows into *count_host (pinned memory for a copy that does
// not block); knn_end, called once the stream has passed that copy, queues
// the exact recomputation of those rows and releases the temporaries.
struct KnnPending {
  bool open = false;  // knn_end still has work to queue
  const double *X = nullptr;
  int n = 0, dim = 0, k = 0;
  int32_t *I = nullptr;
  double *D = nullptr;
  int *fb = nullptr;  // device: [0] count, then the rows
  const int *count_host = nullptr;
  void *bufs[32] = {};
  int nbufs = 0;
};
int knn_begin(hipStream_t s, const double *X, int n, int dim, int k, int32_t *I, double *D, int method,
              int *count_host, KnnPending *kp);
int knn_end(hipStream_t s, KnnPending *kp, int *fallback_rows);
void knn_release(hipStream_t s, KnnPending *kp);  // frees the temporaries without finishing
int launch_perplexity(hipStream_t s, const double *D, int n, int k, double u, float *val,
                      double *beta_out);
int launch_lambda_rescale(hipStream_t s, int n, const int32_t *rowptr, float *val, double lambda,
                          int drop_leaf);
int launch_rowids(hipStream_t s, int n, const int32_t *rowptr, int32_t *row);
int gpu_symmetrize(hipStream_t s, int n, int64_t nnz, const int32_t *rowptr_d,
                   const int32_t *col_d, const float *val_d, int32_t **rowptr_out,
                   int32_t **col_out, float **val_out, int64_t *nnz_out,
                   bool stream_ordered = false);
// gpu_symmetrize in two halves (see KnnPending): sym_begin queues the merge of
// P and P^T and a copy of the merged nonzero count into *runs_host; sym_end,
// once the stream has passed that copy, queues the normalisation and the CSR
// build into stream-ordered outputs, without waiting for them.
struct SymPending {
  bool open = false;
  int n = 0;
  int64_t m2 = 0;
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  double *v0 = nullptr, *v1 = nullptr, *total = nullptr;
  int *nruns = nullptr;
  const int *runs_host = nullptr;
};
int sym_begin(hipStream_t s, int n, int64_t nnz, const int32_t *rowptr_d, const int32_t *col_d,
              const float *val_d, int *runs_host, SymPending *sp);
int sym_end(hipStream_t s, SymPending *sp, int32_t **rowptr_out, int32_t **col_out, float **val_out,
            int64_t *nnz_out, bool stream_ordered);
void sym_release(hipStream_t s, SymPending *sp);
int launch_build_ell(hipStream_t s, int n, int width, const int32_t *rowptr, const int32_t *col,
                     const float *val, int32_t *ecol, float *eval, int32_t *tail_cnt);
int launch_build_tail(hipStream_t s, int n, int width, const int32_t *rowptr, const int32_t *col,
                      const float *val, const int32_t *tail_off, int32_t *trow, int32_t *tcol,
                      float *tval);
int device_exclusive_scan(hipStream_t s, const int32_t *in, int32_t *out, int n);
// order[r] = current index of the point placed at r (Morton order of Y over
// the box [lo, lo + extent]^d); inv = order^-1.
int spatial_order(hipStream_t s, int n, int d, int ds, const float *Y, const double *lo, double extent,
                  int32_t *order, int32_t *inv);
int launch_gather_points(hipStream_t s, int n, int ds, const int32_t *order, const float *in, float *out);
int launch_permute_csr(hipStream_t s, int n, int64_t nnz, const int32_t *perm, const int32_t *iperm,
                       const int32_t *rp, const int32_t *col, const float *val, int32_t *rp2,
                       int32_t *col2, float *val2,
                       bool sort_rows = true, bool device_sort = false);

}  // namespace sgt
