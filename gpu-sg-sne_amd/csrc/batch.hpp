// batch.hpp -- device records and launchers of the batched engine: many
// independent embedding jobs of uneven size advanced together, one launch
// per phase for all of them.
//
// All jobs of one batch share d.  Their points occupy consecutive ranges of
// one arena (Y, uY, gains, Fattr, Frep), their P_sym rows form one
// block-diagonal CSR (global column indices), so the attractive term is a
// single SpMV over the arena.  Each job has its own grid, FFT buffers and
// schedule; the grid of the next iteration is derived on the device from the
// job's bounding box, so an iteration needs no host round trip and can be
// captured in a graph.
//
// Every per-job result is independent of the other jobs in the batch and of
// how the work is cut into launches: the spread accumulates 64-bit fixed-point
// integers (exact, order-free), and every reduction runs over the job's own
// fixed-size chunks in a fixed order.
#pragma once

#include "kernels.hpp"

namespace sgt {

enum BatchJobStatus {
  BJ_EMPTY = 0,    // slot unused
  BJ_INIT = 1,     // admitted: centring and first grid at the next iteration
  BJ_RUN = 2,      // iterating
  BJ_DONE = 3,     // max_iter reached
  BJ_ERR_NUMERIC = -4,
  BJ_ERR_GRID = -7,  // the grid outgrew the job's capacity
};

constexpr int kBatchChunk = 1024;  // points per work item (fixes each job's reduction order)

// One job (device-resident).  The host writes the static part at admission
// and moves off / item0 on compaction; the kernels own the dynamic part.
struct BatchJob {
  int off, n;            // point range [off, off + n) of the arena
  int item0, nitems;     // work items of the job
  int max_iter, early_iter;
  int grid_mode, ng_fixed;
  int ng_cap, zdirect;
  float alpha0, pad0;
  double h;              // grid spacing (h-driven grid size)
  float mean0[4];        // subtracted from Y at admission
  long long *vfix;       // spread grid, fixed point [d+1][ng^d] (ng <= ng_cap)
  float *phi;            // potentials: [ng^d] float4 of channels 0-3, then channel 4 [ng^d] (3-D)
  float2 *W, *Kt1, *Kt2, *Kh;
  // dynamic
  int iter, status;
  double fix_scale, fix_inv;
  double Z;
  double minv[4], maxv[4];
  GridParams g;
};

struct BatchItem {
  int job, begin, end, pad;
};

// Bytes of one job's grid buffers at grid side ng_cap.
size_t batch_grid_bytes(int d, int ng_cap, size_t *vfix_b, size_t *phi_b, size_t *w_b, size_t *k_b);
size_t batch_pass_args_bytes();
int batch_npasses(int d);

// Per-iteration phases (all jobs, persistent work-groups over the items).
int launch_bs2g(hipStream_t s, int d, const BatchJob *jobs, const BatchItem *items, int nitems, const float *Y,
                int blocks, int lds_bytes);
int launch_bconv(hipStream_t s, int d, int zdirect, const void *args, const int *prefix, int njobs, int blocks);
int launch_bg2s(hipStream_t s, int d, int zdirect, const BatchJob *jobs, const BatchItem *items, int nitems,
                const float *Y, float *Frep, double *zpart, int blocks);
int launch_bupdate(hipStream_t s, int d, BatchJob *jobs, const BatchItem *items, int nitems, float *Y, float *uY,
                   float *gains, const float *Fattr, const float *Frep, const double *zpart, double *spart,
                   int blocks);
int launch_bcenter(hipStream_t s, int d, const BatchJob *jobs, const BatchItem *items, int nitems, float *Y,
                   const double *spart, double *mpart, int blocks);
// Bounding box -> grid, fixed-point scale and FFT pass table of every job;
// advances the iteration counters.
int launch_bplan(hipStream_t s, int d, BatchJob *jobs, int njobs, const double *mpart, void *args, int *prefix,
                 const float2 *tw);
// Morton keys of one job's points over its current grid box (key = cell code, value = local index).
int launch_bmorton(hipStream_t s, int d, const BatchJob *jobs, int job, int off, int n, const float *Y,
                   uint32_t *keys, int32_t *vals);
int launch_offset_copy_i32(hipStream_t s, const int32_t *in, int32_t *out, int64_t count, int32_t delta);
// Z of one job from the item partials of the last gather (one wavefront).
int launch_bzeta(hipStream_t s, const BatchJob *jobs, int job, const double *zpart, double *out);
int launch_knn_csr(hipStream_t s, int n, int k, const int32_t *I, int32_t *rowptr, int32_t *col);
int launch_gather_i32(hipStream_t s, int n, const int32_t *order, const int32_t *in, int32_t *out);
int launch_iota(hipStream_t s, int n, int32_t *out);
int launch_invert(hipStream_t s, int n, const int32_t *order, int32_t *inv);
int launch_to_double(hipStream_t s, const float *in, double *out, int64_t m);

}  // namespace sgt
