// batch.hip -- the batched engine behind sgtsne_batch_* (include/sgtsne.h):
// many independent embedding jobs of uneven size advanced together, one
// launch per phase for all of them, iterations replayed from a captured
// HIP graph.  The single-job pipeline it batches is sgtsneCUDA
// (CU/sgtsne.cu:26-128) / kl_minimization (CU/gradient_descend.cu:102-245).
//
// Memory: one arena of points (Y, uY, gains, Fattr, Frep, perm; stride ds per
// point) and one block-diagonal CSR of P_sym with global column indices.  A
// job occupies consecutive rows [off, off + n) and nnz [nnz_off, nnz_off + nnz).
// Jobs are appended at the tail; a removed job leaves a hole (its rows stay
// valid and are still multiplied, harmlessly) until a compaction moves the
// later jobs down.  Each job has its own grid / FFT buffers sized for the grid
// side ng_cap it may reach; the sum of them is bounded by a byte budget.
//
// Schedule: the host knows every job's iteration without reading the device
// (a job spends one batch iteration in INIT, then advances one iteration per
// batch iteration until max_iter), so it can cut the replay at the locality
// re-orders of each job (its iterations 50, 150, 300, 600) without syncing.
#include "common.hpp"
#include "kernels.hpp"
#include "batch.hpp"
#include "../../include/sgtsne.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <vector>

using namespace sgt;

namespace {

constexpr int kReorderIters[] = {50, 150, 300, 600};

struct HostJob {
  bool live = false;
  int off = 0, n = 0;
  int64_t nnz_off = 0, nnz = 0;
  int max_iter = 0;
  int64_t t = 0;         // batch iterations run since admission (INIT included)
  int ng_cap = 0;
  size_t gbytes = 0;
  void *gmem = nullptr;  // vfix | phi | W | Kt1 | Kt2 | Kh
  int reorders = 0;
  // recorded on the iteration stream after the step that ran the job's last
  // iteration: its Y and record can then be read without waiting for later
  // iterations (unless a compaction moved the job since)
  hipEvent_t done_ev = nullptr;
  bool done_marked = false, moved_after_done = false;
};

// A span of the pinned staging ring and the event after the copies reading it.
struct StageSpan {
  size_t off, len;
  hipEvent_t ev;  // nullptr until the copies reading the span are queued
};

constexpr size_t kStageBytes = 16u << 20;  // pinned staging for host -> device updates
constexpr int kMarks = 16;                 // tick marks kept

// One admission from points (sgtsne_batch_add_points_async) whose front end
// is queued in three parts on the front-end stream; sgtsne_batch_add_poll
// queues the next part once the device has passed the previous one, so no
// part waits on the host.  The parts end where the host needs a count from
// the device: the kNN rows left uncertified, then the nonzeros of P + P^T.
// Up to kFrontEnds of them are in flight at once: their parts queue behind
// one another on the stream, so the device runs one front end's part while
// the host takes up another's.
constexpr int kFrontEnds = 8;
struct FrontEnd {
  int stage = 0;  // 0 free, 1 kNN queued, 2 P queued up to its merge, 3 P_sym queued
  int64_t ticket = 0;
  int n = 0, dim = 0, k = 0, max_iter = 0, ng_cap = 0;
  double u = 0.0;
  const float *y_in = nullptr;  // the caller's, read at admission
  float *Xf = nullptr, *vv = nullptr, *svv = nullptr;
  double *Xd = nullptr, *Dd = nullptr;
  int32_t *Id = nullptr, *rp = nullptr, *cc = nullptr, *srp = nullptr, *scc = nullptr;
  int64_t m = 0;
  KnnPending knn;
  SymPending sym;
  hipEvent_t ev = nullptr;  // after the queued part (kept across uses of the slot)
  int *counts = nullptr;    // pinned: [0] uncertified kNN rows, [1] nonzeros of P + P^T
};

}  // namespace

struct sgtsne_batch {
  sgtsne_params p{};
  int d = 2, ds = 2;
  hipStream_t stream = nullptr;  // iterations and every arena write
  hipStream_t prep = nullptr;    // front end of new jobs (kNN, perplexity, P + P^T)
  // reads of finished jobs: the front-end stream, idle whenever such a read
  // runs unless an asynchronous front end is in flight (the read then waits
  // for its queued part).  Two streams per batch, not three: a server's two
  // engines then need four, one per hardware queue (GPU_MAX_HW_QUEUES = 4),
  // and do not depend on which queues earlier streams of the process left
  // busiest
  hipStream_t copy = nullptr;
  FrontEnd fe[kFrontEnds];
  int64_t fe_tickets = 0;
  int *fe_host = nullptr;  // pinned: two counts per front-end slot
  // host -> device updates go through a pinned ring, so queuing them never
  // waits for the iterations queued ahead of them
  char *stage = nullptr;
  size_t stage_head = 0;
  std::deque<StageSpan> stage_live;
  std::vector<hipEvent_t> stage_free;
  // marks recorded on the iteration stream (pacing of queued iterations)
  hipEvent_t marks[kMarks] = {};
  int64_t nmarks = 0;
  int64_t passed_upto = -1;  // the latest mark known to be passed
  // guards the batch's host state: every entry point holds it (the front end
  // of sgtsne_batch_add / _add_points only around the admission itself)
  std::mutex mu;
  // scratch of the re-orders and compactions, reused by each: they all run
  // on the iteration stream, in order, so one buffer serves every call and
  // no allocation or release sits between queued iterations
  char *ws = nullptr;
  size_t ws_bytes = 0;
  int max_jobs = 0;
  int64_t max_points = 0, max_nnz = 0;
  size_t grid_budget = 0, grid_used = 0;

  // arena
  int32_t *rowptr = nullptr, *col = nullptr;
  float *val = nullptr;
  float *Y = nullptr, *uY = nullptr, *gains = nullptr, *Fattr = nullptr, *Frep = nullptr;
  int32_t *perm = nullptr;  // device row -> the caller's index within its job
  int64_t used_points = 0, used_nnz = 0;

  // job table and work items
  BatchJob *jobs_d = nullptr;
  std::vector<BatchJob> jobs_h;  // static part mirrored on the host
  std::vector<BatchJob> jobs_up; // static part as last uploaded (the device's copy, in stream order)
  std::vector<HostJob> hj;
  std::vector<int> arena;        // live slots in ascending off
  BatchItem *items_d = nullptr;
  int items_cap = 0, nitems = 0;
  double *zpart = nullptr, *spart = nullptr, *mpart = nullptr;
  void *args_d = nullptr;
  int *prefix_d = nullptr;
  float2 *tw = nullptr;
  double *scal_d = nullptr;

  // launch shapes (autotuned for the attractive term; the others are
  // persistent work-group counts)
  int spmv_group = 16, spmv_unroll = 2, spmv_rpb = 0, spmv_win = 0;
  int64_t tuned_points = 0;
  bool tune_pinned = false;  // shape pinned by sgtsne_batch_set_attractive
  std::vector<double> tune_ms;
  int s2g_blocks = 1024, s2g_lds = 32768, conv_blocks = 1024, pt_blocks = 2048;

  // captured iteration: two executable graphs used in turn, so that an
  // update only ever touches the one whose launches have all completed
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;  // the one launched now (execs[cur])
  hipGraphExec_t execs[2] = {nullptr, nullptr};
  hipEvent_t exec_done[2] = {nullptr, nullptr};  // after each one's last launch
  int cur = 0;
  // memory released on the iteration stream, returned to the pool once the
  // queued work that used it has passed its event
  struct Deferred {
    void *p[3];
    hipEvent_t ev;
  };
  std::vector<Deferred> deferred;
  bool graph_valid = false;
  bool use_graph = true;
  int64_t graph_replays = 0, graph_updates = 0, graph_instantiations = 0;
};

namespace {

int balloc(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 1);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return SGTSNE_ERR_NOMEM;
  }
  return SGTSNE_OK;
}

template <class T>
void bfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// Grows the iteration-stream scratch to at least `bytes` (rare: only a job
// larger than any before it); queued work may still read the old buffer.
int ws_reserve(sgtsne_batch *b, size_t bytes) {
  if (bytes <= b->ws_bytes) return SGTSNE_OK;
  SGT_HIP(hipStreamSynchronize(b->stream));
  bfree(b->ws);
  b->ws_bytes = 0;
  const size_t cap = std::max(bytes + bytes / 2, (size_t)1 << 20);
  SGT_TRY(balloc((void **)&b->ws, cap));
  b->ws_bytes = cap;
  return SGTSNE_OK;
}

// Consecutive 256-byte-aligned pieces of one buffer; with a null base it
// only adds up the size the same pieces need.
struct Carve {
  char *base;
  size_t off = 0;
  template <class T>
  T *take(size_t count) {
    T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
    off += (sizeof(T) * std::max<size_t>(count, 1) + 255) & ~(size_t)255;
    return p;
  }
};

double bdefault_h(int d) { return d == 1 ? 0.5 : d == 2 ? 0.7 : 1.2; }

// Largest grid side a job may reach: the embedding extent grows with N, so
// small jobs get small grid buffers.  3-D jobs of 10k-20k points reach sides
// of about 60, so their cap leaves room above that (a job that outgrows its
// cap stops with SGTSNE_ERR_FULL).
int auto_ng_cap(int d, int n) {
  if (d == 1) return 512;
  if (d == 2) return n <= 20000 ? 256 : 512;
  return n <= 20000 ? 96 : 128;
}

// The captured iteration no longer matches the arena: the next step
// re-captures it and updates the executable graph in place.
void invalidate_graph(sgtsne_batch *b) { b->graph_valid = false; }

int stage_commit(sgtsne_batch *b);

// Reserves `bytes` of the staging ring, first waiting for the copies that
// still read an overlapping span; nullptr when the request exceeds the ring.
// A span reserved since the last commit has its copy queued but no event
// yet: when the ring wraps onto one, the group is committed and its event
// waited for before the memory is written again.
void *stage_alloc(sgtsne_batch *b, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes > kStageBytes || !b->stage) return nullptr;
  size_t off = b->stage_head;
  if (off + bytes > kStageBytes) off = 0;
  for (const auto &sp : b->stage_live)
    if (!sp.ev && sp.off < off + bytes && off < sp.off + sp.len) {
      if (stage_commit(b) != SGTSNE_OK) return nullptr;  // the caller then copies synchronously
      break;
    }
  for (auto it = b->stage_live.begin(); it != b->stage_live.end();) {
    const bool overlap = it->off < off + bytes && off < it->off + it->len;
    const bool done = it->ev && hipEventQuery(it->ev) == hipSuccess;
    if (overlap || done) {
      if (overlap && !done && it->ev) (void)hipEventSynchronize(it->ev);
      if (it->ev) b->stage_free.push_back(it->ev);
      it = b->stage_live.erase(it);
    } else {
      ++it;
    }
  }
  b->stage_live.push_back(StageSpan{off, bytes, nullptr});
  b->stage_head = off + bytes;
  return b->stage + off;
}

// Records one event after the queued copies that read the spans reserved
// since the last commit.
int stage_commit(sgtsne_batch *b) {
  hipEvent_t ev = nullptr;
  bool any = false;
  for (auto &sp : b->stage_live) any = any || sp.ev == nullptr;
  if (!any) return SGTSNE_OK;
  if (!b->stage_free.empty()) {
    ev = b->stage_free.back();
    b->stage_free.pop_back();
  } else {
    SGT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  SGT_HIP(hipEventRecord(ev, b->stream));
  for (auto &sp : b->stage_live)
    if (sp.ev == nullptr) sp.ev = ev;
  return SGTSNE_OK;
}

// Queues a host -> device copy on the iteration stream; `src` may be reused
// as soon as this returns.
int upload(sgtsne_batch *b, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return SGTSNE_OK;
  void *p = stage_alloc(b, bytes);
  if (!p) {  // larger than the ring: copy and wait
    SGT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
    SGT_HIP(hipStreamSynchronize(b->stream));
    return SGTSNE_OK;
  }
  std::memcpy(p, src, bytes);
  SGT_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, b->stream));
  return SGTSNE_OK;
}

// Destroys the executable graphs (no launch of them may be pending).
void release_graph(sgtsne_batch *b) {
  for (int k = 0; k < 2; ++k) {
    if (b->execs[k]) (void)hipGraphExecDestroy(b->execs[k]);
    b->execs[k] = nullptr;
  }
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_valid = false;
}

// Queues the release of up to three blocks once the iteration stream has
// passed this point (see release_deferred).
int defer_free(sgtsne_batch *b, void *p0, void *p1 = nullptr, void *p2 = nullptr) {
  hipEvent_t ev = nullptr;
  SGT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  SGT_HIP(hipEventRecord(ev, b->stream));
  b->deferred.push_back(sgtsne_batch::Deferred{{p0, p1, p2}, ev});
  return SGTSNE_OK;
}

// Returns to the pool the deferred blocks whose event has completed (all of
// them with `all`, after the caller has drained the stream).
void release_deferred(sgtsne_batch *b, bool all = false) {
  for (auto it = b->deferred.begin(); it != b->deferred.end();) {
    if (all || hipEventQuery(it->ev) == hipSuccess) {
      for (void *q : it->p)
        if (q) (void)hipFreeAsync(q, b->copy);  // idle stream: the block is free at once
      (void)hipEventDestroy(it->ev);
      it = b->deferred.erase(it);
    } else {
      ++it;
    }
  }
}

// Writes the static part of slot j's record (everything before the dynamic
// fields) so that a running job's iteration state is left alone; skipped
// when the device already holds it (an admission at the tail of the arena
// changes only the new job's record, so admitting J jobs makes J uploads,
// not J^2 / 2).
constexpr size_t kJobStatic = offsetof(BatchJob, iter);
static_assert(kJobStatic == 10 * sizeof(int) + 6 * sizeof(float) + sizeof(double) + 6 * sizeof(void *),
              "the static part of a job record has no padding bytes (it is compared with memcmp)");
int write_static(sgtsne_batch *b, int j) {
  if (std::memcmp(&b->jobs_h[j], &b->jobs_up[j], kJobStatic) == 0) return SGTSNE_OK;
  std::memcpy(&b->jobs_up[j], &b->jobs_h[j], kJobStatic);
  return upload(b, b->jobs_d + j, &b->jobs_h[j], kJobStatic);
}

// Host-side check of the tables a step reads, before they are uploaded: the
// running jobs' point and nonzero ranges lie inside the arena and do not
// overlap, and every work item covers at most kBatchChunk points of its own
// job's range, the items of a job being contiguous from its item0.
int validate_tables(const sgtsne_batch *b, const std::vector<BatchItem> &items) {
  std::vector<std::pair<int64_t, int64_t>> pr, zr;
  for (int j : b->arena) {
    const HostJob &h = b->hj[j];
    SGT_CHECK(j >= 0 && j < b->max_jobs && h.live, SGTSNE_ERR_STATE, "arena lists job %d, not live", j);
    SGT_CHECK(h.off >= 0 && h.n > 0 && (int64_t)h.off + h.n <= b->max_points, SGTSNE_ERR_STATE,
              "job %d: points [%d, %d) outside the arena of %lld", j, h.off, h.off + h.n, (long long)b->max_points);
    SGT_CHECK(h.nnz_off >= 0 && h.nnz >= 0 && h.nnz_off + h.nnz <= b->max_nnz, SGTSNE_ERR_STATE,
              "job %d: nonzeros outside the arena", j);
    pr.push_back({h.off, (int64_t)h.off + h.n});
    if (h.nnz > 0) zr.push_back({h.nnz_off, h.nnz_off + h.nnz});
    const BatchJob &J = b->jobs_h[j];
    for (int i = J.item0; i < J.item0 + J.nitems; ++i) {
      SGT_CHECK(i >= 0 && i < (int)items.size(), SGTSNE_ERR_STATE, "job %d: item %d outside the table", j, i);
      const BatchItem &it = items[i];
      SGT_CHECK(it.job == j && it.begin >= h.off && it.begin < it.end && it.end <= h.off + h.n &&
                    it.end - it.begin <= kBatchChunk,
                SGTSNE_ERR_STATE, "job %d: item %d [%d, %d) outside its points", j, i, it.begin, it.end);
    }
  }
  for (auto *v : {&pr, &zr}) {
    std::sort(v->begin(), v->end());
    for (size_t i = 1; i < v->size(); ++i)
      SGT_CHECK((*v)[i - 1].second <= (*v)[i].first, SGTSNE_ERR_STATE, "two running jobs share arena memory");
  }
  return SGTSNE_OK;
}

int rebuild_items(sgtsne_batch *b) {
  std::vector<BatchItem> items;
  for (int j : b->arena) {
    HostJob &h = b->hj[j];
    BatchJob &J = b->jobs_h[j];
    J.item0 = (int)items.size();
    for (int s = 0; s < h.n; s += kBatchChunk)
      items.push_back(BatchItem{j, h.off + s, h.off + std::min(h.n, s + kBatchChunk), 0});
    J.nitems = (int)items.size() - J.item0;
    J.off = h.off;
    J.n = h.n;
  }
  SGT_CHECK((int)items.size() <= b->items_cap, SGTSNE_ERR_ARG, "work items exceed capacity");
  SGT_TRY(validate_tables(b, items));
  b->nitems = (int)items.size();
  // stream-ordered after the iterations already queued, which still read the
  // old table; staged, so nothing here waits for them
  SGT_TRY(upload(b, b->items_d, items.data(), items.size() * sizeof(BatchItem)));
  for (int j : b->arena) SGT_TRY(write_static(b, j));
  SGT_TRY(stage_commit(b));
  invalidate_graph(b);
  return SGTSNE_OK;
}

int attractive(sgtsne_batch *b) {
  const int n = (int)b->used_points;
  if (n == 0) return SGTSNE_OK;
  if (b->spmv_win > 0)
    return launch_attr_csr_win(b->stream, b->d, b->spmv_group, b->spmv_unroll, b->spmv_rpb, b->spmv_win, n,
                               b->rowptr, b->col, b->val, b->Y, b->Fattr);
  if (b->spmv_rpb > 0)
    return launch_attr_csr_blk(b->stream, b->d, b->spmv_group, b->spmv_unroll, b->spmv_rpb, n, b->rowptr, b->col,
                               b->val, b->Y, b->Fattr);
  return launch_attr_csr_seg2(b->stream, b->d, b->spmv_group, b->spmv_unroll, n, b->rowptr, b->col, b->val, b->Y,
                              b->Fattr);
}

// The gradient terms of every running job at its current Y (no update).
int enqueue_terms(sgtsne_batch *b) {
  const int zd = b->p.zeta_mode == SGTSNE_ZETA_DIRECT;
  SGT_TRY(attractive(b));
  SGT_TRY(launch_bs2g(b->stream, b->d, b->jobs_d, b->items_d, b->nitems, b->Y,
                      std::min(b->nitems * (b->d == 3 ? 4 : 1), b->s2g_blocks), b->d == 3 ? 64512 : b->s2g_lds));
  SGT_TRY(launch_bconv(b->stream, b->d, zd, b->args_d, b->prefix_d, b->max_jobs, b->conv_blocks));
  SGT_TRY(launch_bg2s(b->stream, b->d, zd, b->jobs_d, b->items_d, b->nitems, b->Y, b->Frep, b->zpart,
                      std::min(b->nitems, b->pt_blocks)));
  return SGTSNE_OK;
}

int enqueue_iteration(sgtsne_batch *b) {
  const int blocks = std::min(b->nitems, b->pt_blocks);
  SGT_TRY(enqueue_terms(b));
  SGT_TRY(launch_bupdate(b->stream, b->d, b->jobs_d, b->items_d, b->nitems, b->Y, b->uY, b->gains, b->Fattr,
                         b->Frep, b->zpart, b->spart, blocks));
  SGT_TRY(launch_bcenter(b->stream, b->d, b->jobs_d, b->items_d, b->nitems, b->Y, b->spart, b->mpart, blocks));
  SGT_TRY(launch_bplan(b->stream, b->d, b->jobs_d, b->max_jobs, b->mpart, b->args_d, b->prefix_d, b->tw));
  return SGTSNE_OK;
}

// Captures one iteration.  When an executable graph exists, it is updated
// from the new capture (same kernels, new arguments: much cheaper than a new
// instantiation); only a refused update re-instantiates.
int capture(sgtsne_batch *b) {
  b->graph_valid = false;
  if (!b->use_graph || b->nitems == 0) return SGTSNE_OK;
  SGT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int st = enqueue_iteration(b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (st != SGTSNE_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess || g == nullptr) {
    set_error("graph capture failed: %s", hipGetErrorString(e));
    return SGTSNE_ERR_HIP;
  }
  // the other executable graph: its launches all precede the current one's,
  // so waiting for its last one rarely waits at all
  const int k = b->exec ? 1 - b->cur : b->cur;
  if (b->execs[k] && b->exec_done[k]) SGT_HIP(hipEventSynchronize(b->exec_done[k]));
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph = g;
  if (b->execs[k]) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
    if (hipGraphExecUpdate(b->execs[k], g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
      b->cur = k;
      b->exec = b->execs[k];
      b->graph_updates++;
      b->graph_valid = true;
      return SGTSNE_OK;
    }
    (void)hipGetLastError();
    (void)hipGraphExecDestroy(b->execs[k]);  // no launch of it is pending (waited above)
    b->execs[k] = nullptr;
  }
  SGT_HIP(hipGraphInstantiate(&b->execs[k], g, nullptr, nullptr, 0));
  b->cur = k;
  b->exec = b->execs[k];
  b->graph_instantiations++;
  b->graph_valid = true;
  return SGTSNE_OK;
}

int run_iterations(sgtsne_batch *b, int64_t count) {
  if (count <= 0 || b->nitems == 0) return SGTSNE_OK;
  if (b->use_graph && !b->graph_valid) SGT_TRY(capture(b));
  for (int64_t k = 0; k < count; ++k) {
    if (b->graph_valid) {
      SGT_HIP(hipGraphLaunch(b->exec, b->stream));
      b->graph_replays++;
    } else {
      SGT_TRY(enqueue_iteration(b));
    }
  }
  if (b->graph_valid) {
    hipEvent_t &ev = b->exec_done[b->cur];
    if (!ev) SGT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    SGT_HIP(hipEventRecord(ev, b->stream));
  }
  for (int j : b->arena) {
    HostJob &h = b->hj[j];
    const int64_t end = (int64_t)h.max_iter + 1;
    const int64_t t0 = h.t;
    h.t = std::min<int64_t>(h.t + count, end);
    if (t0 < end && h.t >= end && h.done_ev) {
      SGT_HIP(hipEventRecord(h.done_ev, b->stream));
      h.done_marked = true;
    }
  }
  return SGTSNE_OK;
}

// Attractive-kernel shape for the current arena.  The lanes per row stay
// fixed (the engine's spmv_group): every candidate then sums each row in the
// same order, so the choice never changes a result; only rows per
// work-group and unroll vary.
int tune(sgtsne_batch *b) {
  if (b->used_points < 4096) return SGTSNE_OK;
  struct Cand { int unroll, rpb, win; };
  const Cand cands[] = {{1, 0, 0},       {2, 0, 0},       {4, 0, 0},       {8, 0, 0},      {16, 0, 0},     {2, 32, 0},
                        {2, 64, 0},
                        {1, 64, 0},      {2, 128, 0},     {1, 128, 0},     {2, 256, 0},    {1, 256, 4096},
                        {2, 512, 4096},  {1, 512, 8192},  {2, 1024, 8192}, {16, 64, 0},     {16, 256, 0},
                        {8, 256, 0}};
  hipEvent_t e0, e1;
  SGT_HIP(hipEventCreate(&e0));
  SGT_HIP(hipEventCreate(&e1));
  b->tune_ms.clear();
  float best = 1e30f;
  Cand pick{b->spmv_unroll, b->spmv_rpb, b->spmv_win};
  const Cand keep = pick;
  int st = SGTSNE_OK;
  for (const Cand &c : cands) {
    if (c.win > 0 && (size_t)c.win * b->ds * sizeof(float) > 65536) continue;  // window must fit LDS
    b->spmv_unroll = c.unroll;
    b->spmv_rpb = c.rpb;
    b->spmv_win = c.win;
    st = attractive(b);  // warm
    if (st) break;
    (void)hipEventRecord(e0, b->stream);
    for (int r = 0; r < 3 && !st; ++r) st = attractive(b);
    (void)hipEventRecord(e1, b->stream);
    if (st) break;
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= 3.0f;
    b->tune_ms.push_back(ms);
    if (ms < best) {
      best = ms;
      pick = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) {
    b->spmv_unroll = keep.unroll;
    b->spmv_rpb = keep.rpb;
    b->spmv_win = keep.win;
    return st;
  }
  b->spmv_unroll = pick.unroll;
  b->spmv_rpb = pick.rpb;
  b->spmv_win = pick.win;
  b->tuned_points = b->used_points;
  invalidate_graph(b);
  return SGTSNE_OK;
}

// Moves the live jobs down over the holes (stream-ordered, through a staging
// buffer so that overlapping moves are safe).
int compact(sgtsne_batch *b) {
  int64_t cur_p = 0, cur_z = 0;
  size_t need = 0;
  for (int j : b->arena) {
    const HostJob &h = b->hj[j];
    need = std::max(need, std::max((size_t)h.n * b->ds * sizeof(float), (size_t)(h.nnz + h.n + 1) * sizeof(int32_t)));
  }
  SGT_TRY(ws_reserve(b, need));
  void *tmp = b->ws;
  for (int j : b->arena) {
    HostJob &h = b->hj[j];
    if (h.off != cur_p || h.nnz_off != cur_z) {
      if (h.done_marked) h.moved_after_done = true;
      const size_t pb = (size_t)h.n * b->ds * sizeof(float);
      float *arr[3] = {b->Y, b->uY, b->gains};
      for (float *a : arr) {
        SGT_HIP(hipMemcpyAsync(tmp, a + (size_t)h.off * b->ds, pb, hipMemcpyDeviceToDevice, b->stream));
        SGT_HIP(hipMemcpyAsync(a + (size_t)cur_p * b->ds, tmp, pb, hipMemcpyDeviceToDevice, b->stream));
      }
      int32_t *ti = static_cast<int32_t *>(tmp);
      SGT_HIP(hipMemcpyAsync(ti, b->perm + h.off, (size_t)h.n * 4, hipMemcpyDeviceToDevice, b->stream));
      SGT_HIP(hipMemcpyAsync(b->perm + cur_p, ti, (size_t)h.n * 4, hipMemcpyDeviceToDevice, b->stream));
      SGT_HIP(hipMemcpyAsync(ti, b->rowptr + h.off, (size_t)(h.n + 1) * 4, hipMemcpyDeviceToDevice, b->stream));
      SGT_TRY(launch_offset_copy_i32(b->stream, ti, b->rowptr + cur_p, h.n + 1, (int32_t)(cur_z - h.nnz_off)));
      SGT_HIP(hipMemcpyAsync(ti, b->col + h.nnz_off, (size_t)h.nnz * 4, hipMemcpyDeviceToDevice, b->stream));
      SGT_TRY(launch_offset_copy_i32(b->stream, ti, b->col + cur_z, h.nnz, (int32_t)(cur_p - h.off)));
      SGT_HIP(hipMemcpyAsync(tmp, b->val + h.nnz_off, (size_t)h.nnz * 4, hipMemcpyDeviceToDevice, b->stream));
      SGT_HIP(hipMemcpyAsync(b->val + cur_z, tmp, (size_t)h.nnz * 4, hipMemcpyDeviceToDevice, b->stream));
      h.off = (int)cur_p;
      h.nnz_off = cur_z;
    }
    cur_p += h.n;
    cur_z += h.nnz;
  }
  b->used_points = cur_p;
  b->used_nnz = cur_z;
  return rebuild_items(b);
}

// Morton re-order of one job's points (its rows of P, its Y / uY / gains and
// the device->caller map), entirely on the device.
int reorder_job(sgtsne_batch *b, int j) {
  HostJob &h = b->hj[j];
  const int n = h.n;
  const int64_t nnz = h.nnz;
  const hipStream_t s = b->stream;
  size_t tb = 0;
  SGT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (int32_t *)nullptr, (int32_t *)nullptr, n, 0, 32, s));
  uint32_t *k0, *k1;
  int32_t *i0, *order, *inv, *rpl, *coll, *rp2, *col2, *itmp;
  float *val2, *ftmp;
  void *tmp;
  for (int pass = 0; pass < 2; ++pass) {  // sizes, then the pieces of the scratch
    Carve c{pass ? b->ws : nullptr};
    k0 = c.take<uint32_t>(n);
    k1 = c.take<uint32_t>(n);
    i0 = c.take<int32_t>(n);
    order = c.take<int32_t>(n);
    inv = c.take<int32_t>(n);
    tmp = c.take<char>(tb);
    rpl = c.take<int32_t>((size_t)n + 1);
    rp2 = c.take<int32_t>((size_t)n + 1);
    coll = c.take<int32_t>(nnz);
    col2 = c.take<int32_t>(nnz);
    val2 = c.take<float>(nnz);
    ftmp = c.take<float>((size_t)n * b->ds);
    itmp = c.take<int32_t>(n);
    if (pass == 0) SGT_TRY(ws_reserve(b, c.off));
  }
  SGT_TRY(launch_bmorton(s, b->d, b->jobs_d, j, h.off, n, b->Y, k0, i0));
  SGT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, order, n, 0, 32, s));
  SGT_TRY(launch_invert(s, n, order, inv));
  // the job's rows as a local CSR, permuted, written back with offsets
  SGT_TRY(launch_offset_copy_i32(s, b->rowptr + h.off, rpl, n + 1, (int32_t)(-h.nnz_off)));
  SGT_TRY(launch_offset_copy_i32(s, b->col + h.nnz_off, coll, nnz, -h.off));
  // rows re-sorted on the device alone: a host round trip here would wait
  // for every iteration queued ahead of the re-order
  SGT_TRY(launch_permute_csr(s, n, nnz, order, inv, rpl, coll, b->val + h.nnz_off, rp2, col2, val2, true, true));
  SGT_TRY(launch_offset_copy_i32(s, rp2, b->rowptr + h.off, n + 1, (int32_t)h.nnz_off));
  SGT_TRY(launch_offset_copy_i32(s, col2, b->col + h.nnz_off, nnz, h.off));
  if (nnz > 0)
    SGT_HIP(hipMemcpyAsync(b->val + h.nnz_off, val2, sizeof(float) * nnz, hipMemcpyDeviceToDevice, s));
  // point state
  float *arr[3] = {b->Y, b->uY, b->gains};
  for (float *a : arr) {
    SGT_TRY(launch_gather_points(s, n, b->ds, order, a + (size_t)h.off * b->ds, ftmp));
    SGT_HIP(hipMemcpyAsync(a + (size_t)h.off * b->ds, ftmp, sizeof(float) * (size_t)n * b->ds,
                           hipMemcpyDeviceToDevice, s));
  }
  SGT_TRY(launch_gather_i32(s, n, order, b->perm + h.off, itmp));
  SGT_HIP(hipMemcpyAsync(b->perm + h.off, itmp, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  h.reorders++;
  return SGTSNE_OK;
}

// Iterations until the next job reaches a re-order point (or `limit`).
int64_t next_event(const sgtsne_batch *b, int64_t limit) {
  int64_t best = limit;
  if (b->p.reorder < 2) return best;
  for (int j : b->arena) {
    const HostJob &h = b->hj[j];
    for (int r : kReorderIters) {
      const int64_t te = (int64_t)r + 1;  // before the job's iteration r runs
      if (r < h.max_iter && te > h.t) {
        best = std::min(best, te - h.t);
        break;
      }
    }
  }
  return best;
}

int do_events(sgtsne_batch *b) {
  if (b->p.reorder < 2) return SGTSNE_OK;
  for (int j : b->arena) {
    const HostJob &h = b->hj[j];
    for (int r : kReorderIters)
      if (r < h.max_iter && h.t == (int64_t)r + 1) SGT_TRY(reorder_job(b, j));
  }
  return SGTSNE_OK;
}

int check_batch(sgtsne_batch *b) {
  SGT_CHECK(b != nullptr, SGTSNE_ERR_ARG, "batch is NULL");
  return SGTSNE_OK;
}

int check_job(sgtsne_batch *b, int job) {
  SGT_TRY(check_batch(b));
  SGT_CHECK(job >= 0 && job < b->max_jobs && b->hj[job].live, SGTSNE_ERR_ARG, "no job %d", job);
  return SGTSNE_OK;
}

// Places a job whose symmetric P (local CSR on the device, rows starting at 0)
// and initial Y (host, caller order, layout per params) are given.
int admit(sgtsne_batch *b, int n, int64_t nnz, const int32_t *rp_d, const int32_t *col_d, const float *val_d,
          const float *y_in, int max_iter, int ng_cap, int *job_out) {
  SGT_CHECK(n > 0 && nnz >= 0, SGTSNE_ERR_ARG, "bad job size");
  int slot = -1;
  for (int j = 0; j < b->max_jobs; ++j)
    if (!b->hj[j].live) { slot = j; break; }
  SGT_CHECK(slot >= 0, SGTSNE_ERR_FULL, "all %d job slots are in use", b->max_jobs);
  int64_t live_p = 0, live_z = 0;
  for (int j : b->arena) { live_p += b->hj[j].n; live_z += b->hj[j].nnz; }
  SGT_CHECK(live_p + n <= b->max_points && live_z + nnz <= b->max_nnz, SGTSNE_ERR_FULL,
            "arena full (%lld + %d points, %lld + %lld nnz)", (long long)live_p, n, (long long)live_z,
            (long long)nnz);
  if (ng_cap <= 0) ng_cap = auto_ng_cap(b->d, n);
  if (b->p.grid_mode == SGTSNE_GRID_FIXED) ng_cap = std::max(ng_cap, b->p.ng + 2);
  SGT_CHECK(ng_cap >= 16 && ng_cap <= 512, SGTSNE_ERR_ARG, "grid capacity %d outside [16, 512]", ng_cap);
  size_t vb, pb, wb, kb;
  const size_t gbytes = batch_grid_bytes(b->d, ng_cap, &vb, &pb, &wb, &kb);
  SGT_CHECK(b->grid_used + gbytes <= b->grid_budget, SGTSNE_ERR_FULL, "grid budget exhausted");
  if (b->used_points + n > b->max_points || b->used_nnz + nnz > b->max_nnz) SGT_TRY(compact(b));

  HostJob &h = b->hj[slot];
  h = HostJob{};
  h.off = (int)b->used_points;
  h.n = n;
  h.nnz_off = b->used_nnz;
  h.nnz = nnz;
  h.max_iter = max_iter;
  h.ng_cap = ng_cap;
  h.gbytes = gbytes;
  SGT_HIP(hipMallocAsync(&h.gmem, gbytes, b->stream));
  SGT_HIP(hipMemsetAsync(h.gmem, 0, vb, b->stream));  // the spread grid starts (and stays) zero
  char *gm = static_cast<char *>(h.gmem);

  // P rows into the arena (global row offsets and column indices)
  SGT_TRY(launch_offset_copy_i32(b->stream, rp_d, b->rowptr + h.off, n + 1, (int32_t)h.nnz_off));
  SGT_TRY(launch_offset_copy_i32(b->stream, col_d, b->col + h.nnz_off, nnz, h.off));
  if (nnz > 0)
    SGT_HIP(hipMemcpyAsync(b->val + h.nnz_off, val_d, sizeof(float) * nnz, hipMemcpyDeviceToDevice, b->stream));

  // Y (caller order -> AoS stride ds), centred by the admission mean
  const int d = b->d, ds = b->ds;
  std::vector<float> yh((size_t)n * ds, 0.0f);
  std::vector<float> y0;
  const float *yi = y_in;
  if (!yi) {
    std::mt19937_64 gen(b->p.seed + 0x9e3779b97f4a7c15ull * (uint64_t)(slot + 1));
    std::normal_distribution<double> nd(0.0, 1.0);
    y0.resize((size_t)n * d);
    for (auto &v : y0) v = (float)(1e-4 * nd(gen));
    yi = y0.data();
  }
  const bool soa = b->p.layout == SGTSNE_LAYOUT_SOA && y_in != nullptr;
  double mean[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < d; ++a) {
      const float v = soa ? yi[(size_t)a * n + i] : yi[(size_t)i * d + a];
      yh[(size_t)i * ds + a] = v;
      mean[a] += v;
    }
  BatchJob &J = b->jobs_h[slot];
  J = BatchJob{};
  for (int a = 0; a < 4; ++a) J.mean0[a] = a < d ? (float)(mean[a] / n) : 0.0f;
  SGT_TRY(upload(b, b->Y + (size_t)h.off * ds, yh.data(), yh.size() * sizeof(float)));
  SGT_TRY(launch_fill(b->stream, b->uY + (size_t)h.off * ds, (int64_t)n * ds, 0.0f));
  SGT_TRY(launch_fill(b->stream, b->gains + (size_t)h.off * ds, (int64_t)n * ds, 1.0f));
  SGT_TRY(launch_iota(b->stream, n, b->perm + h.off));

  J.off = h.off;
  J.n = n;
  J.max_iter = max_iter;
  J.early_iter = b->p.early_iter;
  J.grid_mode = b->p.grid_mode;
  J.ng_fixed = b->p.ng;
  J.ng_cap = ng_cap;
  J.zdirect = b->p.zeta_mode == SGTSNE_ZETA_DIRECT;
  J.alpha0 = (float)b->p.alpha;
  J.h = b->p.h > 0 ? b->p.h : bdefault_h(d);
  J.vfix = reinterpret_cast<long long *>(gm);
  J.phi = reinterpret_cast<float *>(gm + vb);
  J.W = reinterpret_cast<float2 *>(gm + vb + pb);
  J.Kt1 = reinterpret_cast<float2 *>(gm + vb + pb + wb);
  J.Kt2 = reinterpret_cast<float2 *>(gm + vb + pb + wb + kb);
  J.Kh = reinterpret_cast<float2 *>(gm + vb + pb + wb + 2 * kb);
  J.iter = 0;
  J.status = BJ_INIT;
  SGT_TRY(upload(b, b->jobs_d + slot, &J, sizeof(BatchJob)));
  std::memcpy(&b->jobs_up[slot], &J, kJobStatic);
  if (!h.done_ev) SGT_HIP(hipEventCreateWithFlags(&h.done_ev, hipEventDisableTiming));

  h.live = true;
  b->used_points += n;
  b->used_nnz += nnz;
  b->grid_used += gbytes;
  // rowptr[used_points] is this job's end: the next job starts there
  b->arena.push_back(slot);
  SGT_TRY(rebuild_items(b));  // commits the staged uploads above too
  if (job_out) *job_out = slot;
  return SGTSNE_OK;
}

void fe_release(sgtsne_batch *b, FrontEnd &f);

void destroy_all(sgtsne_batch *b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->prep) (void)hipStreamSynchronize(b->prep);
  for (FrontEnd &f : b->fe) {
    if (f.stage != 0) fe_release(b, f);
    if (f.ev) (void)hipEventDestroy(f.ev);
    f.ev = nullptr;
  }
  release_graph(b);
  release_deferred(b, true);
  for (hipEvent_t &e : b->exec_done)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  // per-job grids come from the stream-ordered pool: release them the same way
  for (auto &h : b->hj)
    if (h.gmem) {
      (void)hipFreeAsync(h.gmem, b->stream);
      h.gmem = nullptr;
    }
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  bfree(b->rowptr);
  bfree(b->col);
  bfree(b->val);
  bfree(b->Y);
  bfree(b->uY);
  bfree(b->gains);
  bfree(b->Fattr);
  bfree(b->Frep);
  bfree(b->perm);
  bfree(b->jobs_d);
  bfree(b->items_d);
  bfree(b->zpart);
  bfree(b->spart);
  bfree(b->mpart);
  bfree(b->args_d);
  bfree(b->prefix_d);
  bfree(b->tw);
  bfree(b->scal_d);
  bfree(b->ws);
  b->ws_bytes = 0;
  if (b->prep) (void)hipStreamSynchronize(b->prep);
  if (b->copy) (void)hipStreamSynchronize(b->copy);
  for (auto &h : b->hj)
    if (h.done_ev) {
      (void)hipEventDestroy(h.done_ev);
      h.done_ev = nullptr;
    }
  for (auto &sp : b->stage_live)
    if (sp.ev) b->stage_free.push_back(sp.ev);
  b->stage_live.clear();
  std::sort(b->stage_free.begin(), b->stage_free.end());  // one event may guard several spans
  b->stage_free.erase(std::unique(b->stage_free.begin(), b->stage_free.end()), b->stage_free.end());
  for (hipEvent_t e : b->stage_free) (void)hipEventDestroy(e);
  b->stage_free.clear();
  if (b->stage) (void)hipHostFree(b->stage);
  b->stage = nullptr;
  for (hipEvent_t &e : b->marks)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  if (b->fe_host) (void)hipHostFree(b->fe_host);
  b->fe_host = nullptr;
  if (b->prep) (void)hipStreamDestroy(b->prep);
  b->copy = nullptr;  // the front-end stream, destroyed above
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

// Reads one job's rows of an arena array back in the caller's order.
int fetch_points(sgtsne_batch *b, int job, const float *src, float *out, int layout, double scale = 1.0) {
  const HostJob &h = b->hj[job];
  const int n = h.n, d = b->d, ds = b->ds;
  std::vector<float> v((size_t)n * ds);
  std::vector<int32_t> pm(n);
  // a finished job's rows no longer change: read them right after its last
  // iteration instead of after everything queued since
  const bool fast = src == b->Y && h.done_marked && !h.moved_after_done;
  const hipStream_t cs = fast ? b->copy : b->stream;
  if (fast) SGT_HIP(hipStreamWaitEvent(b->copy, h.done_ev, 0));
  SGT_HIP(hipMemcpyAsync(v.data(), src + (size_t)h.off * ds, v.size() * sizeof(float), hipMemcpyDeviceToHost, cs));
  SGT_HIP(hipMemcpyAsync(pm.data(), b->perm + h.off, sizeof(int32_t) * n, hipMemcpyDeviceToHost, cs));
  SGT_HIP(hipStreamSynchronize(cs));
  for (int r = 0; r < n; ++r)
    SGT_CHECK(pm[r] >= 0 && pm[r] < n, SGTSNE_ERR_STATE, "job %d: point order corrupt at row %d", job, r);
  for (int r = 0; r < n; ++r) {
    const int i = pm[r];
    for (int a = 0; a < d; ++a) {
      const float x = scale == 1.0 ? v[(size_t)r * ds + a] : (float)(v[(size_t)r * ds + a] * scale);
      if (layout == SGTSNE_LAYOUT_SOA) out[(size_t)a * n + i] = x;
      else out[(size_t)i * d + a] = x;
    }
  }
  return SGTSNE_OK;
}

// Frees the slot for the next front end (its event and pinned counts stay).
void fe_clear(FrontEnd &f) {
  hipEvent_t ev = f.ev;
  int *counts = f.counts;
  f = FrontEnd{};
  f.ev = ev;
  f.counts = counts;
}

// Releases a front end's device temporaries (stream-ordered on the
// front-end stream, after whatever of it is queued) and clears it.
void fe_release(sgtsne_batch *b, FrontEnd &f) {
  const hipStream_t s = b->prep;
  knn_release(s, &f.knn);
  sym_release(s, &f.sym);
  void *bufs[] = {f.Xf, f.Xd, f.Dd, f.Id, f.rp, f.cc, f.vv, f.srp, f.scc, f.svv};
  for (void *q : bufs)
    if (q) (void)hipFreeAsync(q, s);
  fe_clear(f);
}

// Queues the front end's next part; the device has passed the previous one.
// Returns 1 once the job is admitted (*job set), 0 when a further part is
// queued, or an error status (the front end is then released).
int fe_advance(sgtsne_batch *b, FrontEnd &f, int *job) {
  const hipStream_t s = b->prep;
  int st = SGTSNE_OK;
  if (f.stage == 1) {  // kNN rows certified: the uncertified ones, P, and the merge of P and P^T
    st = knn_end(s, &f.knn, nullptr);
    if (!st) st = launch_perplexity(s, f.Dd, f.n, f.k, f.u, f.vv, nullptr);
    if (!st) st = launch_knn_csr(s, f.n, f.k, f.Id, f.rp, f.cc);
    if (!st && b->p.lambda != 1.0) st = launch_lambda_rescale(s, f.n, f.rp, f.vv, b->p.lambda, b->p.drop_leaf);
    if (!st) st = sym_begin(s, f.n, (int64_t)f.n * f.k, f.rp, f.cc, f.vv, f.counts + 1, &f.sym);
  } else if (f.stage == 2) {  // merged count known: P_sym's CSR
    st = sym_end(s, &f.sym, &f.srp, &f.scc, &f.svv, &f.m, true);
    // released in stream order, after the kernels just queued that read them
    void *tmp[] = {f.Xf, f.Xd, f.Dd, f.Id, f.rp, f.cc, f.vv};
    for (void *q : tmp)
      if (q) (void)hipFreeAsync(q, s);
    f.Xf = f.vv = nullptr;
    f.Xd = f.Dd = nullptr;
    f.Id = f.rp = f.cc = nullptr;
  } else {  // P_sym built: admission
    std::lock_guard<std::mutex> lock(b->mu);
    release_deferred(b);
    st = admit(b, f.n, f.m, f.srp, f.scc, f.svv, f.y_in, f.max_iter, f.ng_cap, job);
    if (st) {
      fe_release(b, f);
      return st;
    }
    // the admitted P is read by copies queued on the iteration stream
    st = defer_free(b, f.srp, f.scc, f.svv);
    fe_clear(f);
    return st ? st : 1;
  }
  if (!st && hipEventRecord(f.ev, s) != hipSuccess) {
    set_error("front end: event record failed");
    st = SGTSNE_ERR_HIP;
  }
  if (st) {
    fe_release(b, f);
    return st;
  }
  ++f.stage;
  return 0;
}

}  // namespace

extern "C" {

int sgtsne_batch_create(sgtsne_batch **out, const sgtsne_params *params, int max_jobs, int64_t max_points,
                        int64_t max_nnz, int64_t grid_budget_bytes) {
  SGT_CHECK(out && params, SGTSNE_ERR_ARG, "NULL argument");
  *out = nullptr;
  SGT_CHECK(params->d >= 1 && params->d <= 3, SGTSNE_ERR_ARG, "d must be 1..3");
  SGT_CHECK(max_jobs >= 1 && max_jobs <= 1024, SGTSNE_ERR_ARG, "max_jobs must be 1..1024");
  SGT_CHECK(max_points >= 1 && max_points < (int64_t)INT32_MAX / 4, SGTSNE_ERR_ARG, "bad max_points");
  SGT_CHECK(max_nnz >= 0 && max_nnz < (int64_t)INT32_MAX, SGTSNE_ERR_ARG, "max_nnz exceeds the int32 range");
  SGT_CHECK(params->zeta_mode == SGTSNE_ZETA_DIRECT || params->zeta_mode == SGTSNE_ZETA_ALGEBRAIC, SGTSNE_ERR_ARG,
            "bad zeta_mode");
  SGT_CHECK(params->grid_mode == SGTSNE_GRID_H || params->grid_mode == SGTSNE_GRID_FIXED, SGTSNE_ERR_ARG,
            "bad grid_mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no GPU visible");
    return SGTSNE_ERR_NODEV;
  }
  sgtsne_batch *b = new sgtsne_batch();
  b->p = *params;
  b->d = params->d;
  b->ds = point_stride(params->d);
  b->spmv_group = 32;  // lanes per row, fixed for the engine's life (16 measured 7 % slower at d = 2)
  b->max_jobs = max_jobs;
  b->max_points = max_points;
  b->max_nnz = max_nnz;
  b->grid_budget = grid_budget_bytes > 0 ? (size_t)grid_budget_bytes : ((size_t)64 << 30);
  b->hj.resize(max_jobs);
  b->jobs_h.resize(max_jobs);
  b->jobs_up.resize(max_jobs);  // zero, as the device table is cleared below
  b->items_cap = (int)(max_points / kBatchChunk + max_jobs + 1);
  const int np = batch_npasses(b->d);
  int st = SGTSNE_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (st == SGTSNE_OK && e != hipSuccess) {
      set_error("%s failed: %s", what, hipGetErrorString(e));
      st = SGTSNE_ERR_HIP;
    }
  };
  retain_pool_memory();
  chk(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
  chk(hipStreamCreateWithFlags(&b->prep, hipStreamNonBlocking), "stream");
  b->copy = b->prep;
  for (hipEvent_t &e : b->marks) chk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
  chk(hipHostMalloc((void **)&b->stage, kStageBytes, hipHostMallocDefault), "pinned staging");
  chk(hipHostMalloc((void **)&b->fe_host, sizeof(int) * 2 * kFrontEnds, hipHostMallocDefault), "pinned counts");
  for (int i = 0; i < kFrontEnds; ++i) {
    chk(hipEventCreateWithFlags(&b->fe[i].ev, hipEventDisableTiming), "event");
    b->fe[i].counts = b->fe_host ? b->fe_host + 2 * i : nullptr;
  }
  const size_t P = (size_t)max_points;
  if (!st) st = balloc((void **)&b->rowptr, sizeof(int32_t) * (P + 1));
  if (!st) st = balloc((void **)&b->col, sizeof(int32_t) * std::max<int64_t>(max_nnz, 1));
  if (!st) st = balloc((void **)&b->val, sizeof(float) * std::max<int64_t>(max_nnz, 1));
  float **pts[5] = {&b->Y, &b->uY, &b->gains, &b->Fattr, &b->Frep};
  for (float **q : pts)
    if (!st) st = balloc((void **)q, sizeof(float) * P * b->ds);
  if (!st) st = balloc((void **)&b->perm, sizeof(int32_t) * P);
  if (!st) st = balloc((void **)&b->jobs_d, sizeof(BatchJob) * max_jobs);
  if (!st) st = balloc((void **)&b->items_d, sizeof(BatchItem) * b->items_cap);
  if (!st) st = balloc((void **)&b->zpart, sizeof(double) * 2 * b->items_cap);
  if (!st) st = balloc((void **)&b->spart, sizeof(double) * 4 * b->items_cap);
  if (!st) st = balloc((void **)&b->mpart, sizeof(double) * 8 * b->items_cap);
  if (!st) st = balloc(&b->args_d, batch_pass_args_bytes() * np * max_jobs);
  if (!st) st = balloc((void **)&b->prefix_d, sizeof(int) * np * (max_jobs + 1));
  if (!st) st = balloc((void **)&b->tw, twiddle_count() * sizeof(float2));
  if (!st) st = balloc((void **)&b->scal_d, 8 * sizeof(double));
  if (!st) chk(hipMemsetAsync(b->rowptr, 0, sizeof(int32_t) * (P + 1), b->stream), "memset");
  if (!st) chk(hipMemsetAsync(b->col, 0, sizeof(int32_t) * std::max<int64_t>(max_nnz, 1), b->stream), "memset");
  if (!st) chk(hipMemsetAsync(b->val, 0, sizeof(float) * std::max<int64_t>(max_nnz, 1), b->stream), "memset");
  for (float **q : pts)
    if (!st) chk(hipMemsetAsync(*q, 0, sizeof(float) * P * b->ds, b->stream), "memset");
  if (!st) chk(hipMemsetAsync(b->jobs_d, 0, sizeof(BatchJob) * max_jobs, b->stream), "memset");
  if (!st) chk(hipMemsetAsync(b->prefix_d, 0, sizeof(int) * np * (max_jobs + 1), b->stream), "memset");
  if (!st) st = init_twiddles(b->stream, b->tw);
  if (!st) chk(hipStreamSynchronize(b->stream), "create");
  if (st) {
    destroy_all(b);
    delete b;
    return st;
  }
  *out = b;
  return SGTSNE_OK;
}

int sgtsne_batch_destroy(sgtsne_batch *b) {
  if (!b) return SGTSNE_OK;
  destroy_all(b);
  delete b;
  return SGTSNE_OK;
}

int sgtsne_batch_add(sgtsne_batch *b, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                     const float *val, int symmetrize, const float *y_in, int max_iter, int ng_cap, int *job) {
  SGT_TRY(check_batch(b));
  SGT_CHECK(n > 0 && nnz >= 0 && rowptr && (nnz == 0 || (col && val)), SGTSNE_ERR_ARG, "bad P arguments");
  SGT_CHECK(nnz < (int64_t)INT32_MAX / 2, SGTSNE_ERR_ARG, "nnz exceeds the int32 index range");
  SGT_CHECK(rowptr[0] == 0 && rowptr[n] == nnz, SGTSNE_ERR_ARG, "rowptr inconsistent with nnz");
  for (int i = 0; i < n; i++) SGT_CHECK(rowptr[i] <= rowptr[i + 1], SGTSNE_ERR_ARG, "rowptr decreases at row %d", i);
  {
    int64_t bad = 0;
    for (int64_t t = 0; t < nnz; t++) bad += (col[t] < 0 || col[t] >= n);
    SGT_CHECK(bad == 0, SGTSNE_ERR_ARG, "%lld column indices out of range [0, %d)", (long long)bad, n);
  }
  if (max_iter <= 0) max_iter = b->p.max_iter;
  const hipStream_t s = b->prep;
  int32_t *rp = nullptr, *cc = nullptr;
  float *vv = nullptr;
  // temporaries come from the stream-ordered pool: no release here waits for
  // the iterations queued on the other stream
  SGT_HIP(hipMallocAsync((void **)&rp, sizeof(int32_t) * ((size_t)n + 1), s));
  SGT_HIP(hipMallocAsync((void **)&cc, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1), s));
  SGT_HIP(hipMallocAsync((void **)&vv, sizeof(float) * (size_t)std::max<int64_t>(nnz, 1), s));
  int st = SGTSNE_OK;
  if (hipMemcpyAsync(rp, rowptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      (nnz > 0 && (hipMemcpyAsync(cc, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s) != hipSuccess ||
                   hipMemcpyAsync(vv, val, sizeof(float) * nnz, hipMemcpyHostToDevice, s) != hipSuccess))) {
    set_error("copy of P failed");
    st = SGTSNE_ERR_HIP;
  }
  int32_t *srp = rp, *scc = cc;
  float *svv = vv;
  int64_t m = nnz;
  if (!st && symmetrize) {
    if (b->p.lambda != 1.0) st = launch_lambda_rescale(s, n, rp, vv, b->p.lambda, b->p.drop_leaf);
    if (!st) st = gpu_symmetrize(s, n, nnz, rp, cc, vv, &srp, &scc, &svv, &m, true);
  }
  if (!st && hipStreamSynchronize(s) != hipSuccess) {  // the caller's P is read, and m known
    set_error("preparation of P failed");
    st = SGTSNE_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(b->mu);
  release_deferred(b);
  if (!st) st = admit(b, n, m, srp, scc, svv, y_in, max_iter, ng_cap, job);
  if (srp != rp) {
    (void)hipFreeAsync(rp, s);
    (void)hipFreeAsync(cc, s);
    (void)hipFreeAsync(vv, s);
  }
  // the admitted P is read by copies queued on the iteration stream
  if (st) {
    (void)hipFreeAsync(srp, s);
    (void)hipFreeAsync(scc, s);
    (void)hipFreeAsync(svv, s);
  } else {
    SGT_TRY(defer_free(b, srp, scc, svv));
  }
  return st;
}

int sgtsne_batch_add_points_async(sgtsne_batch *b, const float *X, int n, int dim, int k, double u,
                                  const float *y_in, int max_iter, int ng_cap, int64_t *ticket) {
  SGT_TRY(check_batch(b));
  SGT_CHECK(X && n > k && k >= 1 && dim >= 1 && u > 0, SGTSNE_ERR_ARG, "bad kNN arguments");
  SGT_CHECK(ticket != nullptr, SGTSNE_ERR_ARG, "ticket is NULL");
  FrontEnd *slot = nullptr;
  for (FrontEnd &c : b->fe)
    if (c.stage == 0) {
      slot = &c;
      break;
    }
  SGT_CHECK(slot != nullptr, SGTSNE_ERR_STATE, "%d front ends are in flight: poll one to its admission first",
            kFrontEnds);
  FrontEnd &f = *slot;
  fe_clear(f);
  f.n = n;
  f.dim = dim;
  f.k = k;
  f.u = u;
  f.y_in = y_in;
  f.max_iter = max_iter > 0 ? max_iter : b->p.max_iter;
  f.ng_cap = ng_cap;
  const hipStream_t s = b->prep;
  const size_t nx = (size_t)n * dim;
  int st = SGTSNE_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (st == SGTSNE_OK && e != hipSuccess) {
      set_error("front end: %s failed: %s", what, hipGetErrorString(e));
      st = e == hipErrorOutOfMemory ? SGTSNE_ERR_NOMEM : SGTSNE_ERR_HIP;
    }
  };
  // temporaries come from the stream-ordered pool (see sgtsne_batch_add)
  chk(hipMallocAsync((void **)&f.Xf, nx * sizeof(float), s), "hipMallocAsync");
  chk(hipMallocAsync((void **)&f.Xd, nx * sizeof(double), s), "hipMallocAsync");
  chk(hipMallocAsync((void **)&f.Dd, (size_t)n * (k + 1) * sizeof(double), s), "hipMallocAsync");
  chk(hipMallocAsync((void **)&f.Id, (size_t)n * (k + 1) * sizeof(int32_t), s), "hipMallocAsync");
  chk(hipMallocAsync((void **)&f.rp, ((size_t)n + 1) * sizeof(int32_t), s), "hipMallocAsync");
  chk(hipMallocAsync((void **)&f.cc, (size_t)n * k * sizeof(int32_t), s), "hipMallocAsync");
  chk(hipMallocAsync((void **)&f.vv, (size_t)n * k * sizeof(float), s), "hipMallocAsync");
  chk(hipMemcpyAsync(f.Xf, X, nx * sizeof(float), hipMemcpyHostToDevice, s), "copy of X");
  if (!st) st = launch_to_double(s, f.Xf, f.Xd, (int64_t)nx);
  if (!st) st = knn_begin(s, f.Xd, n, dim, k, f.Id, f.Dd, SGTSNE_KNN_AUTO, f.counts, &f.knn);
  chk(hipEventRecord(f.ev, s), "event record");
  if (st) {
    fe_release(b, f);
    return st;
  }
  f.stage = 1;
  f.ticket = ++b->fe_tickets;
  *ticket = f.ticket;
  return SGTSNE_OK;
}

int sgtsne_batch_add_poll(sgtsne_batch *b, int64_t ticket, int wait, int *job) {
  SGT_TRY(check_batch(b));
  SGT_CHECK(job != nullptr, SGTSNE_ERR_ARG, "job is NULL");
  *job = -1;
  FrontEnd *slot = nullptr;
  for (FrontEnd &c : b->fe)
    if (c.stage != 0 && c.ticket == ticket) slot = &c;
  SGT_CHECK(slot != nullptr, SGTSNE_ERR_ARG, "no front end %lld in flight", (long long)ticket);
  FrontEnd &f = *slot;
  for (;;) {
    const hipError_t e = wait ? hipEventSynchronize(f.ev) : hipEventQuery(f.ev);
    if (e == hipErrorNotReady) return SGTSNE_OK;
    if (e != hipSuccess) {
      set_error("front end failed: %s", hipGetErrorString(e));
      fe_release(b, f);
      return SGTSNE_ERR_HIP;
    }
    const int r = fe_advance(b, f, job);
    if (r == 1) return SGTSNE_OK;
    if (r != 0) return r;
    if (!wait) return SGTSNE_OK;  // the next part is queued; a later poll takes it on
  }
}

int sgtsne_batch_add_points(sgtsne_batch *b, const float *X, int n, int dim, int k, double u, const float *y_in,
                            int max_iter, int ng_cap, int *job) {
  int64_t ticket = 0;
  SGT_TRY(sgtsne_batch_add_points_async(b, X, n, dim, k, u, y_in, max_iter, ng_cap, &ticket));
  int slot = -1;
  const int st = sgtsne_batch_add_poll(b, ticket, 1, &slot);
  if (job) *job = slot;
  return st;
}

int sgtsne_batch_step(sgtsne_batch *b, int iters) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_CHECK(iters >= 0, SGTSNE_ERR_ARG, "iters < 0");
  release_deferred(b);
  if (b->arena.empty()) return SGTSNE_OK;
  if (!b->tune_pinned &&
      (b->tuned_points == 0 || b->used_points > 2 * b->tuned_points || 2 * b->used_points < b->tuned_points))
    SGT_TRY(tune(b));
  int64_t left = iters;
  while (left > 0) {
    const int64_t chunk = next_event(b, left);
    SGT_TRY(run_iterations(b, chunk));
    left -= chunk;
    SGT_TRY(do_events(b));
  }
  return SGTSNE_OK;
}

int sgtsne_batch_sync(sgtsne_batch *b) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_HIP(hipStreamSynchronize(b->stream));
  return SGTSNE_OK;
}

int sgtsne_batch_mark(sgtsne_batch *b, int64_t *mark) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_CHECK(mark != nullptr, SGTSNE_ERR_ARG, "mark is NULL");
  SGT_HIP(hipEventRecord(b->marks[b->nmarks % kMarks], b->stream));
  *mark = b->nmarks++;
  return SGTSNE_OK;
}

int sgtsne_batch_wait(sgtsne_batch *b, int64_t mark) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_CHECK(mark >= 0 && mark < b->nmarks, SGTSNE_ERR_ARG, "no mark %lld", (long long)mark);
  if (mark <= b->passed_upto) return SGTSNE_OK;
  // the slot's event may have been recorded again for a later mark: waiting
  // for that later point covers this one too
  const int64_t slot_mark = b->nmarks - 1 - ((b->nmarks - 1 - mark) % kMarks);
  SGT_HIP(hipEventSynchronize(b->marks[mark % kMarks]));
  b->passed_upto = std::max(b->passed_upto, slot_mark);
  return SGTSNE_OK;
}

int sgtsne_batch_passed(sgtsne_batch *b, int64_t mark, int *passed) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_CHECK(mark >= 0 && mark < b->nmarks && passed, SGTSNE_ERR_ARG, "no mark %lld", (long long)mark);
  if (mark <= b->passed_upto) {
    *passed = 1;
    return SGTSNE_OK;
  }
  // the slot holds this mark or the latest later one recorded into it
  const int64_t slot_mark = b->nmarks - 1 - ((b->nmarks - 1 - mark) % kMarks);
  const hipError_t e = hipEventQuery(b->marks[mark % kMarks]);
  if (e == hipErrorNotReady) {
    *passed = 0;
    return SGTSNE_OK;
  }
  SGT_HIP(e);
  b->passed_upto = std::max(b->passed_upto, slot_mark);
  *passed = 1;
  return SGTSNE_OK;
}

int sgtsne_batch_job_status(sgtsne_batch *b, int job, int *status, int *iter) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_TRY(check_job(b, job));
  int v[2] = {0, 0};
  const HostJob &h = b->hj[job];
  const hipStream_t cs = h.done_marked ? b->copy : b->stream;  // a finished job's record no longer changes
  if (h.done_marked) SGT_HIP(hipStreamWaitEvent(b->copy, h.done_ev, 0));
  SGT_HIP(hipMemcpyAsync(v, &b->jobs_d[job].iter, 2 * sizeof(int), hipMemcpyDeviceToHost, cs));
  SGT_HIP(hipStreamSynchronize(cs));
  if (iter) *iter = v[0];
  if (status) {
    const int s = v[1];
    *status = s == BJ_DONE ? SGTSNE_JOB_DONE
              : s == BJ_ERR_NUMERIC ? SGTSNE_ERR_NUMERIC
              : s == BJ_ERR_GRID ? SGTSNE_ERR_FULL
              : SGTSNE_JOB_RUNNING;
  }
  return SGTSNE_OK;
}

int sgtsne_batch_job_left(sgtsne_batch *b, int job, int64_t *left) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_TRY(check_job(b, job));
  SGT_CHECK(left != nullptr, SGTSNE_ERR_ARG, "left is NULL");
  const HostJob &h = b->hj[job];
  *left = std::max<int64_t>(0, (int64_t)h.max_iter + 1 - h.t);
  return SGTSNE_OK;
}

int sgtsne_batch_get_Y(sgtsne_batch *b, int job, float *y, int layout) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_TRY(check_job(b, job));
  SGT_CHECK(y != nullptr && (layout == SGTSNE_LAYOUT_SOA || layout == SGTSNE_LAYOUT_AOS), SGTSNE_ERR_ARG,
            "bad arguments");
  return fetch_points(b, job, b->Y, y, layout);
}

int sgtsne_batch_terms(sgtsne_batch *b, int job, float *Fattr, float *Frep, double *Z, int *ng, int layout) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_TRY(check_job(b, job));
  SGT_CHECK(layout == SGTSNE_LAYOUT_SOA || layout == SGTSNE_LAYOUT_AOS, SGTSNE_ERR_ARG, "bad layout");
  int st_iter[2] = {0, 0};
  SGT_HIP(hipMemcpyAsync(st_iter, &b->jobs_d[job].iter, 2 * sizeof(int), hipMemcpyDeviceToHost, b->stream));
  SGT_HIP(hipStreamSynchronize(b->stream));
  SGT_CHECK(st_iter[1] == BJ_RUN, SGTSNE_ERR_STATE, "job %d is not iterating (status %d)", job, st_iter[1]);
  SGT_TRY(enqueue_terms(b));
  SGT_TRY(launch_bzeta(b->stream, b->jobs_d, job, b->zpart, b->scal_d));
  double z = 0.0;
  GridParams g{};
  SGT_HIP(hipMemcpyAsync(&z, b->scal_d, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  SGT_HIP(hipMemcpyAsync(&g, &b->jobs_d[job].g, sizeof(GridParams), hipMemcpyDeviceToHost, b->stream));
  SGT_HIP(hipStreamSynchronize(b->stream));
  if (Z) *Z = z;
  if (ng) *ng = g.ng;
  if (Fattr) SGT_TRY(fetch_points(b, job, b->Fattr, Fattr, layout));
  if (Frep) SGT_TRY(fetch_points(b, job, b->Frep, Frep, layout, 1.0 / z));
  return SGTSNE_OK;
}

int sgtsne_batch_remove(sgtsne_batch *b, int job) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_TRY(check_job(b, job));
  // everything below is stream-ordered after the queued iterations
  HostJob &h = b->hj[job];
  BatchJob empty{};
  empty.status = BJ_EMPTY;
  SGT_TRY(upload(b, b->jobs_d + job, &empty, sizeof(BatchJob)));
  std::memcpy(&b->jobs_up[job], &empty, kJobStatic);
  release_deferred(b);
  SGT_TRY(defer_free(b, h.gmem));
  if (h.done_ev) (void)hipEventDestroy(h.done_ev);
  b->grid_used -= h.gbytes;
  b->arena.erase(std::find(b->arena.begin(), b->arena.end(), job));
  const bool tail = h.off + h.n == b->used_points;
  if (tail) {
    b->used_points = h.off;
    b->used_nnz = h.nnz_off;
  }
  h = HostJob{};
  b->jobs_h[job] = BatchJob{};
  // compact when holes hold more than a quarter of the arena
  int64_t live = 0;
  for (int j : b->arena) live += b->hj[j].n;
  if (4 * (b->used_points - live) > b->used_points) SGT_TRY(compact(b));
  else SGT_TRY(rebuild_items(b));
  return SGTSNE_OK;
}

int sgtsne_batch_info(sgtsne_batch *b, int *jobs, int64_t *points, int64_t *nnz, int64_t *grid_bytes,
                      int64_t *graph_replays) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  if (jobs) *jobs = (int)b->arena.size();
  if (points) *points = b->used_points;
  if (nnz) *nnz = b->used_nnz;
  if (grid_bytes) *grid_bytes = (int64_t)b->grid_used;
  if (graph_replays) *graph_replays = b->graph_replays;
  return SGTSNE_OK;
}

int sgtsne_batch_tuning(sgtsne_batch *b, int *group, int *unroll, int *rows_per_block, int *window, int *use_graph) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  if (window) *window = b->spmv_win;
  if (group) *group = b->spmv_group;
  if (unroll) *unroll = b->spmv_unroll;
  if (rows_per_block) *rows_per_block = b->spmv_rpb;
  if (use_graph) *use_graph = b->use_graph ? 1 : 0;
  return SGTSNE_OK;
}

int sgtsne_batch_set_attractive(sgtsne_batch *b, int unroll, int rows_per_block, int window) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_CHECK(unroll >= 0 && rows_per_block >= 0 && window >= 0 && (window == 0 || rows_per_block > 0),
            SGTSNE_ERR_ARG, "bad attractive kernel shape");
  SGT_CHECK((size_t)window * b->ds * sizeof(float) <= 65536, SGTSNE_ERR_ARG, "window does not fit LDS");
  const int keep[3] = {b->spmv_unroll, b->spmv_rpb, b->spmv_win};
  b->spmv_unroll = unroll;
  b->spmv_rpb = rows_per_block;
  b->spmv_win = window;
  // one launch on the current arena checks that the shape exists (it only
  // rewrites Fattr, which every iteration recomputes first)
  int st = attractive(b);
  if (st == SGTSNE_OK && hipStreamSynchronize(b->stream) != hipSuccess) {
    set_error("attractive kernel failed");
    st = SGTSNE_ERR_HIP;
  }
  if (st != SGTSNE_OK) {
    b->spmv_unroll = keep[0];
    b->spmv_rpb = keep[1];
    b->spmv_win = keep[2];
    return st;
  }
  b->tune_pinned = true;
  invalidate_graph(b);
  return SGTSNE_OK;
}

int sgtsne_batch_set_graph(sgtsne_batch *b, int on) {
  SGT_TRY(check_batch(b));
  std::lock_guard<std::mutex> lock(b->mu);
  SGT_HIP(hipStreamSynchronize(b->stream));
  b->use_graph = on != 0;
  release_graph(b);
  return SGTSNE_OK;
}

}  // extern "C"
