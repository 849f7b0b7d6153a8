// engine.hip -- per-context state, the gradient-descent loop and the C-ABI.
//
// One context = one embedding job on the current device with its own stream.
// The loop follows kl_minimization (CU/gradient_descend.cu:102-245;
// HY/gradient_descendHYB.cpp:160-244): eta = 200, momentum 0.5 -> 0.8 after
// iteration 250, alpha -> 1 after iteration early_iter, mean-centring every
// iteration (HYB semantics), h-driven grid size (HY/qq.cpp:159-166).
//
// Per iteration there is one device->host transfer of the (min, max) of Y,
// needed to size the grid; the attractive SpMV is issued before the host
// waits for it, so the GPU stays busy across the round trip.
#include "common.hpp"
#include "kernels.hpp"
#include "grid_device.hpp"
#include "../../include/sgtsne.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <atomic>
#include <random>
#include <thread>
#include <vector>

namespace sgt {

static thread_local char g_err[1024] = "no error";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

static const int kGridSizes[] = {SGT_GRID_SIZE_LIST};

int best_grid_size(int nGrid) {
  return best_grid_size_in(kGridSizes, (int)(sizeof(kGridSizes) / sizeof(kGridSizes[0])), nGrid);
}

// CUSP's compute_optimal_entries_per_row rule (the reference calls it at
// CU/matrix_converter.cpp:237): the ELL width is the smallest w at which the
// rows longer than w are fewer than max(4096, n / 3).
static int optimal_ell_width(int n, const std::vector<int32_t> &rowptr) {
  int maxlen = 0;
  for (int i = 0; i < n; i++) maxlen = std::max(maxlen, rowptr[i + 1] - rowptr[i]);
  std::vector<int64_t> hist((size_t)maxlen + 1, 0);
  for (int i = 0; i < n; i++) hist[rowptr[i + 1] - rowptr[i]]++;
  int64_t rows = n;
  for (int w = 0; w < maxlen; w++) {
    rows -= hist[w];  // rows with more than w entries
    if (3.0 * (double)rows < (double)n || rows < 4096) return w;
  }
  return maxlen;
}

struct PhaseEvents {
  hipEvent_t e[6];  // attr start, attr end, s2g end, conv end, g2s end, update end
  bool valid = false;
};

}  // namespace sgt

using namespace sgt;

struct sgtsne_ctx {
  sgtsne_params p{};
  hipStream_t stream = nullptr;
  int device = 0;
  int n = 0, d = 2, ds = 2;
  int64_t nnz = 0;
  int iter = 0;
  bool have_P = false, have_Y = false;
  int points_n = 0;

  // P (CSR always; COO rows / hybrid as requested)
  int32_t *rowptr = nullptr, *col = nullptr;
  float *val = nullptr;
  int32_t *coo_row = nullptr;
  int ell_w = 0;
  int32_t *ell_col = nullptr;
  float *ell_val = nullptr;
  int64_t tail_nnz = 0;
  int32_t *tail_row = nullptr, *tail_col = nullptr;
  float *tail_val = nullptr;
  int seg_group = 16;
  int seg_unroll = 0;  // 0: plain segmented kernel; else the unrolled form
  int seg_rpb = 0;     // > 0: blocked form with this many rows per work-group
  int seg_win = 0;     // > 0: windowed blocked form, this many points of Y staged in LDS;
                       // -1: column-localised blocks (loc_* below); <= -64: chunk-cached
                       // blocked form with chunks of -seg_win points (cc_* below)
  // column-localised blocks: row ranges, their distinct columns, 16-bit local indices
  int loc_nblk = 0, loc_cap = 0;
  int32_t *loc_brow = nullptr, *loc_bu = nullptr, *loc_ucol = nullptr;
  uint16_t *loc_lidx = nullptr;
  double loc_reuse = 0.0;  // nonzeros per staged column
  // chunk-cached blocked form: per range of cc_rpb rows, the cc_nslots chunks
  // of 2^cc_chs points of Y its neighbours fall in most, staged in LDS
  int cc_rpb = 0, cc_chs = 0, cc_nslots = 0;
  int32_t *cc_slots = nullptr;
  std::vector<double> tune_ms;  // autotuner timings (ms per candidate)
  std::vector<std::array<int, 4>> tune_cands;
  bool tune_pinned = false;  // the caller fixed the attractive kernel: no re-tuning

  // per-point state, AoS with stride ds
  float *Y = nullptr, *uY = nullptr, *gains = nullptr, *Fattr = nullptr, *Frep = nullptr;

  // reductions
  RedWork rw{};
  DevScalars *sc = nullptr;       // device
  DevScalars *sc_host = nullptr;  // pinned
  hipEvent_t ev_scal = nullptr;

  // grid workspace
  size_t grid_cap = 0, w_cap = 0, k_cap = 0;
  float *V = nullptr, *Phi = nullptr;
  float *s2g_part = nullptr;  // per-block private grids of the spread
  int64_t s2g_cap = 0;        // floats
  ConvBuffers cb{};

  // timing
  bool timing = true;
  PhaseEvents pe[2];
  double acc_ms[7] = {0, 0, 0, 0, 0, 0, 0};
  double *time_info = nullptr;  // optional per-iteration output (sgtsne_run)
  int time_info_base = 0;
  GridParams last_grid{};

  // compute_error history
  std::vector<double> errors;

  // locality order: device row r holds the caller's point perm[r]
  bool permuted = false;
  int reorders = 0;
  int reorders_declined = 0;  // Morton orders measured slower than the current one
  std::vector<int32_t> perm, iperm;
};

namespace {

int alloc_dev(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 1);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return SGTSNE_ERR_NOMEM;
  }
  return SGTSNE_OK;
}

template <class T>
void free_dev(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void free_loc(sgtsne_ctx *c) {
  free_dev(c->loc_brow);
  free_dev(c->loc_bu);
  free_dev(c->loc_ucol);
  free_dev(c->loc_lidx);
  c->loc_nblk = 0;
  if (c->seg_win == -1) c->seg_win = 0;
}

void free_cc(sgtsne_ctx *c) {
  free_dev(c->cc_slots);
  c->cc_rpb = c->cc_chs = c->cc_nslots = 0;
  if (c->seg_win <= -2) c->seg_win = 0;
}

// LDS of the chunk-cached form: 128 KB of cached points (plus a byte per chunk)
constexpr size_t kCcPointBytes = 131072;

int cc_slots_for(const sgtsne_ctx *c, int chs) {
  return (int)std::min<size_t>(127, (kCcPointBytes / ((size_t)c->ds * sizeof(float))) >> chs);
}

bool cc_fits(const sgtsne_ctx *c, int chs) {
  const int ns = cc_slots_for(c, chs);
  return ns >= 1 && ((size_t)ns << chs) * c->ds * sizeof(float) + (size_t)div_up(c->n, 1 << chs) <= 147456;
}

int log2_exact(int v) {
  int k = 0;
  while ((1 << k) < v) ++k;
  return (1 << k) == v ? k : -1;
}

// The chunk table of the chunk-cached form for ranges of rpb rows and chunks of
// 2^chs points, for the current order of P (kept while the parameters match).
int build_chunk_cache(sgtsne_ctx *c, int rpb, int chs) {
  if (c->cc_slots && c->cc_rpb == rpb && c->cc_chs == chs) return SGTSNE_OK;
  free_dev(c->cc_slots);
  c->cc_rpb = c->cc_chs = c->cc_nslots = 0;
  SGT_CHECK(rpb > 0 && chs >= 6 && chs <= 16 && cc_fits(c, chs), SGTSNE_ERR_ARG,
            "chunk cache with chunks of %d points does not fit LDS", 1 << chs);
  const int nranges = div_up(c->n, rpb), nchunks = div_up(c->n, 1 << chs), nslots = cc_slots_for(c, chs);
  int32_t *counts = nullptr;
  SGT_TRY(alloc_dev((void **)&counts, sizeof(int32_t) * (size_t)nranges * nchunks));
  int st = alloc_dev((void **)&c->cc_slots, sizeof(int32_t) * (size_t)nranges * nslots);
  if (st == SGTSNE_OK) st = launch_chunk_cache(c->stream, c->n, rpb, chs, nslots, c->rowptr, c->col, counts, c->cc_slots);
  if (st == SGTSNE_OK && hipStreamSynchronize(c->stream) != hipSuccess) {
    set_error("building the chunk table failed");
    st = SGTSNE_ERR_HIP;
  }
  free_dev(counts);
  if (st != SGTSNE_OK) {
    free_dev(c->cc_slots);
    return st;
  }
  c->cc_rpb = rpb;
  c->cc_chs = chs;
  c->cc_nslots = nslots;
  return SGTSNE_OK;
}

void free_P(sgtsne_ctx *c) {
  free_loc(c);
  free_cc(c);
  free_dev(c->rowptr);
  free_dev(c->col);
  free_dev(c->val);
  free_dev(c->coo_row);
  free_dev(c->ell_col);
  free_dev(c->ell_val);
  free_dev(c->tail_row);
  free_dev(c->tail_col);
  free_dev(c->tail_val);
  c->have_P = false;
  c->nnz = 0;
  c->ell_w = 0;
  c->tail_nnz = 0;
}

int check_params(const sgtsne_params *p) {
  SGT_CHECK(p != nullptr, SGTSNE_ERR_ARG, "params is NULL");
  SGT_CHECK(p->d >= 1 && p->d <= 3, SGTSNE_ERR_ARG, "d must be 1, 2 or 3 (got %d)", p->d);
  SGT_CHECK(p->format == SGTSNE_FORMAT_CSR || p->format == SGTSNE_FORMAT_COO ||
                p->format == SGTSNE_FORMAT_HYBRID || p->format == SGTSNE_FORMAT_CSR_SEG,
            SGTSNE_ERR_ARG, "format %d unsupported (BSR is a stub in the reference)", p->format);
  SGT_CHECK(p->bs == 0, SGTSNE_ERR_ARG, "block formats are not supported (bs must be 0)");
  SGT_CHECK(p->max_iter >= 0 && p->early_iter >= -1, SGTSNE_ERR_ARG, "bad iteration counts");
  SGT_CHECK(p->lambda > 0, SGTSNE_ERR_ARG, "lambda must be positive");
  SGT_CHECK(p->grid_mode == SGTSNE_GRID_H || p->grid_mode == SGTSNE_GRID_FIXED, SGTSNE_ERR_ARG,
            "bad grid_mode");
  SGT_CHECK(p->layout == SGTSNE_LAYOUT_SOA || p->layout == SGTSNE_LAYOUT_AOS, SGTSNE_ERR_ARG,
            "bad layout");
  SGT_CHECK(p->zeta_mode == SGTSNE_ZETA_DIRECT || p->zeta_mode == SGTSNE_ZETA_ALGEBRAIC, SGTSNE_ERR_ARG,
            "bad zeta_mode");
  SGT_CHECK(p->spread >= SGTSNE_SPREAD_AUTO && p->spread <= SGTSNE_SPREAD_MFMA, SGTSNE_ERR_ARG, "bad spread");
  return SGTSNE_OK;
}

double default_h(int d) { return d == 1 ? 0.5 : d == 2 ? 0.7 : 1.2; }

// Host side of one iteration's grid mapping (HY/qq.cpp:159-166, HY/nuconv.cpp:28-47).
int make_grid(sgtsne_ctx *c, const DevScalars &s, GridParams &g) {
  const int d = c->d;
  double maxy = 0.0;
  for (int a = 0; a < d; a++) {
    SGT_CHECK(std::isfinite(s.minv[a]) && std::isfinite(s.maxv[a]), SGTSNE_ERR_NUMERIC,
              "embedding is not finite (iteration %d)", c->iter);
    maxy = std::max(maxy, s.maxv[a] - s.minv[a]);
  }
  SGT_CHECK(maxy > 0.0, SGTSNE_ERR_NUMERIC, "embedding collapsed to a point (iteration %d)", c->iter);
  int nGrid;
  if (c->p.grid_mode == SGTSNE_GRID_FIXED) {
    nGrid = best_grid_size(c->p.ng);
  } else {
    const double h = c->p.h > 0 ? c->p.h : default_h(d);
    nGrid = std::max((int)std::ceil(maxy / h), 14);
    nGrid = best_grid_size(nGrid);
  }
  grid_from_box(d, nGrid, s.minv, s.maxv, maxy, g);
  return SGTSNE_OK;
}

int ensure_grid(sgtsne_ctx *c, const GridParams &g) {
  const int m = c->d + 2;  // charges 0..d, plus the direct-Z channel of Phi
  const size_t G = c->d == 1 ? (size_t)g.ng : c->d == 2 ? (size_t)g.ng * g.ng : (size_t)g.ng * g.ng * g.ng;
  const size_t gb = G * m * sizeof(float);
  size_t wb = 0, kb = 0;
  conv_workspace_bytes(c->d, g.ng, g.L, &wb, &kb);
  const int64_t need = s2g_scratch_need(g, c->n, c->p.spread);
  if (gb > c->grid_cap || wb > c->w_cap || kb > c->k_cap || need > c->s2g_cap) {
    SGT_HIP(hipStreamSynchronize(c->stream));
  }
  if (need > c->s2g_cap) {
    free_dev(c->s2g_part);
    SGT_TRY(alloc_dev((void **)&c->s2g_part, (size_t)need * sizeof(float)));
    c->s2g_cap = need;
  }
  if (gb > c->grid_cap) {
    free_dev(c->V);
    free_dev(c->Phi);
    SGT_TRY(alloc_dev((void **)&c->V, gb));
    SGT_TRY(alloc_dev((void **)&c->Phi, gb));
    c->grid_cap = gb;
  }
  if (wb > c->w_cap) {
    free_dev(c->cb.W);
    SGT_TRY(alloc_dev((void **)&c->cb.W, wb));
    c->w_cap = wb;
  }
  if (kb > c->k_cap) {
    free_dev(c->cb.Kt1);
    free_dev(c->cb.Kt2);
    free_dev(c->cb.Kh);
    SGT_TRY(alloc_dev((void **)&c->cb.Kt1, kb));
    SGT_TRY(alloc_dev((void **)&c->cb.Kt2, kb));
    SGT_TRY(alloc_dev((void **)&c->cb.Kh, kb));
    c->k_cap = kb;
  }
  return SGTSNE_OK;
}

int attractive(sgtsne_ctx *c) {
  switch (c->p.format) {
    case SGTSNE_FORMAT_CSR_SEG:
      if (c->seg_win <= -2)
        return launch_attr_csr_cc(c->stream, c->d, c->seg_group, c->seg_unroll, c->cc_rpb, c->cc_chs, c->cc_nslots,
                                  c->cc_slots, c->n, c->rowptr, c->col, c->val, c->Y, c->Fattr);
      if (c->seg_win < 0)
        return launch_attr_csr_loc(c->stream, c->d, c->seg_group, c->loc_nblk, c->loc_cap, c->loc_brow, c->loc_bu,
                                   c->loc_ucol, c->rowptr, c->loc_lidx, c->val, c->Y, c->Fattr);
      if (c->seg_win > 0)
        return launch_attr_csr_win(c->stream, c->d, c->seg_group, c->seg_unroll, c->seg_rpb, c->seg_win, c->n,
                                   c->rowptr, c->col, c->val, c->Y, c->Fattr);
      if (c->seg_rpb > 0)
        return launch_attr_csr_blk(c->stream, c->d, c->seg_group, c->seg_unroll, c->seg_rpb, c->n, c->rowptr,
                                   c->col, c->val, c->Y, c->Fattr);
      if (c->seg_unroll > 0)
        return launch_attr_csr_seg2(c->stream, c->d, c->seg_group, c->seg_unroll, c->n, c->rowptr, c->col,
                                    c->val, c->Y, c->Fattr);
      return launch_attr_csr_seg(c->stream, c->d, c->seg_group, c->n, c->rowptr, c->col, c->val, c->Y, c->Fattr);
    case SGTSNE_FORMAT_CSR:
      return launch_attr_csr_row(c->stream, c->d, c->n, c->rowptr, c->col, c->val, c->Y, c->Fattr);
    case SGTSNE_FORMAT_COO:
      SGT_TRY(launch_fill(c->stream, c->Fattr, (int64_t)c->n * c->ds, 0.0f));
      return launch_attr_coo(c->stream, c->d, c->nnz, c->coo_row, c->col, c->val, c->Y, c->Fattr);
    case SGTSNE_FORMAT_HYBRID:
      SGT_TRY(launch_attr_ell(c->stream, c->d, c->n, c->ell_w, c->ell_col, c->ell_val, c->Y, c->Fattr));
      return launch_attr_coo(c->stream, c->d, c->tail_nnz, c->tail_row, c->tail_col, c->tail_val, c->Y, c->Fattr);
    default:
      SGT_CHECK(false, SGTSNE_ERR_ARG, "bad format");
  }
}

// Repulsive term for the current (centred) Y: S2G -> convolution -> G2S.
int repulsive(sgtsne_ctx *c, const GridParams &g, PhaseEvents *pe) {
  const int zdirect = c->p.zeta_mode == SGTSNE_ZETA_DIRECT;
  SGT_TRY(ensure_grid(c, g));
  int64_t need = 0;
  SGT_TRY(launch_s2g(c->stream, g, c->n, c->Y, c->V, c->s2g_part, c->s2g_cap, &need, c->p.spread));
  if (pe) SGT_HIP(hipEventRecord(pe->e[2], c->stream));
  SGT_TRY(launch_conv(c->stream, g, zdirect, c->V, c->Phi, c->cb));
  if (pe) SGT_HIP(hipEventRecord(pe->e[3], c->stream));
  SGT_TRY(launch_g2s_force(c->stream, g, c->n, c->Y, c->Phi, c->Frep, zdirect, c->rw, c->sc));
  if (pe) SGT_HIP(hipEventRecord(pe->e[4], c->stream));
  c->last_grid = g;
  return SGTSNE_OK;
}

int center_and_fetch(sgtsne_ctx *c, int apply_mean, PhaseEvents *pe = nullptr) {
  SGT_TRY(launch_center_minmax(c->stream, c->d, c->n, c->Y, apply_mean, c->rw, c->sc));
  if (pe) {
    SGT_HIP(hipEventRecord(pe->e[5], c->stream));
    pe->valid = true;
  }
  SGT_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  SGT_HIP(hipEventRecord(c->ev_scal, c->stream));
  return SGTSNE_OK;
}

void harvest_timing(sgtsne_ctx *c, PhaseEvents &pe, int it) {
  if (!pe.valid) return;
  float t[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < 5; k++) (void)hipEventElapsedTime(&t[k], pe.e[k], pe.e[k + 1]);
  // t: attr, s2g(+fill), conv, g2s, update(+centre)
  c->acc_ms[0] += t[0];
  c->acc_ms[1] += t[1];
  c->acc_ms[2] += t[2];
  c->acc_ms[3] += t[3];
  c->acc_ms[4] += t[4];
  c->acc_ms[5] += t[0] + t[1] + t[2] + t[3] + t[4];
  c->acc_ms[6] += 1;
  if (c->time_info) {
    double *ti = c->time_info + 7 * (size_t)(it - c->time_info_base);
    ti[0] = t[0];
    ti[1] = t[1];
    ti[2] = t[2];
    ti[3] = t[3];
    ti[4] = 0.0;  // forces are fused into G2S
    ti[5] = t[1] + t[2] + t[3];
    ti[6] = 0.0;
  }
  pe.valid = false;
}

// Iterations at which a re-order of the points along a Morton curve of the
// embedding is tried (reorder >= 2, CSR formats).  Once the clusters have
// formed, the neighbours of a point in P are its neighbours in Y, and the
// gathers of the attractive term touch small windows of memory; before that
// (large N converges slowly) the breadth-first order of P is better.  So the
// new order is kept only when the attractive kernel measures faster on it.
constexpr int kSpatialReorderIters[] = {50, 150, 300, 600};

int tune_attractive(sgtsne_ctx *c);
int build_localised(sgtsne_ctx *c);

// Mean ms of the current attractive kernel over 3 launches (after a warm one)
// on the given P and Y; the context's own arrays are restored afterwards.
int time_attractive(sgtsne_ctx *c, int32_t *rp, int32_t *col, float *val, float *Y, float *ms) {
  int32_t *k_rp = c->rowptr, *k_col = c->col;
  float *k_val = c->val, *k_Y = c->Y;
  // the column-localised blocks and the chunk table belong to one order:
  // compare orders with the grid-stride kernel of the same lanes per row
  // (and, for the chunk-cached form, the same steps)
  const int k_win = c->seg_win, k_unroll = c->seg_unroll, k_rpb = c->seg_rpb;
  if (c->seg_win < 0) {
    if (c->seg_win == -1) c->seg_unroll = 0;
    c->seg_win = 0;
    c->seg_rpb = 0;
  }
  c->rowptr = rp;
  c->col = col;
  c->val = val;
  c->Y = Y;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess ? SGTSNE_OK : SGTSNE_ERR_HIP;
  if (!st) st = attractive(c);
  if (!st) (void)hipEventRecord(e0, c->stream);
  for (int r = 0; r < 3 && !st; r++) st = attractive(c);
  if (!st) (void)hipEventRecord(e1, c->stream);
  if (!st && hipEventSynchronize(e1) != hipSuccess) st = SGTSNE_ERR_HIP;
  float t = 0.f;
  if (!st) (void)hipEventElapsedTime(&t, e0, e1);
  *ms = t / 3.0f;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  c->rowptr = k_rp;
  c->col = k_col;
  c->val = k_val;
  c->Y = k_Y;
  c->seg_win = k_win;
  c->seg_unroll = k_unroll;
  c->seg_rpb = k_rpb;
  if (st == SGTSNE_ERR_HIP) set_error("timing of the attractive kernel failed");
  return st;
}

int spatial_reorder(sgtsne_ctx *c) {
  const int n = c->n, d = c->d, ds = c->ds;
  SGT_HIP(hipEventSynchronize(c->ev_scal));
  const DevScalars sh = *c->sc_host;
  double lo[3] = {0, 0, 0}, extent = 0.0;
  for (int a = 0; a < d; a++) {
    lo[a] = sh.minv[a];
    extent = std::max(extent, sh.maxv[a] - sh.minv[a]);
  }
  if (!(extent > 0.0) || !std::isfinite(extent)) return SGTSNE_OK;  // the step reports it
  int32_t *order = nullptr, *inv = nullptr, *rp2 = nullptr, *col2 = nullptr;
  float *val2 = nullptr, *tmp = nullptr;
  SGT_TRY(alloc_dev((void **)&order, sizeof(int32_t) * (size_t)n));
  SGT_TRY(alloc_dev((void **)&inv, sizeof(int32_t) * (size_t)n));
  SGT_TRY(alloc_dev((void **)&rp2, sizeof(int32_t) * ((size_t)n + 1)));
  SGT_TRY(alloc_dev((void **)&col2, sizeof(int32_t) * (size_t)std::max<int64_t>(c->nnz, 1)));
  SGT_TRY(alloc_dev((void **)&val2, sizeof(float) * (size_t)std::max<int64_t>(c->nnz, 1)));
  SGT_TRY(alloc_dev((void **)&tmp, sizeof(float) * (size_t)n * ds));
  int st = spatial_order(c->stream, n, d, ds, c->Y, lo, extent, order, inv);
  if (!st) st = launch_permute_csr(c->stream, n, c->nnz, order, inv, c->rowptr, c->col, c->val, rp2, col2, val2);
  if (!st) st = launch_gather_points(c->stream, n, ds, order, c->Y, tmp);
  // keep the new order only if the attractive term runs faster on it
  float t_old = 0.f, t_new = 0.f;
  if (!st) st = time_attractive(c, c->rowptr, c->col, c->val, c->Y, &t_old);
  if (!st) st = time_attractive(c, rp2, col2, val2, tmp, &t_new);
  if (!st && !(t_new < 0.98f * t_old)) {
    (void)hipStreamSynchronize(c->stream);
    free_dev(order);
    free_dev(inv);
    free_dev(tmp);
    free_dev(rp2);
    free_dev(col2);
    free_dev(val2);
    c->reorders_declined++;
    return SGTSNE_OK;
  }
  if (!st) std::swap(c->Y, tmp);
  float **state[2] = {&c->uY, &c->gains};
  for (int k = 0; k < 2 && !st; k++) {
    st = launch_gather_points(c->stream, n, ds, order, *state[k], tmp);
    std::swap(*state[k], tmp);
  }
  std::vector<int32_t> ord(n);
  if (!st && hipMemcpyAsync(ord.data(), order, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
    set_error("copy of the order failed");
    st = SGTSNE_ERR_HIP;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && !st) {
    set_error("re-ordering failed");
    st = SGTSNE_ERR_HIP;
  }
  free_dev(order);
  free_dev(inv);
  free_dev(tmp);
  if (st) {
    free_dev(rp2);
    free_dev(col2);
    free_dev(val2);
    return st;
  }
  free_dev(c->rowptr);
  free_dev(c->col);
  free_dev(c->val);
  c->rowptr = rp2;
  c->col = col2;
  c->val = val2;
  // compose with the caller's order
  if (c->perm.empty()) {
    c->perm.resize(n);
    for (int r = 0; r < n; r++) c->perm[r] = r;
  }
  std::vector<int32_t> np(n);
  for (int r = 0; r < n; r++) np[r] = c->perm[ord[r]];
  c->perm.swap(np);
  c->iperm.assign(n, 0);
  for (int r = 0; r < n; r++) c->iperm[c->perm[r]] = r;
  c->permuted = true;
  c->reorders++;
  // the column-localised blocks follow the order
  const bool had_loc = c->loc_nblk > 0, using_loc = c->seg_win == -1;
  if (had_loc) {
    SGT_TRY(build_localised(c));
    if (using_loc && c->loc_nblk > 0) c->seg_win = -1;
  }
  // so does the chunk table
  if (c->cc_slots) {
    const int rpb = c->cc_rpb, chs = c->cc_chs;
    free_dev(c->cc_slots);
    SGT_TRY(build_chunk_cache(c, rpb, chs));
  }
  // the first re-order changes the gather pattern most: choose the kernel again
  if (c->reorders == 1 && c->p.format == SGTSNE_FORMAT_CSR_SEG && !c->tune_pinned) SGT_TRY(tune_attractive(c));
  return SGTSNE_OK;
}

int iterate(sgtsne_ctx *c) {
  const int it = c->iter;
  if (c->p.reorder >= 2 && (c->p.format == SGTSNE_FORMAT_CSR_SEG || c->p.format == SGTSNE_FORMAT_CSR))
    for (int r : kSpatialReorderIters)
      if (it == r) SGT_TRY(spatial_reorder(c));
  const double alpha = it <= c->p.early_iter ? c->p.alpha : 1.0;
  const double momentum = it <= 250 ? 0.5 : 0.8;
  const double eta = 200.0;
  PhaseEvents *pe = c->timing ? &c->pe[it & 1] : nullptr;
  if (pe) SGT_HIP(hipEventRecord(pe->e[0], c->stream));
  SGT_TRY(attractive(c));
  if (pe) SGT_HIP(hipEventRecord(pe->e[1], c->stream));
  // wait for the (min, max) of this iteration's Y (issued by the previous centring)
  SGT_HIP(hipEventSynchronize(c->ev_scal));
  if (c->timing) harvest_timing(c, c->pe[(it + 1) & 1], it - 1);
  const DevScalars s = *c->sc_host;
  GridParams g{};
  SGT_TRY(make_grid(c, s, g));
  SGT_TRY(repulsive(c, g, pe));
  if (c->p.compute_error > 0 && it % 10 == 0) {
    // relative L2 error of the interpolated repulsion (CU/gradient_descend.cu:206-215)
    SGT_HIP(hipStreamSynchronize(c->stream));
    const size_t nb = (size_t)c->n * c->ds;
    std::vector<float> fi(nb), fe(nb);
    DevScalars sh{};
    SGT_HIP(hipMemcpy(fi.data(), c->Frep, nb * sizeof(float), hipMemcpyDeviceToHost));
    SGT_HIP(hipMemcpy(&sh, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
    float *tmp = nullptr;
    SGT_TRY(alloc_dev((void **)&tmp, nb * sizeof(float)));
    int st = launch_rep_exact(c->stream, c->d, c->n, c->Y, tmp, c->rw, c->sc);
    if (st == SGTSNE_OK) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipMemcpy(fe.data(), tmp, nb * sizeof(float), hipMemcpyDeviceToHost);
      DevScalars se{};
      (void)hipMemcpy(&se, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost);
      double num = 0, den = 0;
      for (int i = 0; i < c->n; i++)
        for (int a = 0; a < c->d; a++) {
          double x = fi[(size_t)i * c->ds + a] / sh.Z, y = fe[(size_t)i * c->ds + a] / se.zexact;
          num += (x - y) * (x - y);
          den += y * y;
        }
      c->errors.push_back(den > 0 ? std::sqrt(num / den) : 0.0);
    }
    free_dev(tmp);
    SGT_TRY(st);
  }
  SGT_TRY(launch_update(c->stream, c->d, c->n, c->Y, c->uY, c->gains, c->Fattr, c->Frep,
                        (float)alpha, (float)momentum, (float)eta, c->rw, c->sc));
  SGT_TRY(center_and_fetch(c, 1, pe));
  c->iter++;
  return SGTSNE_OK;
}

// caller order <-> device order (perm empty: identity)
void pack_y(const float *y, int layout, int n, int d, int ds, const std::vector<int32_t> &perm,
            std::vector<float> &out) {
  out.assign((size_t)n * ds, 0.0f);
  for (int r = 0; r < n; r++) {
    const int i = perm.empty() ? r : perm[r];
    for (int a = 0; a < d; a++)
      out[(size_t)r * ds + a] = layout == SGTSNE_LAYOUT_AOS ? y[(size_t)i * d + a] : y[(size_t)a * n + i];
  }
}

void unpack_y(const std::vector<float> &in, int layout, int n, int d, int ds, const std::vector<int32_t> &perm,
              float *y, double scale = 1.0) {
  for (int r = 0; r < n; r++) {
    const int i = perm.empty() ? r : perm[r];
    for (int a = 0; a < d; a++) {
      float v = (float)(in[(size_t)r * ds + a] * scale);
      if (layout == SGTSNE_LAYOUT_AOS) y[(size_t)i * d + a] = v;
      else y[(size_t)a * n + i] = v;
    }
  }
}

// Breadth-first order of the graph of P (Cuthill-McKee without the degree
// sort): each component from its lowest-degree vertex, neighbours in column
// order.  Rows that are neighbours in P end up close in memory, so the
// gathers of the attractive term stay in cache.
// Reverse Cuthill-McKee order of the graph of P: per connected component, a
// breadth-first sweep from a pseudo-peripheral vertex (George-Liu: repeat
// the sweep from a minimum-degree vertex of the last level while the depth
// grows) that takes the unvisited neighbours of each vertex in increasing
// degree, the whole order then reversed.  Compared with a plain breadth-first
// order it keeps the neighbours of consecutive rows in a much narrower band
// of indices, so the gathers of consecutive rows share cache lines.
namespace {

// Breadth-first levels from `root` over unvisited vertices (mark != stamp),
// level by level on `nthreads` host threads: a vertex joins the next level
// on the thread whose atomic exchange of its mark succeeds.  Returns the
// depth and leaves the last level (as a set: its order depends on timing) in
// `last`; both are what the sequential sweep gives.
int bfs_levels(int root, const int32_t *rp, const int32_t *col, std::vector<int32_t> &mark, int stamp,
               std::vector<int32_t> &last, int nthreads) {
  std::vector<int32_t> front{root}, next;
  mark[root] = stamp;
  int depth = 0;
  std::vector<std::vector<int32_t>> part(nthreads);
  for (;;) {
    const size_t nf = front.size();
    // small levels: one thread (spawning costs more than they take)
    const int T = nf < 4096 ? 1 : nthreads;
    auto work = [&](int t) {
      std::vector<int32_t> &out = part[t];
      out.clear();
      const size_t b0 = nf * t / T, b1 = nf * (t + 1) / T;
      for (size_t h = b0; h < b1; ++h) {
        const int v = front[h];
        for (int e = rp[v]; e < rp[v + 1]; e++) {
          const int u = col[e];
          if (__atomic_load_n(&mark[u], __ATOMIC_RELAXED) != stamp &&
              __atomic_exchange_n(&mark[u], stamp, __ATOMIC_RELAXED) != stamp)
            out.push_back(u);
        }
      }
    };
    if (T == 1) {
      work(0);
    } else {
      std::vector<std::thread> pool;
      for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
      work(0);
      for (auto &th : pool) th.join();
    }
    next.clear();
    for (int t = 0; t < T; ++t) next.insert(next.end(), part[t].begin(), part[t].end());
    if (next.empty()) {
      last.swap(front);
      return depth;
    }
    ++depth;
    front.swap(next);
  }
}

}  // namespace

void bfs_order(int n, const int32_t *rp, const int32_t *col, std::vector<int32_t> &perm) {
  perm.clear();
  perm.reserve(n);
  std::vector<int32_t> deg(n), mark(n, 0), last;
  const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<uint8_t> placed(n, 0);
  for (int i = 0; i < n; i++) deg[i] = rp[i + 1] - rp[i];
  // candidate starts in increasing degree (ties by index)
  std::vector<int32_t> by_deg(n);
  for (int i = 0; i < n; i++) by_deg[i] = i;
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  int stamp = 1;
  std::vector<std::pair<int32_t, int32_t>> nb;
  for (int cand : by_deg) {
    if (placed[cand]) continue;
    // pseudo-peripheral vertex of cand's component
    int root = cand;
    int depth = bfs_levels(root, rp, col, mark, ++stamp, last, nthreads);
    for (int sweep = 0; sweep < 4; sweep++) {
      int best = last[0];
      for (const int v : last)
        if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
      const int d2 = bfs_levels(best, rp, col, mark, ++stamp, last, nthreads);
      if (d2 <= depth) break;
      depth = d2;
      root = best;
    }
    // Cuthill-McKee sweep of the component from root
    const size_t first = perm.size();
    perm.push_back(root);
    placed[root] = 1;
    for (size_t head = first; head < perm.size(); ++head) {
      const int v = perm[head];
      nb.clear();
      for (int t = rp[v]; t < rp[v + 1]; t++) {
        const int u = col[t];
        if (!placed[u]) {
          placed[u] = 1;
          nb.emplace_back(deg[u], u);
        }
      }
      std::sort(nb.begin(), nb.end());
      for (const auto &e : nb) perm.push_back(e.second);
    }
  }
  std::reverse(perm.begin(), perm.end());
}

int apply_locality_order(sgtsne_ctx *c) {
  const int n = c->n;
  std::vector<int32_t> rp((size_t)n + 1), col((size_t)std::max<int64_t>(c->nnz, 1));
  // P was written on the context's (non-blocking) stream, which a plain
  // hipMemcpy on the null stream does not wait for: read it back in order
  SGT_HIP(hipMemcpyAsync(rp.data(), c->rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, c->stream));
  if (c->nnz > 0)
    SGT_HIP(hipMemcpyAsync(col.data(), c->col, sizeof(int32_t) * (size_t)c->nnz, hipMemcpyDeviceToHost, c->stream));
  SGT_HIP(hipStreamSynchronize(c->stream));
  // the sweep indexes host arrays with these: check them first
  SGT_CHECK(rp[0] == 0 && rp[n] == c->nnz, SGTSNE_ERR_STATE, "device P inconsistent (rowptr ends %d, %d)", rp[0],
            rp[n]);
  for (int i = 0; i < n; i++)
    SGT_CHECK(rp[i] <= rp[i + 1], SGTSNE_ERR_STATE, "device P inconsistent: rowptr decreases at row %d", i);
  for (int64_t t = 0; t < c->nnz; t++)
    SGT_CHECK(col[t] >= 0 && col[t] < n, SGTSNE_ERR_STATE, "device P inconsistent: column %d at %lld", col[t],
              (long long)t);
  bfs_order(n, rp.data(), col.data(), c->perm);
  c->iperm.assign(n, 0);
  for (int r = 0; r < n; r++) c->iperm[c->perm[r]] = r;
  int32_t *dperm = nullptr, *diperm = nullptr, *rp2 = nullptr, *col2 = nullptr;
  float *val2 = nullptr;
  SGT_TRY(alloc_dev((void **)&dperm, sizeof(int32_t) * (size_t)n));
  SGT_TRY(alloc_dev((void **)&diperm, sizeof(int32_t) * (size_t)n));
  SGT_TRY(alloc_dev((void **)&rp2, sizeof(int32_t) * ((size_t)n + 1)));
  SGT_TRY(alloc_dev((void **)&col2, sizeof(int32_t) * (size_t)std::max<int64_t>(c->nnz, 1)));
  SGT_TRY(alloc_dev((void **)&val2, sizeof(float) * (size_t)std::max<int64_t>(c->nnz, 1)));
  SGT_HIP(hipMemcpyAsync(dperm, c->perm.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  SGT_HIP(hipMemcpyAsync(diperm, c->iperm.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  int st = launch_permute_csr(c->stream, n, c->nnz, dperm, diperm, c->rowptr, c->col, c->val, rp2, col2, val2);
  (void)hipStreamSynchronize(c->stream);
  free_dev(dperm);
  free_dev(diperm);
  free_dev(c->rowptr);
  free_dev(c->col);
  free_dev(c->val);
  c->rowptr = rp2;
  c->col = col2;
  c->val = val2;
  c->permuted = true;
  return st;
}

// Per-point state, sized by n (kept across P uploads of the same n).
int alloc_points(sgtsne_ctx *c) {
  const size_t nb = (size_t)c->n * c->ds;
  if (c->Y && c->points_n == c->n) return SGTSNE_OK;
  free_dev(c->Y);
  free_dev(c->uY);
  free_dev(c->gains);
  free_dev(c->Fattr);
  free_dev(c->Frep);
  SGT_TRY(alloc_dev((void **)&c->Y, nb * sizeof(float)));
  SGT_TRY(alloc_dev((void **)&c->uY, nb * sizeof(float)));
  SGT_TRY(alloc_dev((void **)&c->gains, nb * sizeof(float)));
  SGT_TRY(alloc_dev((void **)&c->Fattr, nb * sizeof(float)));
  SGT_TRY(alloc_dev((void **)&c->Frep, nb * sizeof(float)));
  SGT_HIP(hipMemsetAsync(c->Y, 0, nb * sizeof(float), c->stream));
  SGT_HIP(hipMemsetAsync(c->Fattr, 0, nb * sizeof(float), c->stream));
  SGT_HIP(hipMemsetAsync(c->Frep, 0, nb * sizeof(float), c->stream));
  c->points_n = c->n;
  return SGTSNE_OK;
}

// Autotuner of the attractive kernel: times every (lanes per row, unroll)
// candidate on the uploaded matrix and keeps the fastest.  Runs once per
// upload; the result is reported by sgtsne_ctx_tuning.
// Column-localised blocks of the current P (host build, once per order):
// consecutive rows are grouped while their distinct columns stay within
// `cap` (the LDS copy of Y) and the rows within 1024; each block's distinct
// columns are sorted, so the LDS staging loads of dense column runs coalesce.
int build_localised(sgtsne_ctx *c) {
  free_loc(c);
  const int n = c->n;
  const int cap = c->d == 3 ? 4096 : 8192;
  const int rmax = 1024;
  std::vector<int32_t> rp((size_t)n + 1), col((size_t)std::max<int64_t>(c->nnz, 1));
  SGT_HIP(hipStreamSynchronize(c->stream));
  SGT_HIP(hipMemcpy(rp.data(), c->rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost));
  if (c->nnz > 0) SGT_HIP(hipMemcpy(col.data(), c->col, sizeof(int32_t) * (size_t)c->nnz, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++)
    if (rp[i + 1] - rp[i] > cap) return SGTSNE_OK;  // a row wider than the LDS copy: form unavailable
  std::vector<int32_t> brow, bu, ucol, mark(n, -1), pos(n, 0);
  std::vector<uint16_t> lidx((size_t)std::max<int64_t>(c->nnz, 1));
  brow.push_back(0);
  bu.push_back(0);
  int r = 0, blk = 0;
  std::vector<int32_t> cur;
  while (r < n) {
    cur.clear();
    const int rstart = r;
    while (r < n && r - rstart < rmax) {
      int fresh = 0;
      for (int t = rp[r]; t < rp[r + 1]; t++) fresh += mark[col[t]] != blk;
      if ((int)cur.size() + fresh > cap && r > rstart) break;
      for (int t = rp[r]; t < rp[r + 1]; t++)
        if (mark[col[t]] != blk) {
          mark[col[t]] = blk;
          cur.push_back(col[t]);
        }
      ++r;
    }
    std::sort(cur.begin(), cur.end());
    for (size_t k = 0; k < cur.size(); k++) pos[cur[k]] = (int32_t)k;
    for (int t = rp[rstart]; t < rp[r]; t++) lidx[t] = (uint16_t)pos[col[t]];
    ucol.insert(ucol.end(), cur.begin(), cur.end());
    brow.push_back(r);
    bu.push_back((int32_t)ucol.size());
    ++blk;
  }
  c->loc_nblk = blk;
  c->loc_cap = cap;
  SGT_TRY(alloc_dev((void **)&c->loc_brow, sizeof(int32_t) * brow.size()));
  SGT_TRY(alloc_dev((void **)&c->loc_bu, sizeof(int32_t) * bu.size()));
  SGT_TRY(alloc_dev((void **)&c->loc_ucol, sizeof(int32_t) * std::max<size_t>(ucol.size(), 1)));
  SGT_TRY(alloc_dev((void **)&c->loc_lidx, sizeof(uint16_t) * lidx.size()));
  SGT_HIP(hipMemcpy(c->loc_brow, brow.data(), sizeof(int32_t) * brow.size(), hipMemcpyHostToDevice));
  SGT_HIP(hipMemcpy(c->loc_bu, bu.data(), sizeof(int32_t) * bu.size(), hipMemcpyHostToDevice));
  if (!ucol.empty()) SGT_HIP(hipMemcpy(c->loc_ucol, ucol.data(), sizeof(int32_t) * ucol.size(), hipMemcpyHostToDevice));
  SGT_HIP(hipMemcpy(c->loc_lidx, lidx.data(), sizeof(uint16_t) * lidx.size(), hipMemcpyHostToDevice));
  c->loc_reuse = ucol.empty() ? 0.0 : (double)c->nnz / (double)ucol.size();
  return SGTSNE_OK;
}

int tune_attractive(sgtsne_ctx *c) {
  // lanes per row, unroll (0 = plain), rows per work-group (0 = grid-stride),
  // points of Y staged in LDS (0 = none)
  struct Cand { int g, u, rpb, win; };
  const Cand cands[] = {{16, 0, 0, 0},     {32, 0, 0, 0},      {16, 4, 0, 0},      {32, 2, 0, 0},
                        {64, 2, 0, 0},     {16, 1, 64, 0},     {32, 1, 64, 0},     {32, 2, 64, 0},
                        {16, 2, 128, 0},   {32, 2, 256, 0},    {64, 1, 256, 0},    {16, 1, 256, 4096},
                        {16, 2, 512, 4096}, {32, 1, 512, 4096}, {16, 1, 512, 8192}, {16, 2, 1024, 8192},
                        {32, 2, 1024, 8192}, {16, 0, 0, -1},    {32, 0, 0, -1},    {64, 0, 0, -1},
                        {16, 8, 0, 0},     {32, 8, 0, 0},      {64, 8, 0, 0},      {16, 16, 0, 0},
                        {32, 16, 0, 0},    {32, 16, 64, 0},    {32, 16, 256, 0},   {16, 16, 128, 0},
                        {32, 8, 256, 0},   {32, 16, 2048, -512}, {64, 8, 2048, -512}};
  if (c->n < 4096) return SGTSNE_OK;  // the heuristic choice is fine for small jobs
  // the column-localised candidates run only when their blocks exist (built
  // when the caller pins that form): on kNN graphs of high-dimensional data
  // no order gives their blocks enough column reuse to pay for the build
  hipEvent_t e0, e1;
  SGT_HIP(hipEventCreate(&e0));
  SGT_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bg = c->seg_group, bu = c->seg_unroll, brpb = c->seg_rpb, bw = c->seg_win;
  c->tune_ms.clear();
  c->tune_cands.clear();
  int st = SGTSNE_OK;
  for (const Cand &k : cands) {
    if (k.win > 0 && (size_t)k.win * c->ds * sizeof(float) > 65536) continue;  // window must fit LDS
    if (k.win == -1 && c->loc_nblk == 0) continue;                             // no localised blocks
    const int chs = k.win <= -2 ? log2_exact(-k.win) : 0;
    if (k.win <= -2) {  // chunk-cached: its table for these ranges and chunks
      if (!cc_fits(c, chs) || k.rpb >= c->n) continue;
      st = build_chunk_cache(c, k.rpb, chs);
      if (st != SGTSNE_OK) break;
    }
    float ms = 0.f, tot = 0.f;
    for (int rep = 0; rep < 3 && st == SGTSNE_OK; rep++) {
      (void)hipEventRecord(e0, c->stream);
      st = k.win <= -2 ? launch_attr_csr_cc(c->stream, c->d, k.g, k.u, k.rpb, chs, c->cc_nslots, c->cc_slots, c->n,
                                            c->rowptr, c->col, c->val, c->Y, c->Fattr)
           : k.win < 0   ? launch_attr_csr_loc(c->stream, c->d, k.g, c->loc_nblk, c->loc_cap, c->loc_brow, c->loc_bu,
                                             c->loc_ucol, c->rowptr, c->loc_lidx, c->val, c->Y, c->Fattr)
           : k.win > 0 ? launch_attr_csr_win(c->stream, c->d, k.g, k.u, k.rpb, k.win, c->n, c->rowptr, c->col,
                                             c->val, c->Y, c->Fattr)
           : k.rpb > 0 ? launch_attr_csr_blk(c->stream, c->d, k.g, k.u, k.rpb, c->n, c->rowptr, c->col, c->val,
                                             c->Y, c->Fattr)
           : k.u > 0   ? launch_attr_csr_seg2(c->stream, c->d, k.g, k.u, c->n, c->rowptr, c->col, c->val, c->Y,
                                              c->Fattr)
                       : launch_attr_csr_seg(c->stream, c->d, k.g, c->n, c->rowptr, c->col, c->val, c->Y, c->Fattr);
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0) tot += ms;  // the first run warms the caches
    }
    if (st != SGTSNE_OK) break;
    c->tune_ms.push_back(tot / 2);
    c->tune_cands.push_back({k.g, k.u, k.rpb, k.win});
    if (tot / 2 < best) { best = tot / 2; bg = k.g; bu = k.u; brpb = k.rpb; bw = k.win; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SGT_TRY(st);
  // the winner's chunk table (the last one built may belong to another candidate)
  if (bw <= -2) SGT_TRY(build_chunk_cache(c, brpb, log2_exact(-bw)));
  else free_dev(c->cc_slots), c->cc_rpb = c->cc_chs = c->cc_nslots = 0;
  c->seg_group = bg;
  c->seg_unroll = bu;
  c->seg_rpb = brpb;
  c->seg_win = bw;
  return SGTSNE_OK;
}

int build_formats(sgtsne_ctx *c) {
  if (c->p.format == SGTSNE_FORMAT_COO) {
    SGT_TRY(alloc_dev((void **)&c->coo_row, sizeof(int32_t) * (size_t)std::max<int64_t>(c->nnz, 1)));
    SGT_TRY(launch_rowids(c->stream, c->n, c->rowptr, c->coo_row));
  } else if (c->p.format == SGTSNE_FORMAT_HYBRID) {
    std::vector<int32_t> rp((size_t)c->n + 1);
    SGT_HIP(hipMemcpyAsync(rp.data(), c->rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, c->stream));
    SGT_HIP(hipStreamSynchronize(c->stream));
    int w = c->p.ell_width > 0 ? c->p.ell_width : optimal_ell_width(c->n, rp);
    c->ell_w = w;
    int64_t tail = 0;
    for (int i = 0; i < c->n; i++) tail += std::max(0, rp[i + 1] - rp[i] - w);
    c->tail_nnz = tail;
    SGT_TRY(alloc_dev((void **)&c->ell_col, sizeof(int32_t) * (size_t)w * c->n));
    SGT_TRY(alloc_dev((void **)&c->ell_val, sizeof(float) * (size_t)w * c->n));
    int32_t *cnt = nullptr, *off = nullptr;
    SGT_TRY(alloc_dev((void **)&cnt, sizeof(int32_t) * (size_t)c->n));
    SGT_TRY(alloc_dev((void **)&off, sizeof(int32_t) * (size_t)c->n));
    SGT_TRY(alloc_dev((void **)&c->tail_row, sizeof(int32_t) * (size_t)std::max<int64_t>(tail, 1)));
    SGT_TRY(alloc_dev((void **)&c->tail_col, sizeof(int32_t) * (size_t)std::max<int64_t>(tail, 1)));
    SGT_TRY(alloc_dev((void **)&c->tail_val, sizeof(float) * (size_t)std::max<int64_t>(tail, 1)));
    int st = launch_build_ell(c->stream, c->n, w, c->rowptr, c->col, c->val, c->ell_col, c->ell_val, cnt);
    if (st == SGTSNE_OK) st = device_exclusive_scan(c->stream, cnt, off, c->n);
    if (st == SGTSNE_OK)
      st = launch_build_tail(c->stream, c->n, w, c->rowptr, c->col, c->val, off, c->tail_row, c->tail_col, c->tail_val);
    (void)hipStreamSynchronize(c->stream);
    free_dev(cnt);
    free_dev(off);
    SGT_TRY(st);
  }
  // segmented CSR: lanes per row from the mean row length (power of two, 4..64)
  int g = 4;
  const double mean = c->n > 0 ? (double)c->nnz / c->n : 0.0;
  while (g < 64 && g * 6 < mean) g <<= 1;
  c->seg_group = g;
  return SGTSNE_OK;
}

}  // namespace

extern "C" {

int sgtsne_params_default(sgtsne_params *p) {
  SGT_CHECK(p != nullptr, SGTSNE_ERR_ARG, "params is NULL");
  std::memset(p, 0, sizeof(*p));
  p->d = 2;
  p->lambda = 1.0;
  p->alpha = 12.0;
  p->max_iter = 1000;
  p->early_iter = 250;
  p->n = 0;
  p->h = -1.0;
  p->drop_leaf = 0;
  p->np = 0;
  p->format = SGTSNE_FORMAT_CSR_SEG;
  p->bs = 0;
  p->compute_error = -1;
  p->ng = 50;
  p->grid_mode = SGTSNE_GRID_H;
  p->layout = SGTSNE_LAYOUT_SOA;
  p->seed = 0;
  p->ell_width = 0;
  p->zeta_mode = SGTSNE_ZETA_DIRECT;
  p->reorder = 2;
  p->spread = SGTSNE_SPREAD_AUTO;
  return SGTSNE_OK;
}

const char *sgtsne_last_error(void) { return g_err; }

const char *sgtsne_version(void) { return "sgtsne-mi355x 0.1 (gfx950, HIP)"; }

int sgtsne_ctx_create(sgtsne_ctx **out, const sgtsne_params *params) {
  SGT_CHECK(out != nullptr, SGTSNE_ERR_ARG, "ctx pointer is NULL");
  *out = nullptr;
  SGT_TRY(check_params(params));
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  SGT_CHECK(e == hipSuccess && ndev > 0, SGTSNE_ERR_NODEV, "no HIP device visible");
  sgtsne_ctx *c = new sgtsne_ctx();
  c->p = *params;
  c->d = params->d;
  c->ds = point_stride(c->d);
  (void)hipGetDevice(&c->device);
  int st = SGTSNE_OK;
  auto fail = [&](int code) {
    sgtsne_ctx_destroy(c);
    return code;
  };
  retain_pool_memory();
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(SGTSNE_ERR_HIP);
  }
  c->rw.max_blocks = 65536;
  if ((st = alloc_dev((void **)&c->rw.part, sizeof(double) * 8 * c->rw.max_blocks))) return fail(st);
  if ((st = alloc_dev((void **)&c->rw.ticket, sizeof(unsigned) * 4))) return fail(st);
  if ((st = alloc_dev((void **)&c->sc, sizeof(DevScalars)))) return fail(st);
  if (hipMemset(c->rw.ticket, 0, sizeof(unsigned) * 4) != hipSuccess ||
      hipMemset(c->sc, 0, sizeof(DevScalars)) != hipSuccess) {
    set_error("hipMemset failed");
    return fail(SGTSNE_ERR_HIP);
  }
  if (hipHostMalloc((void **)&c->sc_host, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    return fail(SGTSNE_ERR_NOMEM);
  }
  std::memset(c->sc_host, 0, sizeof(DevScalars));
  if (hipEventCreateWithFlags(&c->ev_scal, hipEventDisableTiming) != hipSuccess) {
    set_error("hipEventCreate failed");
    return fail(SGTSNE_ERR_HIP);
  }
  for (auto &pe : c->pe)
    for (auto &ev : pe.e)
      if (hipEventCreate(&ev) != hipSuccess) {
        set_error("hipEventCreate failed");
        return fail(SGTSNE_ERR_HIP);
      }
  if ((st = alloc_dev((void **)&c->cb.tw, sizeof(float2) * twiddle_count()))) return fail(st);
  if ((st = init_twiddles(c->stream, c->cb.tw))) return fail(st);
  *out = c;
  return SGTSNE_OK;
}

int sgtsne_ctx_destroy(sgtsne_ctx *c) {
  if (!c) return SGTSNE_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_P(c);
  free_dev(c->Y);
  free_dev(c->uY);
  free_dev(c->gains);
  free_dev(c->Fattr);
  free_dev(c->Frep);
  free_dev(c->rw.part);
  free_dev(c->rw.ticket);
  free_dev(c->sc);
  free_dev(c->V);
  free_dev(c->Phi);
  free_dev(c->s2g_part);
  free_dev(c->cb.W);
  free_dev(c->cb.Kt1);
  free_dev(c->cb.Kt2);
  free_dev(c->cb.Kh);
  free_dev(c->cb.tw);
  if (c->sc_host) (void)hipHostFree(c->sc_host);
  if (c->ev_scal) (void)hipEventDestroy(c->ev_scal);
  for (auto &pe : c->pe)
    for (auto &ev : pe.e)
      if (ev) (void)hipEventDestroy(ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SGTSNE_OK;
}

int sgtsne_ctx_upload_P(sgtsne_ctx *c, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                        const float *val, int symmetrize) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  SGT_CHECK(n > 0 && nnz >= 0 && rowptr && (nnz == 0 || (col && val)), SGTSNE_ERR_ARG, "bad P arguments");
  SGT_CHECK(nnz < (int64_t)INT32_MAX, SGTSNE_ERR_ARG, "nnz exceeds the int32 index range");
  SGT_CHECK(rowptr[0] == 0 && rowptr[n] == nnz, SGTSNE_ERR_ARG, "rowptr inconsistent with nnz");
  // every index is checked on the host: the kernels gather Y[col] unguarded
  for (int i = 0; i < n; i++)
    SGT_CHECK(rowptr[i] <= rowptr[i + 1], SGTSNE_ERR_ARG, "rowptr decreases at row %d", i);
  {
    int bad = 0;
    for (int64_t t = 0; t < nnz; t++) bad += (col[t] < 0 || col[t] >= n);
    SGT_CHECK(bad == 0, SGTSNE_ERR_ARG, "%d column indices out of range [0, %d)", bad, n);
  }
  if (c->have_Y && c->n != n) {
    set_error("P has %d rows but Y has %d points", n, c->n);
    return SGTSNE_ERR_ARG;
  }
  free_P(c);
  c->n = n;
  c->have_Y = false;  // the device order of Y follows P: set Y again
  c->permuted = false;
  c->perm.clear();
  c->iperm.clear();
  int32_t *rp = nullptr, *cc = nullptr;
  float *vv = nullptr;
  SGT_TRY(alloc_dev((void **)&rp, sizeof(int32_t) * ((size_t)n + 1)));
  SGT_TRY(alloc_dev((void **)&cc, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
  SGT_TRY(alloc_dev((void **)&vv, sizeof(float) * (size_t)std::max<int64_t>(nnz, 1)));
  SGT_HIP(hipMemcpyAsync(rp, rowptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, c->stream));
  if (nnz > 0) {
    SGT_HIP(hipMemcpyAsync(cc, col, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, c->stream));
    SGT_HIP(hipMemcpyAsync(vv, val, sizeof(float) * (size_t)nnz, hipMemcpyHostToDevice, c->stream));
  }
  if (symmetrize) {
    if (c->p.lambda != 1.0) SGT_TRY(launch_lambda_rescale(c->stream, n, rp, vv, c->p.lambda, c->p.drop_leaf));
    int64_t m = 0;
    int st = gpu_symmetrize(c->stream, n, nnz, rp, cc, vv, &c->rowptr, &c->col, &c->val, &m);
    (void)hipStreamSynchronize(c->stream);
    free_dev(rp);
    free_dev(cc);
    free_dev(vv);
    SGT_TRY(st);
    c->nnz = m;
  } else {
    c->rowptr = rp;
    c->col = cc;
    c->val = vv;
    c->nnz = nnz;
  }
  if (c->p.reorder) SGT_TRY(apply_locality_order(c));
  SGT_TRY(build_formats(c));
  SGT_TRY(alloc_points(c));
  if (c->p.format == SGTSNE_FORMAT_CSR_SEG) SGT_TRY(tune_attractive(c));
  SGT_HIP(hipStreamSynchronize(c->stream));
  c->have_P = true;
  return SGTSNE_OK;
}

int sgtsne_ctx_get_P(sgtsne_ctx *c, int64_t *nnz, int32_t *rowptr, int32_t *col, float *val) {
  SGT_CHECK(c != nullptr && nnz != nullptr, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(c->have_P, SGTSNE_ERR_STATE, "no P uploaded");
  *nnz = c->nnz;
  if (!rowptr && !col && !val) return SGTSNE_OK;
  const int n = c->n;
  std::vector<int32_t> rp((size_t)n + 1), cc((size_t)std::max<int64_t>(c->nnz, 1));
  std::vector<float> vv((size_t)std::max<int64_t>(c->nnz, 1));
  SGT_HIP(hipMemcpy(rp.data(), c->rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost));
  if (c->nnz > 0) {
    SGT_HIP(hipMemcpy(cc.data(), c->col, sizeof(int32_t) * (size_t)c->nnz, hipMemcpyDeviceToHost));
    SGT_HIP(hipMemcpy(vv.data(), c->val, sizeof(float) * (size_t)c->nnz, hipMemcpyDeviceToHost));
  }
  if (!c->permuted) {
    if (rowptr) std::copy(rp.begin(), rp.end(), rowptr);
    if (col) std::copy(cc.begin(), cc.begin() + c->nnz, col);
    if (val) std::copy(vv.begin(), vv.begin() + c->nnz, val);
    return SGTSNE_OK;
  }
  // back to the caller's order, columns ascending within each row
  std::vector<int32_t> orp((size_t)n + 1, 0);
  for (int i = 0; i < n; i++) {
    const int r = c->iperm[i];
    orp[i + 1] = orp[i] + (rp[r + 1] - rp[r]);
  }
  std::vector<std::pair<int32_t, float>> row;
  for (int i = 0; i < n; i++) {
    const int r = c->iperm[i];
    row.clear();
    for (int t = rp[r]; t < rp[r + 1]; t++) row.emplace_back(c->perm[cc[t]], vv[t]);
    std::sort(row.begin(), row.end(), [](const std::pair<int32_t, float> &a, const std::pair<int32_t, float> &b) {
      return a.first < b.first;
    });
    for (size_t e = 0; e < row.size(); e++) {
      if (col) col[orp[i] + e] = row[e].first;
      if (val) val[orp[i] + e] = row[e].second;
    }
  }
  if (rowptr) std::copy(orp.begin(), orp.end(), rowptr);
  return SGTSNE_OK;
}

int sgtsne_ctx_set_Y(sgtsne_ctx *c, const float *y, int layout, int reset_state) {
  SGT_CHECK(c != nullptr && y != nullptr, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(layout == SGTSNE_LAYOUT_SOA || layout == SGTSNE_LAYOUT_AOS, SGTSNE_ERR_ARG, "bad layout");
  int n = c->n;
  SGT_CHECK(n > 0, SGTSNE_ERR_STATE, "upload P before Y (n unknown)");
  const size_t nb = (size_t)n * c->ds;
  SGT_TRY(alloc_points(c));
  if (!c->have_Y) reset_state = 1;
  std::vector<float> h;
  pack_y(y, layout, n, c->d, c->ds, c->perm, h);
  double mean[4] = {0, 0, 0, 0};
  for (int a = 0; a < c->d; a++) {
    double s = 0;
    for (int i = 0; i < n; i++) s += h[(size_t)i * c->ds + a];
    mean[a] = s / n;
  }
  SGT_HIP(hipMemcpyAsync(c->Y, h.data(), nb * sizeof(float), hipMemcpyHostToDevice, c->stream));
  SGT_HIP(hipStreamSynchronize(c->stream));
  DevScalars s{};
  SGT_HIP(hipMemcpy(&s, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
  for (int a = 0; a < 4; a++) s.mean[a] = mean[a];
  SGT_HIP(hipMemcpy(c->sc, &s, sizeof(DevScalars), hipMemcpyHostToDevice));
  if (reset_state) {
    SGT_TRY(launch_fill(c->stream, c->uY, (int64_t)nb, 0.0f));
    SGT_TRY(launch_fill(c->stream, c->gains, (int64_t)nb, 1.0f));
    c->iter = 0;
  }
  SGT_TRY(center_and_fetch(c, 1));
  SGT_HIP(hipStreamSynchronize(c->stream));
  c->have_Y = true;
  return SGTSNE_OK;
}

int sgtsne_ctx_get_Y(sgtsne_ctx *c, float *y, int layout) {
  SGT_CHECK(c != nullptr && y != nullptr, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(c->have_Y, SGTSNE_ERR_STATE, "Y not set");
  std::vector<float> h((size_t)c->n * c->ds);
  SGT_HIP(hipStreamSynchronize(c->stream));
  SGT_HIP(hipMemcpy(h.data(), c->Y, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  unpack_y(h, layout, c->n, c->d, c->ds, c->perm, y);
  return SGTSNE_OK;
}

int sgtsne_ctx_step(sgtsne_ctx *c, int iters) {
  SGT_CHECK(c != nullptr && iters >= 0, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(c->have_P && c->have_Y, SGTSNE_ERR_STATE, "upload P and set Y before stepping");
  for (auto &v : c->acc_ms) v = 0.0;
  for (auto &pe : c->pe) pe.valid = false;
  int st = SGTSNE_OK;
  for (int k = 0; k < iters && st == SGTSNE_OK; k++) st = iterate(c);
  SGT_HIP(hipStreamSynchronize(c->stream));
  if (c->timing && c->iter > 0) harvest_timing(c, c->pe[(c->iter - 1) & 1], c->iter - 1);
  return st;
}

int sgtsne_ctx_terms(sgtsne_ctx *c, float *Fattr, float *Frep, double *Z, int *ng, int layout) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  SGT_CHECK(c->have_P && c->have_Y, SGTSNE_ERR_STATE, "upload P and set Y first");
  SGT_HIP(hipStreamSynchronize(c->stream));
  SGT_TRY(center_and_fetch(c, 0));
  SGT_HIP(hipEventSynchronize(c->ev_scal));
  SGT_TRY(attractive(c));
  GridParams g{};
  SGT_TRY(make_grid(c, *c->sc_host, g));
  SGT_TRY(repulsive(c, g, nullptr));
  SGT_HIP(hipStreamSynchronize(c->stream));
  DevScalars s{};
  SGT_HIP(hipMemcpy(&s, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
  const size_t nb = (size_t)c->n * c->ds;
  std::vector<float> h(nb);
  if (Fattr) {
    SGT_HIP(hipMemcpy(h.data(), c->Fattr, nb * sizeof(float), hipMemcpyDeviceToHost));
    unpack_y(h, layout, c->n, c->d, c->ds, c->perm, Fattr);
  }
  if (Frep) {
    SGT_HIP(hipMemcpy(h.data(), c->Frep, nb * sizeof(float), hipMemcpyDeviceToHost));
    unpack_y(h, layout, c->n, c->d, c->ds, c->perm, Frep, 1.0 / s.Z);
  }
  if (Z) *Z = s.Z;
  if (ng) *ng = g.ng;
  return SGTSNE_OK;
}

int sgtsne_ctx_kl(sgtsne_ctx *c, double alpha, double *kl) {
  SGT_CHECK(c != nullptr && kl != nullptr, SGTSNE_ERR_ARG, "bad arguments");
  SGT_CHECK(c->have_P && c->have_Y, SGTSNE_ERR_STATE, "upload P and set Y first");
  // Z of the current Y
  SGT_TRY(sgtsne_ctx_terms(c, nullptr, nullptr, nullptr, nullptr, SGTSNE_LAYOUT_AOS));
  SGT_TRY(launch_kl(c->stream, c->d, c->n, c->rowptr, c->col, c->val, c->Y, alpha, c->sc, c->rw, c->sc));
  SGT_HIP(hipStreamSynchronize(c->stream));
  DevScalars s{};
  SGT_HIP(hipMemcpy(&s, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
  *kl = s.kl;
  return SGTSNE_OK;
}

int sgtsne_ctx_repulsive_exact(sgtsne_ctx *c, float *Frep, double *Z, int layout) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  SGT_CHECK(c->have_Y, SGTSNE_ERR_STATE, "Y not set");
  SGT_HIP(hipStreamSynchronize(c->stream));
  const size_t nb = (size_t)c->n * c->ds;
  float *tmp = nullptr;
  SGT_TRY(alloc_dev((void **)&tmp, nb * sizeof(float)));
  int st = launch_rep_exact(c->stream, c->d, c->n, c->Y, tmp, c->rw, c->sc);
  std::vector<float> h(nb);
  DevScalars s{};
  if (st == SGTSNE_OK) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipMemcpy(h.data(), tmp, nb * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipMemcpy(&s, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost);
  }
  free_dev(tmp);
  SGT_TRY(st);
  if (Frep) unpack_y(h, layout, c->n, c->d, c->ds, c->perm, Frep, 1.0 / s.zexact);
  if (Z) *Z = s.zexact;
  return SGTSNE_OK;
}

int sgtsne_ctx_timing(sgtsne_ctx *c, double *out7) {
  SGT_CHECK(c != nullptr && out7 != nullptr, SGTSNE_ERR_ARG, "bad arguments");
  for (int k = 0; k < 7; k++) out7[k] = c->acc_ms[k];
  return SGTSNE_OK;
}

int sgtsne_ctx_info(sgtsne_ctx *c, int *iter, int *n, int *d, int64_t *nnz) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  if (iter) *iter = c->iter;
  if (n) *n = c->n;
  if (d) *d = c->d;
  if (nnz) *nnz = c->nnz;
  return SGTSNE_OK;
}

int sgtsne_ctx_tuning(sgtsne_ctx *c, int *group, int *unroll, int *rows_per_block, int *window, double *ms,
                      int max, int *count) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  if (rows_per_block) *rows_per_block = c->seg_rpb;
  if (window) *window = c->seg_win;
  if (group) *group = c->seg_group;
  if (unroll) *unroll = c->seg_unroll;
  if (count) *count = (int)c->tune_ms.size();
  if (ms)
    for (int k = 0; k < std::min(max, (int)c->tune_ms.size()); k++) ms[k] = c->tune_ms[k];
  return SGTSNE_OK;
}

int sgtsne_ctx_set_attractive(sgtsne_ctx *c, int group, int unroll, int rows_per_block, int window) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  SGT_CHECK(c->have_P && c->p.format == SGTSNE_FORMAT_CSR_SEG, SGTSNE_ERR_STATE,
            "pinning the attractive kernel needs P uploaded in the CSR_SEG format");
  const int chs = window <= -2 ? log2_exact(-window) : 0;
  SGT_CHECK(unroll >= 0 && rows_per_block >= 0 &&
                (window >= 0    ? (window == 0 || rows_per_block > 0)
                 : window == -1 ? (unroll == 0 && rows_per_block == 0)
                                : (rows_per_block > 0 && chs >= 6 && chs <= 16)),
            SGTSNE_ERR_ARG, "bad attractive kernel shape");
  if (window <= -2) SGT_TRY(build_chunk_cache(c, rows_per_block, chs));
  if (window == -1 && c->loc_nblk == 0) {
    SGT_TRY(build_localised(c));
    SGT_CHECK(c->loc_nblk > 0, SGTSNE_ERR_ARG, "a row is wider than the column-localised blocks");
  }
  const int keep[4] = {c->seg_group, c->seg_unroll, c->seg_rpb, c->seg_win};
  c->seg_group = group;
  c->seg_unroll = unroll;
  c->seg_rpb = rows_per_block;
  c->seg_win = window;
  // one launch checks that the shape exists (it only rewrites Fattr)
  int st = attractive(c);
  if (st == SGTSNE_OK && hipStreamSynchronize(c->stream) != hipSuccess) {
    set_error("attractive kernel failed");
    st = SGTSNE_ERR_HIP;
  }
  if (st != SGTSNE_OK) {
    c->seg_group = keep[0];
    c->seg_unroll = keep[1];
    c->seg_rpb = keep[2];
    c->seg_win = keep[3];
    return st;
  }
  c->tune_pinned = true;
  return SGTSNE_OK;
}

int sgtsne_ctx_reorders(sgtsne_ctx *c, int *applied, int *declined) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  if (applied) *applied = c->reorders;
  if (declined) *declined = c->reorders_declined;
  return SGTSNE_OK;
}

int sgtsne_ctx_set_timing(sgtsne_ctx *c, int on) {
  SGT_CHECK(c != nullptr, SGTSNE_ERR_ARG, "ctx is NULL");
  c->timing = on != 0;
  return SGTSNE_OK;
}

int sgtsne_ctx_errors(sgtsne_ctx *c, double *out, int max, int *count) {
  SGT_CHECK(c != nullptr && count != nullptr, SGTSNE_ERR_ARG, "bad arguments");
  *count = (int)c->errors.size();
  if (out)
    for (int k = 0; k < std::min(max, *count); k++) out[k] = c->errors[k];
  return SGTSNE_OK;
}

int sgtsne_run(int n, int64_t nnz, const int32_t *rowptr, const int32_t *col, const float *val,
               const sgtsne_params *params, const float *y_in, float *y_out, double *time_info) {
  SGT_TRY(check_params(params));
  SGT_CHECK(y_out != nullptr, SGTSNE_ERR_ARG, "y_out is NULL");
  sgtsne_ctx *c = nullptr;
  SGT_TRY(sgtsne_ctx_create(&c, params));
  int st = sgtsne_ctx_upload_P(c, n, nnz, rowptr, col, val, 1);
  std::vector<float> y0;
  const float *yi = y_in;
  if (st == SGTSNE_OK && !y_in) {
    // N(0, 1e-4^2) start: seed 0 draws the reference's own sequence
    // (CU/utils.cu:185-192); another seed draws from a seeded mt19937_64
    y0.resize((size_t)n * params->d);
    if (params->seed == 0) {
      st = sgtsne_random_start(n, params->d, params->layout, y0.data());
    } else {
      std::mt19937_64 gen(params->seed);
      std::normal_distribution<double> nd(0.0, 1.0);
      for (auto &v : y0) v = (float)(1e-4 * nd(gen));
    }
    yi = y0.data();
  }
  if (st == SGTSNE_OK) st = sgtsne_ctx_set_Y(c, yi, params->layout, 1);
  if (st == SGTSNE_OK) {
    c->time_info = time_info;
    c->time_info_base = 0;
    c->timing = true;
    auto t0 = std::chrono::steady_clock::now();
    st = sgtsne_ctx_step(c, params->max_iter);
    auto t1 = std::chrono::steady_clock::now();
    if (time_info) time_info[7 * (size_t)params->max_iter] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  if (st == SGTSNE_OK) st = sgtsne_ctx_get_Y(c, y_out, params->layout);
  sgtsne_ctx_destroy(c);
  return st;
}

int sgtsne_locality_order(int n, const int32_t *rowptr, const int32_t *col, int32_t *perm) {
  SGT_CHECK(n > 0 && rowptr && perm && (rowptr[n] == 0 || col), SGTSNE_ERR_ARG, "bad arguments");
  for (int i = 0; i < n; i++) SGT_CHECK(rowptr[i] <= rowptr[i + 1], SGTSNE_ERR_ARG, "rowptr decreases at row %d", i);
  for (int64_t t = 0; t < rowptr[n]; t++) SGT_CHECK(col[t] >= 0 && col[t] < n, SGTSNE_ERR_ARG, "column out of range");
  std::vector<int32_t> p;
  bfs_order(n, rowptr, col, p);
  std::copy(p.begin(), p.end(), perm);
  return SGTSNE_OK;
}

int sgtsne_copy_probe(int64_t bytes, int reps, double *ms_per_copy) {
  SGT_CHECK(bytes >= 16 && bytes % 16 == 0 && reps >= 1 && ms_per_copy, SGTSNE_ERR_ARG, "bad probe arguments");
  hipStream_t s = nullptr;
  SGT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = alloc_dev((void **)&a, (size_t)bytes);
  if (!st) st = alloc_dev((void **)&b, (size_t)bytes);
  if (!st && (hipMemsetAsync(a, 0, (size_t)bytes, s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
              hipEventCreate(&e1) != hipSuccess)) {
    set_error("probe set-up failed");
    st = SGTSNE_ERR_HIP;
  }
  const int64_t cnt = bytes / 16;
  if (!st) st = launch_stream_copy(s, a, b, cnt);  // warm
  if (!st) (void)hipEventRecord(e0, s);
  for (int r = 0; r < reps && !st; r++) st = launch_stream_copy(s, a, b, cnt);
  if (!st) (void)hipEventRecord(e1, s);
  if (hipStreamSynchronize(s) != hipSuccess && !st) {
    set_error("probe copy failed");
    st = SGTSNE_ERR_HIP;
  }
  float t = 0.f;
  if (!st) (void)hipEventElapsedTime(&t, e0, e1);
  *ms_per_copy = t / reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  free_dev(a);
  free_dev(b);
  (void)hipStreamDestroy(s);
  return st;
}

int sgtsne_knn(const float *X, int n, int dim, int k, int method, int32_t *I, double *D, int *fallback_rows) {
  SGT_CHECK(X && I && D, SGTSNE_ERR_ARG, "NULL argument");
  SGT_CHECK(method >= SGTSNE_KNN_AUTO && method <= SGTSNE_KNN_FILTERED, SGTSNE_ERR_ARG, "unknown kNN method %d", method);
  SGT_CHECK(n > k && k >= 1 && dim >= 1, SGTSNE_ERR_ARG, "need n > k >= 1 and dim >= 1");
  hipStream_t s = nullptr;
  SGT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const size_t nx = (size_t)n * dim;
  std::vector<double> xd(nx);
  for (size_t t = 0; t < nx; t++) xd[t] = (double)X[t];
  double *Xd = nullptr, *Dd = nullptr;
  int32_t *Id = nullptr;
  int st = alloc_dev((void **)&Xd, nx * sizeof(double));
  if (!st) st = alloc_dev((void **)&Dd, (size_t)n * (k + 1) * sizeof(double));
  if (!st) st = alloc_dev((void **)&Id, (size_t)n * (k + 1) * sizeof(int32_t));
  if (!st && hipMemcpyAsync(Xd, xd.data(), nx * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error("copy of X failed");
    st = SGTSNE_ERR_HIP;
  }
  if (!st) st = launch_knn(s, Xd, n, dim, k, Id, Dd, method, fallback_rows);
  if (!st && (hipMemcpyAsync(I, Id, (size_t)n * (k + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipMemcpyAsync(D, Dd, (size_t)n * (k + 1) * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)) {
    set_error("copy of kNN result failed");
    st = SGTSNE_ERR_HIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && !st) {
    set_error("kNN kernel failed");
    st = SGTSNE_ERR_HIP;
  }
  free_dev(Xd);
  free_dev(Dd);
  free_dev(Id);
  (void)hipStreamDestroy(s);
  return st;
}

int sgtsne_knn_exact(const float *X, int n, int dim, int k, int32_t *I, double *D) {
  return sgtsne_knn(X, n, dim, k, SGTSNE_KNN_AUTO, I, D, nullptr);
}

int sgtsne_grid_conv(int d, int ng, double h, int zdirect, const float *V, float *Phi) {
  SGT_CHECK(V && Phi, SGTSNE_ERR_ARG, "NULL argument");
  SGT_CHECK(d >= 1 && d <= 3 && ng >= 4 && ng <= 512 && h > 0, SGTSNE_ERR_ARG,
            "bad grid_conv sizes (d=%d ng=%d)", d, ng);
  const int m = d + 1, mo = zdirect ? d + 2 : d + 1;
  GridParams g{};
  g.d = d;
  g.ng = ng;
  g.nGrid = ng - 2;
  const int L = fft_length_for(ng);
  g.L = L;
  g.hgrid = h;
  const size_t G = d == 1 ? (size_t)ng : d == 2 ? (size_t)ng * ng : (size_t)ng * ng * ng;
  size_t wb = 0, kb = 0;
  conv_workspace_bytes(d, ng, L, &wb, &kb);
  hipStream_t s = nullptr;
  SGT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  ConvBuffers cb{};
  float *Vd = nullptr, *Pd = nullptr;
  int st = alloc_dev((void **)&Vd, G * m * sizeof(float));
  if (!st) st = alloc_dev((void **)&Pd, G * mo * sizeof(float));
  if (!st) st = alloc_dev((void **)&cb.W, wb);
  if (!st) st = alloc_dev((void **)&cb.Kt1, kb);
  if (!st) st = alloc_dev((void **)&cb.Kt2, kb);
  if (!st) st = alloc_dev((void **)&cb.Kh, kb);
  if (!st) st = alloc_dev((void **)&cb.tw, twiddle_count() * sizeof(float2));
  if (!st) st = init_twiddles(s, cb.tw);
  if (!st && hipMemcpyAsync(Vd, V, G * m * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error("copy of V failed");
    st = SGTSNE_ERR_HIP;
  }
  if (!st) st = launch_conv(s, g, zdirect, Vd, Pd, cb);
  if (!st && hipMemcpyAsync(Phi, Pd, G * mo * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess) {
    set_error("copy of Phi failed");
    st = SGTSNE_ERR_HIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && !st) {
    set_error("convolution failed");
    st = SGTSNE_ERR_HIP;
  }
  free_dev(Vd);
  free_dev(Pd);
  free_dev(cb.W);
  free_dev(cb.Kt1);
  free_dev(cb.Kt2);
  free_dev(cb.Kh);
  free_dev(cb.tw);
  (void)hipStreamDestroy(s);
  return st;
}

int sgtsne_perplexity_equalize(const int32_t *I, const double *D, int n, int k, double u,
                               int32_t *rowptr, int32_t *col, float *val) {
  SGT_CHECK(I && D && rowptr && col && val, SGTSNE_ERR_ARG, "NULL argument");
  SGT_CHECK(n > 0 && k >= 1 && u > 0, SGTSNE_ERR_ARG, "bad sizes");
  hipStream_t s = nullptr;
  SGT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  double *Dd = nullptr;
  float *vd = nullptr;
  int st = alloc_dev((void **)&Dd, (size_t)n * (k + 1) * sizeof(double));
  if (!st) st = alloc_dev((void **)&vd, (size_t)n * k * sizeof(float));
  if (!st && hipMemcpyAsync(Dd, D, (size_t)n * (k + 1) * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error("copy of D failed");
    st = SGTSNE_ERR_HIP;
  }
  if (!st) st = launch_perplexity(s, Dd, n, k, u, vd, nullptr);
  if (!st && hipMemcpyAsync(val, vd, (size_t)n * k * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess) {
    set_error("copy of P failed");
    st = SGTSNE_ERR_HIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && !st) {
    set_error("perplexity kernel failed");
    st = SGTSNE_ERR_HIP;
  }
  free_dev(Dd);
  free_dev(vd);
  (void)hipStreamDestroy(s);
  SGT_TRY(st);
  // kNN ids -> CSR without the self column (CU/perplexity_search.cu:100-112)
  for (int i = 0; i < n; i++) {
    rowptr[i] = i * k;
    for (int e = 0; e < k; e++) col[(size_t)i * k + e] = I[(size_t)i * (k + 1) + e + 1];
  }
  rowptr[n] = n * k;
  return SGTSNE_OK;
}

}  // extern "C"
