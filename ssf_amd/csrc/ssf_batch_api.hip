// C ABI of the batched solver: many assembled matrices (independent systems, each
// with R right-hand sides) solved together. The whole solve is one hipGraph
// launch: the graph holds as many iteration chunks as the batch's previous solve
// needed (plus one), each chunk ending with the stop tests and a new work list
// of the systems still running, so the persistent kernels skip finished systems
// and the chunks after the last pair has stopped cost a few dispatches each. A
// batch's first solve, and one that needs more chunks, replays a one-chunk graph
// from a host loop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ssf_batch.hpp"

using namespace ssf;

namespace {

template <typename T>
hipError_t balloc(T **p, size_t count)
{
  *p = nullptr;
  return hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}
template <typename T>
void bfree(T *&p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
struct StateArrays {
  BatchState<T> s;
  hipError_t alloc(int64_t nq, int64_t nsys)
  {
    hipError_t e = hipSuccess;
    auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    c(balloc(&s.nory, nq)); c(balloc(&s.nom, 2 * nq)); c(balloc(&s.iter, nq));
    c(balloc(&s.done, nq)); c(balloc(&s.status, nq)); c(balloc(&s.pend, nq));
    c(balloc(&s.pend_status, nq)); c(balloc(&s.brk, nq)); c(balloc(&s.ani, nq));
    c(balloc(&s.ares, nq)); c(balloc(&s.sys_done, nsys));
    return e;
  }
  void release()
  {
    bfree(s.nory); bfree(s.nom); bfree(s.iter); bfree(s.done); bfree(s.status); bfree(s.pend);
    bfree(s.pend_status); bfree(s.brk); bfree(s.ani); bfree(s.ares); bfree(s.sys_done);
  }
};

}  // namespace

constexpr int HIST_CAP = 1 << 16;  // chunks recorded per solve

struct ssf_batch {
  ssf_ctx *ctx = nullptr;
  int nsys = 0, R = 1, nblk = 0;
  int64_t nrows = 0, nnz = 0;
  std::vector<int64_t> row0;  // first global row of each system, nsys + 1
  int *rowptr = nullptr, *col = nullptr;
  double *val = nullptr, *diag = nullptr;
  float *val32 = nullptr;  // float(S A S), S = diag(1/sqrt|a_ii|) (mix.isd)
  int *blk_sys = nullptr, *blk_r0 = nullptr, *blk_r1 = nullptr, *sys_blk0 = nullptr;
  // SELL-64 slices of the systems, concatenated
  int nslice = 0;
  int64_t stot = 0;
  int *soff = nullptr, *scol = nullptr, *slice_row0 = nullptr, *blk_s0 = nullptr, *blk_s1 = nullptr;
  short *scol16 = nullptr;
  double *sval64 = nullptr;
  float *sval32 = nullptr;
  // vectors
  double *x = nullptr, *r = nullptr, *d = nullptr, *p = nullptr, *h = nullptr, *y = nullptr,
         *part = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *d32 = nullptr, *p32 = nullptr,
        *y32 = nullptr, *part32 = nullptr;
  StateArrays<double> st64;
  StateArrays<float> st32;
  BatchMixDev mix;
  int *active = nullptr;
  int *active_host = nullptr;  // pinned, 2 entries
  // work list of the persistent kernels and the per-chunk record of unfinished
  // pairs (hist[k] after chunk k, ctr chunks so far)
  int *wl = nullptr, *nwl = nullptr, *hist = nullptr, *ctr = nullptr;
  BatchTune tune64, tune32;
  // [precision]: the one-chunk graph and the whole-solve graph of sgraph_n chunks
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t sgraph[2] = {nullptr, nullptr};
  int sgraph_n[2] = {0, 0};
  int graph_chunk[2] = {0, 0};
  double graph_res[2] = {0.0, 0.0}, graph_zero[2] = {0.0, 0.0};
  int64_t learned[2] = {0, 0};  // chunks the last solve needed (0: none yet)
  BatchDev dev() const
  {
    BatchDev b;
    b.nsys = nsys;
    b.R = R;
    b.nblk = nblk;
    b.nrows = nrows;
    b.rowptr = rowptr;
    b.col = col;
    b.blk_sys = blk_sys;
    b.blk_r0 = blk_r0;
    b.blk_r1 = blk_r1;
    b.sys_blk0 = sys_blk0;
    b.soff = soff;
    b.scol = scol;
    b.scol16 = scol16;
    b.slice_row0 = slice_row0;
    b.blk_s0 = blk_s0;
    b.blk_s1 = blk_s1;
    b.wl = wl;
    b.nwl = nwl;
    return b;
  }
  template <typename T>
  BatchVec<T> vec() const;
};

template <>
BatchVec<double> ssf_batch::vec<double>() const
{
  BatchVec<double> v;
  v.val = val; v.diag = diag; v.x = x; v.r = r; v.d = d; v.p = p; v.h = h; v.y = y; v.part = part;
  v.sval = sval64;
  return v;
}
template <>
BatchVec<float> ssf_batch::vec<float>() const
{
  BatchVec<float> v;
  v.val = val32; v.x = x32; v.r = r32; v.d = d32; v.p = p32;  // no diagonal, h = r
  v.y = y32; v.part = part32;
  v.sval = sval32;
  return v;
}

static void batch_release(ssf_batch *b)
{
  for (int k = 0; k < 2; k++) {
    if (b->graph[k]) (void)hipGraphExecDestroy(b->graph[k]);
    if (b->sgraph[k]) (void)hipGraphExecDestroy(b->sgraph[k]);
  }
  bfree(b->rowptr); bfree(b->col); bfree(b->val); bfree(b->diag); bfree(b->val32); bfree(b->mix.isd);
  bfree(b->blk_sys); bfree(b->blk_r0); bfree(b->blk_r1); bfree(b->sys_blk0);
  bfree(b->soff); bfree(b->scol); bfree(b->scol16); bfree(b->slice_row0); bfree(b->blk_s0);
  bfree(b->blk_s1); bfree(b->sval64); bfree(b->sval32);
  bfree(b->x); bfree(b->r); bfree(b->d); bfree(b->p); bfree(b->h); bfree(b->y); bfree(b->part);
  bfree(b->x32); bfree(b->r32); bfree(b->d32); bfree(b->p32); bfree(b->y32);
  bfree(b->part32);
  b->st64.release();
  b->st32.release();
  BatchMixDev &m = b->mix;
  bfree(m.nory64); bfree(m.rr64); bfree(m.best); bfree(m.scale); bfree(m.delta2); bfree(m.bad);
  bfree(m.since_best); bfree(m.final); bfree(m.status); bfree(m.updates); bfree(m.sys_final);
  bfree(m.norm_a);
  bfree(b->active);
  bfree(b->wl); bfree(b->nwl); bfree(b->hist); bfree(b->ctr);
  if (b->active_host) (void)hipHostFree(b->active_host);
  b->active_host = nullptr;
}

extern "C" int ssf_batch_destroy(ssf_batch *b)
{
  SSF_ENTER();
  if (!b) return SSF_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_release(b);
  b->ctx->live_matrices--;
  delete b;
  return SSF_OK;
}

extern "C" int ssf_batch_create(ssf_ctx *ctx, int64_t nsys, ssf_matrix *const *mats, int32_t nrhs,
                                ssf_batch **out)
{
  SSF_ENTER();
  SSF_CHECK(ctx && mats && out && nsys > 0, SSF_EINVAL, "ssf_batch_create: bad argument");
  SSF_CHECK(nrhs == 1 || nrhs == 2 || nrhs == 4 || nrhs == 8, SSF_EINVAL,
            "ssf_batch_create: nrhs must be 1, 2, 4 or 8");
  int64_t nrows = 0, nnz = 0;
  for (int64_t s = 0; s < nsys; s++) {
    SSF_CHECK(mats[s] && mats[s]->ctx == ctx, SSF_EINVAL, "ssf_batch_create: matrix of another context");
    SSF_CHECK(mats[s]->values_set, SSF_ESTATE, "ssf_batch_create: matrix values were never set");
    SSF_CHECK(mats[s]->rowptr && mats[s]->col && mats[s]->val, SSF_EUNSUP,
              "ssf_batch_create: the matrix keeps no full CSR");
    SSF_CHECK(mats[s]->n > 0 && mats[s]->nslice > 0, SSF_EINVAL,
              "ssf_batch_create: every system needs at least one unknown");
    nrows += mats[s]->n;
    nnz += mats[s]->nnz;
  }
  int64_t stot_all = 0;
  for (int64_t s = 0; s < nsys; s++) stot_all += mats[s]->sell_tot;
  SSF_CHECK(nnz < (1LL << 31) && stot_all < (1LL << 31) && nrows * nrhs < (1LL << 40), SSF_EUNSUP,
            "ssf_batch_create: batch larger than 2^31 matrix entries");
  SSF_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  ssf_batch *b = new ssf_batch();
  b->ctx = ctx;
  ctx->live_matrices++;
  b->nsys = (int)nsys;
  b->R = nrhs;
  b->nrows = nrows;
  b->nnz = nnz;
  b->row0.assign(nsys + 1, 0);
  for (int64_t s = 0; s < nsys; s++) b->row0[s + 1] = b->row0[s] + mats[s]->n;
  // block table: each block covers consecutive 64-row slices of one system
  // (2048 / nrhs rows at most), so its rows are one slice group of the SELL
  // layout and its scalar partials belong to one system
  const int spb = (int)std::max<int64_t>(1, std::max<int64_t>(64, 2048 / nrhs) / 64);
  std::vector<int> bsys, br0, br1, bs0, bs1, sblk0(nsys + 1, 0), srow0;
  std::vector<int64_t> sbase(nsys + 1, 0);
  int gslice = 0;
  for (int64_t s = 0; s < nsys; s++) {
    sblk0[s] = (int)bsys.size();
    const int64_t n = mats[s]->n;
    const int ns = mats[s]->nslice;
    for (int k = 0; k < ns; k += spb) {
      bsys.push_back((int)s);
      bs0.push_back(gslice + k);
      bs1.push_back(gslice + std::min(ns, k + spb));
      br0.push_back((int)(b->row0[s] + 64LL * k));
      br1.push_back((int)(b->row0[s] + std::min<int64_t>(n, 64LL * (k + spb))));
    }
    for (int k = 0; k < ns; k++) srow0.push_back((int)(b->row0[s] + 64LL * k));
    gslice += ns;
    sbase[s + 1] = sbase[s] + mats[s]->sell_tot;
  }
  sblk0[nsys] = (int)bsys.size();
  b->nblk = (int)bsys.size();
  b->nslice = gslice;
  b->stot = sbase[nsys];
  bool all16 = true;
  for (int64_t s = 0; s < nsys; s++) all16 = all16 && mats[s]->sell_col16 != nullptr;
  const int64_t nq = nsys * nrhs, nv = nrows * nrhs;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(balloc(&b->rowptr, nrows + 1)); c(balloc(&b->col, nnz)); c(balloc(&b->val, nnz));
  c(balloc(&b->diag, nrows));
  c(balloc(&b->blk_sys, b->nblk)); c(balloc(&b->blk_r0, b->nblk)); c(balloc(&b->blk_r1, b->nblk));
  c(balloc(&b->sys_blk0, nsys + 1));
  c(balloc(&b->x, nv)); c(balloc(&b->r, nv)); c(balloc(&b->d, nv)); c(balloc(&b->p, nv));
  c(balloc(&b->h, nv)); c(balloc(&b->y, nv)); c(balloc(&b->part, 3 * (int64_t)b->nblk * nrhs));
  c(b->st64.alloc(nq, nsys));
  c(balloc(&b->active, 1));
  c(balloc(&b->wl, b->nblk)); c(balloc(&b->nwl, 1)); c(balloc(&b->hist, HIST_CAP));
  c(balloc(&b->ctr, 1));
  c(balloc(&b->soff, b->nslice + 1)); c(balloc(&b->slice_row0, b->nslice));
  c(balloc(&b->blk_s0, b->nblk)); c(balloc(&b->blk_s1, b->nblk)); c(balloc(&b->sval64, b->stot));
  if (all16) c(balloc(&b->scol16, b->stot));
  else c(balloc(&b->scol, b->stot));
  if (e == hipSuccess) c(hipHostMalloc((void **)&b->active_host, 2 * sizeof(int), hipHostMallocDefault));
  if (e == hipSuccess) {
    c(hipMemcpy(b->blk_sys, bsys.data(), bsys.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipMemcpy(b->blk_r0, br0.data(), br0.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipMemcpy(b->blk_r1, br1.data(), br1.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipMemcpy(b->sys_blk0, sblk0.data(), sblk0.size() * sizeof(int), hipMemcpyHostToDevice));
    // block-diagonal CSR: rowptr and columns shifted by the system offsets
    int64_t k0 = 0;
    for (int64_t s = 0; s < nsys && e == hipSuccess; s++) {
      const ssf_matrix *m = mats[s];
      c(offset_copy(m->n, m->rowptr, b->rowptr + b->row0[s], (int)k0, st));
      c(offset_copy(m->nnz, m->col, b->col + k0, (int)b->row0[s], st));
      c(hipMemcpyAsync(b->val + k0, m->val, m->nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
      c(hipMemcpyAsync(b->diag + b->row0[s], m->diag, m->n * sizeof(double), hipMemcpyDeviceToDevice, st));
      k0 += m->nnz;
    }
    const int last = (int)nnz;
    c(hipMemcpyAsync(b->rowptr + nrows, &last, sizeof(int), hipMemcpyHostToDevice, st));
    // SELL slices: offsets shifted by each system's base, columns as they are
    // (16-bit ones are relative to their slice) or shifted to global rows
    std::vector<int> soff_all(b->nslice + 1, 0), tmp;
    int g = 0;
    for (int64_t s = 0; s < nsys && e == hipSuccess; s++) {
      const ssf_matrix *m = mats[s];
      tmp.resize(m->nslice + 1);
      c(hipMemcpy(tmp.data(), m->sell_off, (m->nslice + 1) * sizeof(int), hipMemcpyDeviceToHost));
      for (int k = 0; k < m->nslice; k++) soff_all[g + k] = (int)(sbase[s] + tmp[k]);
      g += m->nslice;
      c(hipMemcpyAsync(b->sval64 + sbase[s], m->sell_v64, m->sell_tot * sizeof(double),
                       hipMemcpyDeviceToDevice, st));
      if (all16)
        c(hipMemcpyAsync(b->scol16 + sbase[s], m->sell_col16, m->sell_tot * sizeof(short),
                         hipMemcpyDeviceToDevice, st));
      else
        c(offset_copy(m->sell_tot, m->sell_col, b->scol + sbase[s], (int)b->row0[s], st));
    }
    soff_all[b->nslice] = (int)b->stot;
    c(hipMemcpy(b->soff, soff_all.data(), soff_all.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipMemcpy(b->slice_row0, srow0.data(), srow0.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipMemcpy(b->blk_s0, bs0.data(), bs0.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipMemcpy(b->blk_s1, bs1.data(), bs1.size() * sizeof(int), hipMemcpyHostToDevice));
    c(hipStreamSynchronize(st));
  }
  if (e != hipSuccess) {
    const std::string msg = std::string("ssf_batch_create: ") + hipGetErrorString(e);
    ssf_batch_destroy(b);
    set_error(msg);
    return e == hipErrorOutOfMemory ? SSF_ENOMEM : SSF_EHIP;
  }
  // SELL SpMM by default; the CSR alternative with 4 lanes per right-hand side.
  // The default does not depend on the batch's contents, so a system's
  // summation order (and result) is the same alone and among others; only
  // ssf_batch_autotune changes it.
  b->tune64.format = 2;
  b->tune64.lanes = std::min(64, 4 * nrhs);
  b->tune32 = b->tune64;
  *out = b;
  return SSF_OK;
}

static int ensure_mixed(ssf_batch *b)
{
  if (b->val32) return SSF_OK;
  const int64_t nq = (int64_t)b->nsys * b->R, nv = b->nrows * b->R;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(balloc(&b->val32, b->nnz)); c(balloc(&b->mix.isd, b->nrows));
  c(balloc(&b->x32, nv)); c(balloc(&b->r32, nv)); c(balloc(&b->d32, nv)); c(balloc(&b->p32, nv));
  c(balloc(&b->y32, nv));
  // (4 partials per block and right-hand side: the fused FP32 step's dots)
  c(balloc(&b->part32, 4 * (int64_t)b->nblk * b->R));
  c(b->st32.alloc(nq, b->nsys));
  BatchMixDev &m = b->mix;
  c(balloc(&m.nory64, nq)); c(balloc(&m.rr64, nq)); c(balloc(&m.best, nq)); c(balloc(&m.scale, nq));
  c(balloc(&m.delta2, nq)); c(balloc(&m.bad, nq)); c(balloc(&m.since_best, nq));
  c(balloc(&m.final, nq)); c(balloc(&m.status, nq)); c(balloc(&m.updates, nq));
  c(balloc(&m.sys_final, b->nsys));
  c(balloc(&m.norm_a, b->nsys));
  m.active = b->active;
  c(balloc(&b->sval32, b->stot));
  if (e == hipSuccess) c(inv_sqrt_diag((int)b->nrows, b->diag, m.isd, b->ctx->stream));
  if (e == hipSuccess)
    c(csr_scale32((int)b->nrows, b->rowptr, b->col, b->val, m.isd, b->val32, b->ctx->stream));
  if (e == hipSuccess)
    c(batch_sell_scale32(b->dev(), b->nslice, b->sval64, m.isd, b->sval32, b->ctx->stream));
  if (e == hipSuccess) c(batch_norm_inf(b->dev(), b->val, b->part, m.norm_a, b->ctx->stream));
  SSF_HIP(e);
  return SSF_OK;
}

// one chunk: `c` iterations, then the count of unfinished pairs into `active`
// (the mixed solve ends its chunk with the FP64 refresh, which counts them)
static hipError_t enqueue_chunk(ssf_batch *b, int mixed, int64_t c, const ssf_solve_params *p)
{
  hipStream_t s = b->ctx->stream;
  const BatchDev B = b->dev();
  hipError_t e = hipSuccess;
  for (int64_t q = 0; q < c && e == hipSuccess; q++) {
    if (mixed)
      e = batch_iteration<float>(b->tune32, B, b->vec<float>(), b->st32.s, 1.0f, 1e-30f, true, s);
    else
      e = batch_iteration<double>(b->tune64, B, b->vec<double>(), b->st64.s, p->res, p->zero,
                                  false, s);
  }
  if (e != hipSuccess) return e;
  if (mixed)
    e = batch_mixed_refresh(b->tune64, B, b->vec<double>(), b->vec<float>(), b->st32.s, b->mix,
                            p->res, s);
  else {
    e = zero_int(b->active, s);
    if (e == hipSuccess) e = batch_flags<double>(B, b->st64.s, b->active, s);
  }
  // the next chunk's work list: the systems not yet finished
  if (e == hipSuccess)
    e = batch_compact(B, mixed ? b->mix.sys_final : b->st64.s.sys_done, b->wl, b->nwl, b->active,
                      b->hist, b->ctr, HIST_CAP, s);
  return e;
}

// nch full chunks captured into one graph (kernel arguments include res and
// zero); the graph holds kernel nodes only
static int capture_chunks(ssf_batch *b, int mixed, int chunk, int64_t nch,
                          const ssf_solve_params *p, hipGraphExec_t *out)
{
  hipStream_t s = b->ctx->stream;
  SSF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int64_t k = 0; k < nch && e == hipSuccess; k++) e = enqueue_chunk(b, mixed, chunk, p);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  SSF_HIP(e);
  return SSF_OK;
}

// the one-chunk graph of the host loop, and the whole-solve graph of nch chunks
// (nch <= 1: none); both are rebuilt when the chunk, res or zero change
static int solve_graphs(ssf_batch *b, int mixed, int chunk, int64_t nch, const ssf_solve_params *p)
{
  const bool same = b->graph_chunk[mixed] == chunk && b->graph_res[mixed] == p->res &&
                    b->graph_zero[mixed] == p->zero;
  if (!same) {
    for (hipGraphExec_t *g : {&b->graph[mixed], &b->sgraph[mixed]})
      if (*g) (void)hipGraphExecDestroy(*g);
    b->graph[mixed] = b->sgraph[mixed] = nullptr;
    b->sgraph_n[mixed] = 0;
    b->learned[mixed] = 0;
    b->graph_chunk[mixed] = chunk;
    b->graph_res[mixed] = p->res;
    b->graph_zero[mixed] = p->zero;
  }
  if (!b->graph[mixed]) {
    int rc = capture_chunks(b, mixed, chunk, 1, p, &b->graph[mixed]);
    if (rc) return rc;
  }
  if (nch > 1 && b->sgraph_n[mixed] != nch) {
    if (b->sgraph[mixed]) (void)hipGraphExecDestroy(b->sgraph[mixed]);
    b->sgraph[mixed] = nullptr;
    b->sgraph_n[mixed] = 0;
    int rc = capture_chunks(b, mixed, chunk, nch, p, &b->sgraph[mixed]);
    if (rc) return rc;
    b->sgraph_n[mixed] = (int)nch;
  }
  return SSF_OK;
}

extern "C" int ssf_batch_solve_dev(ssf_batch *b, const ssf_solve_params *p, const double *rhs_dev,
                                   double *lhs_dev, ssf_solve_info *infos)
{
  SSF_ENTER();
  SSF_CHECK(b && p && rhs_dev && lhs_dev && infos, SSF_EINVAL, "ssf_batch_solve_dev: NULL argument");
  SSF_CHECK(p->linsolver == SSF_SOLVER_CG && p->prec == SSF_PREC_DIAG, SSF_EUNSUP,
            "ssf_batch_solve_dev: the batched solver is Jacobi-PCG (cg + diagprec)");
  SSF_HIP(hipSetDevice(b->ctx->device));
  const bool mixed = p->precision == SSF_PRECISION_MIXED;
  if (mixed) {
    int rc = ensure_mixed(b);
    if (rc) return rc;
  }
  hipStream_t s = b->ctx->stream;
  const int64_t nv = b->nrows * b->R, nq = (int64_t)b->nsys * b->R;
  const BatchDev B = b->dev();
  SSF_HIP(hipMemcpyAsync(b->y, rhs_dev, nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (p->iv) SSF_HIP(hipMemcpyAsync(b->x, lhs_dev, nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  else SSF_HIP(hipMemsetAsync(b->x, 0, nv * sizeof(double), s));
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(p->inner_ni > 0 ? p->inner_ni : 32,
                                                                 std::max<int64_t>(p->ni, 1)));
  SSF_CHECK(mixed || !p->iv, SSF_EUNSUP, "ssf_batch_solve_dev: the FP64 batched solve starts from x = 0");
  // every block works in the set-up; the chunk record starts empty
  SSF_HIP(zero_int(b->ctr, s));
  SSF_HIP(batch_compact(B, nullptr, b->wl, b->nwl, nullptr, nullptr, nullptr, 0, s));
  if (mixed) {
    SSF_HIP(hipMemsetAsync(b->x32, 0, nv * sizeof(float), s));
    const double delta = (p->inner_res > 0.0 && p->inner_res < 1.0) ? p->inner_res : 1e-4;
    SSF_HIP(batch_mixed_start(b->tune64, b->tune32, B, b->vec<double>(), b->vec<float>(), b->st32.s,
                              b->mix, p->res, delta, s));
    SSF_HIP(batch_compact(B, b->mix.sys_final, b->wl, b->nwl, nullptr, nullptr, nullptr, 0, s));
  } else {
    SSF_HIP(batch_init<double>(B, b->vec<double>(), b->st64.s, p->zero, s));
  }
  const int mk = mixed ? 1 : 0;
  int rc = solve_graphs(b, mk, chunk, 0, p);  // drops graphs and record of other parameters
  if (rc) return rc;
  // the whole solve in one launch: the chunks the last solve of these
  // parameters needed, plus one (never more than ni iterations)
  const int64_t nfull = p->ni / chunk;
  const int64_t G = b->learned[mk] > 0 ? std::min<int64_t>(b->learned[mk] + 1, nfull) : 0;
  rc = solve_graphs(b, mk, chunk, G, p);
  if (rc) return rc;
  int64_t it = 0;
  bool finished = false;
  if (G > 1) {
    SSF_HIP(hipGraphLaunch(b->sgraph[mk], s));
    SSF_HIP(hipMemcpyAsync(&b->active_host[0], b->active, sizeof(int), hipMemcpyDeviceToHost, s));
    SSF_HIP(hipStreamSynchronize(s));
    it = G * chunk;
    finished = b->active_host[0] == 0;
  }
  hipEvent_t ev[2];
  SSF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SSF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int k = 0;
  bool pending = false;
  hipError_t e = hipSuccess;
  while (!finished) {
    bool launched = false;
    if (it < p->ni) {
      // the last chunk is shortened so that exactly ni iterations run
      const int64_t c = std::min<int64_t>(chunk, p->ni - it);
      e = c == chunk ? hipGraphLaunch(b->graph[mk], s) : enqueue_chunk(b, mk, c, p);
      if (e != hipSuccess) break;
      it += c;
      (void)hipMemcpyAsync(&b->active_host[k & 1], b->active, sizeof(int), hipMemcpyDeviceToHost, s);
      (void)hipEventRecord(ev[k & 1], s);
      launched = true;
    }
    if (pending) {
      (void)hipEventSynchronize(ev[(k - 1) & 1]);
      if (b->active_host[(k - 1) & 1] == 0) break;
    }
    if (!launched) break;
    pending = true;
    k++;
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  SSF_HIP(e);
  {
    // chunks this solve needed: the first chunk after which no pair was running
    int nch = 0;
    SSF_HIP(hipMemcpy(&nch, b->ctr, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> hist(std::min(nch, HIST_CAP));
    if (!hist.empty())
      SSF_HIP(hipMemcpy(hist.data(), b->hist, hist.size() * sizeof(int), hipMemcpyDeviceToHost));
    int64_t need = nch;
    for (size_t c = 0; c < hist.size(); c++)
      if (hist[c] == 0) {
        need = (int64_t)c + 1;
        break;
      }
    b->learned[mk] = need;
  }
  SSF_HIP(hipMemcpyAsync(lhs_dev, b->x, nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  // per-pair results
  std::vector<int> status(nq), done(nq);
  std::vector<long long> iter(nq), ani(nq);
  if (mixed) {
    std::vector<int> fin(nq);
    std::vector<long long> upd(nq);
    std::vector<double> rr(nq), ny(nq);
    SSF_HIP(hipMemcpy(fin.data(), b->mix.final, nq * sizeof(int), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(status.data(), b->mix.status, nq * sizeof(int), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(upd.data(), b->mix.updates, nq * sizeof(long long), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(rr.data(), b->mix.rr64, nq * sizeof(double), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(ny.data(), b->mix.nory64, nq * sizeof(double), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(iter.data(), b->st32.s.iter, nq * sizeof(long long), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < nq; q++) {
      ssf_solve_info &in = infos[q];
      in.status = fin[q] ? status[q] : 0;
      in.iterations = iter[q];
      in.outer = upd[q];
      in.ani = iter[q];
      in.ares = ny[q] > 0 ? rr[q] / ny[q] : 0.0;
      in.rel_residual = ny[q] > 0 ? std::sqrt(rr[q] / ny[q]) : 0.0;
    }
  } else {
    std::vector<double> ares(nq);
    SSF_HIP(hipMemcpy(status.data(), b->st64.s.status, nq * sizeof(int), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(iter.data(), b->st64.s.iter, nq * sizeof(long long), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(ani.data(), b->st64.s.ani, nq * sizeof(long long), hipMemcpyDeviceToHost));
    SSF_HIP(hipMemcpy(ares.data(), b->st64.s.ares, nq * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < nq; q++) {
      ssf_solve_info &in = infos[q];
      in.status = status[q];
      in.outer = 0;
      in.iterations = status[q] == 3 ? 0 : (status[q] ? ani[q] + 1 : iter[q]);
      if (status[q]) {
        in.ani = ani[q];
        in.ares = ares[q];
      }
      in.rel_residual = -1.0;
    }
  }
  SSF_HIP(hipStreamSynchronize(s));
  return SSF_OK;
}

extern "C" int ssf_batch_solve(ssf_batch *b, const ssf_solve_params *p, const double *rhs,
                               double *lhs, ssf_solve_info *infos)
{
  SSF_ENTER();
  SSF_CHECK(b && p && rhs && lhs && infos, SSF_EINVAL, "ssf_batch_solve: NULL argument");
  SSF_HIP(hipSetDevice(b->ctx->device));
  const int64_t nv = b->nrows * b->R;
  double *y = nullptr, *x = nullptr;
  SSF_HIP(balloc(&y, nv));
  hipError_t e = balloc(&x, nv);
  if (e != hipSuccess) {
    bfree(y);
    SSF_HIP(e);
  }
  hipStream_t s = b->ctx->stream;
  int rc = SSF_OK;
  e = hipMemcpyAsync(y, rhs, nv * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && p->iv) e = hipMemcpyAsync(x, lhs, nv * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) rc = ssf_batch_solve_dev(b, p, y, x, infos);
  if (e == hipSuccess && rc == SSF_OK) {
    e = hipMemcpyAsync(lhs, x, nv * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  bfree(y);
  bfree(x);
  SSF_HIP(e);
  return rc;
}

extern "C" int ssf_batch_sizes(const ssf_batch *b, int64_t *nrows, int64_t *nnz, int64_t *row0)
{
  SSF_ENTER();
  SSF_CHECK(b, SSF_EINVAL, "ssf_batch_sizes: NULL batch");
  if (nrows) *nrows = b->nrows;
  if (nnz) *nnz = b->nnz;
  if (row0) std::memcpy(row0, b->row0.data(), (b->nsys + 1) * sizeof(int64_t));
  return SSF_OK;
}

// SpMM format of the batch: SELL-64 (format 2) or the CSR row-vector kernel with
// R..64 lanes per row (format 0); the autotuner times the candidates on this
// batch and keeps the fastest. tune_out (may be NULL) receives {format, lanes}.
extern "C" int ssf_batch_autotune(ssf_batch *b, int precision, int reps, int32_t *tune_out)
{
  SSF_ENTER();
  SSF_CHECK(b, SSF_EINVAL, "ssf_batch_autotune: NULL batch");
  SSF_HIP(hipSetDevice(b->ctx->device));
  if (precision) {
    int rc = ensure_mixed(b);
    if (rc) return rc;
  }
  if (reps <= 0) reps = 5;
  hipStream_t s = b->ctx->stream;
  const BatchDev B = b->dev();
  // time the products over every block
  SSF_HIP(batch_compact(B, nullptr, b->wl, b->nwl, nullptr, nullptr, nullptr, 0, s));
  hipEvent_t t0, t1;
  SSF_HIP(hipEventCreate(&t0));
  SSF_HIP(hipEventCreate(&t1));
  float best = 1e30f;
  BatchTune bt = precision ? b->tune32 : b->tune64;
  std::vector<BatchTune> cand;
  {
    BatchTune t;
    t.format = 2;
    t.lanes = bt.lanes;
    cand.push_back(t);
  }
  for (int l = b->R; l <= 64; l *= 2) {
    BatchTune t;
    t.format = 0;
    t.lanes = l;
    cand.push_back(t);
  }
  for (const BatchTune &t : cand) {
    auto launch = [&]() -> hipError_t {
      if (precision)
        return bspmm<float>(t, 0, B, b->val32, b->sval32, b->d32, b->p32, nullptr, nullptr, nullptr, s);
      return bspmm<double>(t, 0, B, b->val, b->sval64, b->d, b->p, nullptr, nullptr, nullptr, s);
    };
    SSF_HIP(launch());
    SSF_HIP(hipEventRecord(t0, s));
    for (int i = 0; i < reps; i++) SSF_HIP(launch());
    SSF_HIP(hipEventRecord(t1, s));
    SSF_HIP(hipEventSynchronize(t1));
    float ms = 0.f;
    SSF_HIP(hipEventElapsedTime(&ms, t0, t1));
    if (ms < best) {
      best = ms;
      bt = t;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (precision) b->tune32 = bt;
  else b->tune64 = bt;
  // a new tiling invalidates the captured chunk
  const int k = precision ? 1 : 0;
  if (b->graph[k]) (void)hipGraphExecDestroy(b->graph[k]);
  if (b->sgraph[k]) (void)hipGraphExecDestroy(b->sgraph[k]);
  b->graph[k] = b->sgraph[k] = nullptr;
  b->sgraph_n[k] = 0;
  b->graph_chunk[k] = 0;
  if (tune_out) {
    tune_out[0] = bt.format;
    tune_out[1] = bt.lanes;
  }
  return SSF_OK;
}
