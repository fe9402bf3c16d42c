// Served solves: a request loop that admits element-mesh solve requests under an
// HBM budget and solves them with continuous batching on one GPU.
//
//  submit (any host thread) -> pending queue (FIFO)
//  prep workers (host threads, one HIP stream each): pattern of the request
//    (symcomprow::initiate semantics, GEFEL/scr.cpp:556), wait until the request's
//    device bytes fit the budget, upload, element kernels + assembly by entry on
//    the device, right-hand side with the Dirichlet lift, solve vectors
//    -> ready queue
//  solve loop (one host thread, its own stream): ready requests join the resident
//    set as slots; every step runs `chunk` iterations of the cg_prec recurrence
//    for all slots in one launch sequence; finished slots are retired (solution
//    copied to the host, device memory and budget released) and their places
//    are taken by the next ready requests at the following step.
//    In the mixed form a step is `chunk` iterations of every slot's FP32
//    recurrence followed by every slot's FP64 refresh (reliable update), which
//    decides which slots are finished.
//
// A request's device data stays where it was allocated: the kernels reach each
// slot through a descriptor, so admission and retirement move no matrix data.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ssf_serve.hpp"

using namespace ssf;

namespace {

using clk = std::chrono::steady_clock;

struct Request {
  int64_t ticket = 0;
  // payload (copied at submit)
  int kind = 0;
  int64_t n = 0, nn = 0, ne = 0, nne = 0, ndofe = 0;
  std::vector<double> xyz, mat_a, mat_b, thick, pv, rhs;
  std::vector<int64_t> conn, cn;
  // result
  std::vector<double> x;
  ssf_solve_info info{};
  int rc = SSF_OK;
  std::string err;
  double t_submit = 0, t_admit = 0, t_ready = 0, t_done = 0;
  bool finished = false;
  // device
  ssf_matrix *m = nullptr;
  // FP64: y, x, r, d, p, h (6 n doubles); mixed: y, x, r (3 n doubles), then
  // x32, r32, d32, p32 (4 n floats), isd (n doubles); then ||A||_inf (1 double)
  double *vec = nullptr;
  int64_t bytes = 0;
  int sid = -1;
};

template <typename T>
hipError_t salloc(T **p, size_t count)
{
  *p = nullptr;
  return hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}
template <typename T>
void sfree(T *&p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

struct ssf_server {
  int device = 0;
  ssf_ctx *ctx = nullptr;  // solve stream
  std::vector<ssf_ctx *> prep_ctx;
  ssf_solve_params params{};
  int chunk = 32;
  int64_t budget = 0, used = 0, peak = 0;
  clk::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work;    // pending / ready / stop
  std::condition_variable cv_budget;  // budget released
  std::condition_variable cv_done;    // a request finished
  std::deque<Request *> pending, ready;
  std::map<int64_t, std::unique_ptr<Request>> all;
  int64_t next_ticket = 1;
  int preparing = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  std::thread loop;

  // solve-loop state (loop thread only)
  std::vector<Request *> active;
  std::vector<int> free_sids;
  int max_slots = 0;  // resident requests at most (size of the per-slot state)
  int cap_slots = 0, cap_blk = 0;
  SlotDesc *d_slots = nullptr;
  int *d_blk_slot = nullptr, *d_blk_s0 = nullptr, *d_blk_s1 = nullptr;
  double *d_part = nullptr;
  SlotState st{};
  int *d_active = nullptr;
  int *h_done = nullptr;  // pinned: done flags by sid
  int nblk = 0;
  // statistics
  int64_t completed = 0, steps = 0;
  double t_pattern = 0, t_create = 0, t_assemble = 0;  // summed over requests (prep threads)
  int64_t half_requests = 0;  // requests whose solve streams the half stencil form

  double now() const
  {
    return std::chrono::duration<double>(clk::now() - t0).count();
  }
};

namespace {

// device bytes of a resident request: SELL slices (offsets, colbytes per entry of
// columns: 0 in the stencil form, 2 or 4; FP64 values; FP32 values in the mixed
// form), diagonal, the solve vectors (six FP64 ones, or three FP64 and six FP32
// ones)
// half: values per precision of the half stencil form the solve streams instead of
// the slices' values (0: the slices' values are kept)
int64_t resident_bytes(int64_t n, int64_t tot, int colbytes, bool mixed, int64_t half = 0)
{
  const int64_t vals = half > 0 ? (8 + (mixed ? 4 : 0)) * half : (8 + (mixed ? 4 : 0)) * tot;
  return 4 * ((n + 63) / 64 + 1) + colbytes * tot + vals + 8 * n + 48 * n + (1 << 16);
}

// the half form a served request keeps: 13 lower slots from the row or 13 lower
// neighbour nodes of 3 unknowns (the instances of the served kernels)
bool served_half(const ssf_matrix *m)
{
  return m->half_nl == 13 && (m->half_b == 1 || m->half_b == 3) && m->half_v64 &&
         (m->half_b == 1 ? m->sten_d != nullptr : m->half_m_d != nullptr);
}

// and while it is laid out and assembled: the device pattern, the layout build's
// temporaries (row index, sort keys and values, twice), SELL columns of both
// widths, SELL positions and values of the reference layout, the assembly plan's
// build and the element-matrix buffer
int64_t prep_bytes(const Request &q, int64_t negm, int64_t tot, bool mixed)
{
  // the half form while both layouts are held: at most tot + 2^14 values per
  // precision (its padding reaches one stencil offset beyond the last slice)
  const int64_t half = (8 + (mixed ? 4 : 0)) * (tot + (1 << 14));
  return resident_bytes(q.n, tot, 4, mixed) + half + 2 * tot + 4 * (q.n + 1) + 4 * negm +
         28 * negm + 24 * q.n + 16 * negm + 8 * negm + asm_bytes(q.ne, (int)q.ndofe, negm) +
         8 * q.ne;
}

// s: the stream the release is ordered on (device memory is stream-ordered)
void release_device(Request *q, hipStream_t s)
{
  if (q->m) ssf_matrix_destroy(q->m);
  q->m = nullptr;
  if (q->vec) (void)hipFreeAsync(q->vec, s);
  q->vec = nullptr;
}

// prep worker: requests from the pending queue to the ready queue
void prep_main(ssf_server *S, int w)
{
  (void)hipSetDevice(S->device);
  ssf_ctx *ctx = S->prep_ctx[w];
  const bool mixed = S->params.precision == SSF_PRECISION_MIXED;
  while (true) {
    Request *q = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty(); });
      if (S->pending.empty()) return;  // stop
      q = S->pending.front();
      S->pending.pop_front();
      S->preparing++;
    }
    int rc = SSF_OK;
    std::string err;
    // FIFO over the budget: wait until `bytes` fit, then hold them
    auto reserve = [&](int64_t bytes) {
      std::unique_lock<std::mutex> lk(S->mu);
      if (bytes > S->budget) {
        err = "ssf_server: request larger than the whole HBM budget";
        return SSF_ENOMEM;
      }
      S->cv_budget.wait(lk, [&] { return S->used + bytes <= S->budget; });
      S->used += bytes;
      S->peak = std::max(S->peak, S->used);
      return SSF_OK;
    };
    auto release = [&](int64_t bytes) {
      std::unique_lock<std::mutex> lk(S->mu);
      S->used -= bytes;
      S->cv_budget.notify_all();
    };
    // 1. the symcomprow pattern, on the device (its sort holds transient bytes)
    const auto tp0 = clk::now();
    int *d_adr = nullptr, *d_ci = nullptr;
    int64_t negm = 0, tot = 0;
    {
      const int64_t pbytes = pattern_build_bytes(q->ne, (int)q->ndofe) + 4 * q->ne * q->ndofe;
      rc = reserve(pbytes);
      if (rc == SSF_OK) {
        std::vector<int> cn32(q->ne * q->ndofe);
        bool bad = false;
        for (int64_t k = 0; k < q->ne * q->ndofe; k++) {
          bad |= q->cn[k] > q->n;
          cn32[k] = (int)q->cn[k];
        }
        int *d_cn = nullptr;
        hipError_t e = bad ? hipErrorInvalidValue
                           : hipMallocAsync((void **)&d_cn, cn32.size() * sizeof(int), ctx->stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(d_cn, cn32.data(), cn32.size() * sizeof(int), hipMemcpyHostToDevice,
                             ctx->stream);
        if (bad) {
          rc = SSF_EINVAL;
          err = "ssf_server: code number larger than the number of unknowns";
        } else if (e != hipSuccess) {
          rc = SSF_EHIP;
          err = std::string("ssf_server: ") + hipGetErrorString(e);
        } else {
          rc = pattern_build_dev(q->n, q->ne, (int)q->ndofe, d_cn, ctx->stream, &d_adr, &d_ci,
                                 &negm);
          if (rc) err = ssf_last_error();
        }
        if (rc == SSF_OK && 2 * negm >= (1LL << 31)) {
          rc = SSF_EUNSUP;
          err = "ssf_server: more than 2^31 entries in the full matrix";
        }
        if (rc == SSF_OK) {
          e = sell_entries_dev(q->n, negm, d_adr, d_ci, ctx->stream, &tot);
          if (e != hipSuccess) {
            rc = SSF_EHIP;
            err = std::string("ssf_server: ") + hipGetErrorString(e);
          }
        }
        if (d_cn) (void)hipFreeAsync(d_cn, ctx->stream);
        (void)hipStreamSynchronize(ctx->stream);
        release(pbytes);
      }
    }
    const double dt_pattern = std::chrono::duration<double>(clk::now() - tp0).count();
    double dt_create = 0, dt_assemble = 0;
    // 2. the request's set-up bytes, then the matrix laid out from the pattern
    if (rc == SSF_OK) {
      q->bytes = prep_bytes(*q, negm, tot, mixed);
      rc = reserve(q->bytes);
      if (rc == SSF_OK) {
        std::unique_lock<std::mutex> lk(S->mu);
        q->t_admit = S->now();
      }
    }
    if (rc != SSF_OK) {
      if (d_adr) (void)hipFreeAsync(d_adr, ctx->stream);
      if (d_ci) (void)hipFreeAsync(d_ci, ctx->stream);
    }
    if (rc == SSF_OK) {
      const auto tc0 = clk::now();
      rc = matrix_create_sell_dev(ctx, q->n, d_adr, d_ci, negm, &q->m, true);
      dt_create = std::chrono::duration<double>(clk::now() - tc0).count();
      const auto ta0 = clk::now();
      if (rc == SSF_OK && q->m->nslice == 0) {
        set_error("ssf_server: the matrix has no SELL layout");
        rc = SSF_EUNSUP;
      }
      std::vector<double> lift(q->n, 0.0);
      if (rc == SSF_OK)
        rc = ssf_matrix_assemble(q->m, q->kind, q->nn, q->xyz.data(), q->ne, q->conn.data(),
                                 q->cn.data(), q->mat_a.data(),
                                 q->mat_b.empty() ? nullptr : q->mat_b.data(),
                                 q->thick.empty() ? nullptr : q->thick.data(),
                                 q->pv.empty() ? nullptr : q->pv.data(), (int64_t)q->pv.size(),
                                 q->pv.empty() ? nullptr : lift.data());
      if (rc == SSF_OK) {
        matrix_drop_assembly(q->m, true);
        matrix_keep_sell_only(q->m);
        dt_assemble = std::chrono::duration<double>(clk::now() - ta0).count();
        for (int64_t i = 0; i < q->n; i++) lift[i] += q->rhs[i];
        hipError_t e = hipMallocAsync((void **)&q->vec,
                                      (6 * (size_t)std::max<int64_t>(q->n, 1) + 1) * sizeof(double),
                                      ctx->stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(q->vec, lift.data(), q->n * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream);
        if (e == hipSuccess && mixed) {
          // the scaled FP32 operator float(S A S) of the SELL values and
          // S = 1/sqrt|diag|; ||A||_inf of the FP64 values, for the rounding
          // floor of the FP64 residual
          ssf_matrix *m = q->m;
          e = hipMallocAsync((void **)&m->sell_v32, std::max<int64_t>(m->sell_tot, 1) * sizeof(float),
                             ctx->stream);
          if (e == hipSuccess)
            e = mixed_operator(m->csr(), m->diag, nullptr, nullptr, m->sell_v64, m->sell_v32,
                               q->vec + 5 * q->n, ctx->stream);
          double *part = nullptr;
          if (e == hipSuccess) e = hipMallocAsync((void **)&part, 1024 * sizeof(double), ctx->stream);
          if (e == hipSuccess)
            e = sell_row_absmax((int)q->n, m->sell_off, m->sell_v64, part, 1024, q->vec + 6 * q->n,
                                ctx->stream);
          if (part) (void)hipFreeAsync(part, ctx->stream);
          if (e == hipSuccess && served_half(m) && matrix_half_build(m, 1) != SSF_OK)
            e = hipErrorUnknown;
        }
        // with the half form the solve streams it alone: the slices' values go
        int64_t half = 0;
        if (e == hipSuccess && served_half(q->m) && (!mixed || q->m->half_v32)) {
          half = matrix_half_entries(q->m);
          dfree(q->m->sell_v64, ctx->stream);
          dfree(q->m->sell_v32, ctx->stream);
        } else if (e == hipSuccess) {
          dfree(q->m->half_v64, ctx->stream);
          dfree(q->m->half_v32, ctx->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
          rc = SSF_EHIP;
          err = std::string("ssf_server: ") + hipGetErrorString(e);
        } else {  // only the solve structures are left: keep their bytes reserved
          const int64_t keep =
              resident_bytes(q->n, q->m->sell_tot,
                             !q->m->sten.empty() ? 0 : q->m->sell_col16 ? 2 : 4, mixed, half);
          std::unique_lock<std::mutex> lk(S->mu);
          if (keep < q->bytes) {
            S->used -= q->bytes - keep;
            q->bytes = keep;
            S->cv_budget.notify_all();
          }
        }
      } else {
        err = ssf_last_error();
      }
      // the payload is no longer needed
      std::vector<double>().swap(q->xyz);
      std::vector<int64_t>().swap(q->conn);
      std::vector<int64_t>().swap(q->cn);
      std::vector<double>().swap(q->mat_a);
      std::vector<double>().swap(q->mat_b);
      std::vector<double>().swap(q->thick);
      std::vector<double>().swap(q->pv);
      std::vector<double>().swap(q->rhs);
    }
    std::unique_lock<std::mutex> lk(S->mu);
    S->preparing--;
    S->t_pattern += dt_pattern;
    S->t_create += dt_create;
    S->t_assemble += dt_assemble;
    if (rc == SSF_OK && q->m && !q->m->sell_v64 && q->m->half_v64) S->half_requests++;
    if (rc != SSF_OK) {
      lk.unlock();
      release_device(q, ctx->stream);
      lk.lock();
      if (q->t_admit > 0) S->used -= q->bytes;
      q->rc = rc;
      q->err = err;
      q->finished = true;
      q->t_done = S->now();
      S->cv_budget.notify_all();
      S->cv_done.notify_all();
      S->cv_work.notify_all();
      continue;
    }
    q->t_ready = S->now();
    S->ready.push_back(q);
    S->cv_work.notify_all();
  }
}

hipError_t ensure_tables(ssf_server *S, int nslot, int nblk)
{
  hipError_t e = hipSuccess;
  if (nslot > S->cap_slots) {
    sfree(S->d_slots);
    int cap = std::max(64, S->cap_slots);
    while (cap < nslot) cap *= 2;
    e = salloc(&S->d_slots, cap);
    if (e != hipSuccess) return e;
    S->cap_slots = cap;
  }
  if (nblk > S->cap_blk) {
    sfree(S->d_blk_slot);
    sfree(S->d_blk_s0);
    sfree(S->d_blk_s1);
    sfree(S->d_part);
    int cap = std::max(1024, S->cap_blk);
    while (cap < nblk) cap *= 2;
    if ((e = salloc(&S->d_blk_slot, cap)) != hipSuccess) return e;
    if ((e = salloc(&S->d_blk_s0, cap)) != hipSuccess) return e;
    if ((e = salloc(&S->d_blk_s1, cap)) != hipSuccess) return e;
    if ((e = salloc(&S->d_part, 4 * (size_t)cap)) != hipSuccess) return e;
    S->cap_blk = cap;
  }
  return e;
}

// descriptors and block table of the resident set (fresh marks the newcomers)
hipError_t upload_tables(ssf_server *S, const std::vector<char> &fresh)
{
  std::vector<SlotDesc> desc(S->active.size());
  std::vector<int> bs, b0, b1;
  for (size_t k = 0; k < S->active.size(); k++) {
    Request *q = S->active[k];
    const ssf_matrix *m = q->m;
    SlotDesc &D = desc[k];
    D.soff = m->sell_off;
    D.scol = m->sell_col;
    D.scol16 = m->sell_col16;
    D.sten = m->sten.empty() ? nullptr : m->sten_d;
    D.sten_n = (int)m->sten.size();
    D.sten_b = m->sten_b;
    D.sval = m->sell_v64;
    D.half_v = m->sell_v64 ? nullptr : m->half_v64;  // the half form replaced the slices' values
    D.half_v32 = m->sell_v64 ? nullptr : m->half_v32;
    D.half_off = m->half_b == 1 ? m->sten_d : m->half_m_d;
    D.half_pad = m->half_pad;
    D.half_b = m->half_b;
    D.diag = m->diag;
    D.y = q->vec;
    D.x = q->vec + q->n;
    D.r = q->vec + 2 * q->n;
    D.d = q->vec + 3 * q->n;
    D.p = q->vec + 4 * q->n;
    D.h = q->vec + 5 * q->n;
    {  // the mixed form's FP32 vectors follow its three FP64 ones
      float *f = (float *)(q->vec + 3 * q->n);
      D.x32 = f;
      D.r32 = f + q->n;
      D.d32 = f + 2 * q->n;
      D.p32 = f + 3 * q->n;
      D.isd = q->vec + 5 * q->n;  // the two n-float slots after p32
      D.sval32 = m->sell_v32;
      D.norm_a = q->vec + 6 * q->n;
    }
    D.n = (int)q->n;
    D.sid = q->sid;
    D.fresh = fresh[k];
    D.blk0 = (int)bs.size();
    const int ns = m->nslice;
    // a request in the half form of 3-unknown nodes takes 12 row slices per
    // block: 4 node slices, one per wavefront
    const int spb = (D.half_v && m->half_b == 3) ? 12 : SERVE_SLICES_PER_BLOCK;
    for (int t = 0; t < ns; t += spb) {
      bs.push_back((int)k);
      b0.push_back(t);
      b1.push_back(std::min(ns, t + spb));
    }
    D.nblk = (int)bs.size() - D.blk0;
  }
  S->nblk = (int)bs.size();
  hipError_t e = ensure_tables(S, (int)desc.size(), S->nblk);
  if (e != hipSuccess) return e;
  hipStream_t s = S->ctx->stream;
  if (!desc.empty()) {
    e = hipMemcpyAsync(S->d_slots, desc.data(), desc.size() * sizeof(SlotDesc),
                       hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(S->d_blk_slot, bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(S->d_blk_s0, b0.data(), b0.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(S->d_blk_s1, b1.data(), b1.size() * sizeof(int), hipMemcpyHostToDevice, s);
    // the host vectors go out of scope: wait for the copies
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  return e;
}

ServeDev serve_view(ssf_server *S)
{
  ServeDev v;
  v.nslot = (int)S->active.size();
  v.nblk = S->nblk;
  v.slots = S->d_slots;
  v.blk_slot = S->d_blk_slot;
  v.blk_s0 = S->d_blk_s0;
  v.blk_s1 = S->d_blk_s1;
  v.part = S->d_part;
  return v;
}

void finish(ssf_server *S, Request *q, int rc, const std::string &err)
{
  std::unique_lock<std::mutex> lk(S->mu);
  q->rc = rc;
  q->err = err;
  q->finished = true;
  q->t_done = S->now();
  S->used -= q->bytes;
  S->completed++;
  S->cv_budget.notify_all();
  S->cv_done.notify_all();
}

void loop_main(ssf_server *S)
{
  (void)hipSetDevice(S->device);
  hipStream_t s = S->ctx->stream;
  const ssf_solve_params P = S->params;
  const bool mixed = P.precision == SSF_PRECISION_MIXED;
  const double delta = (P.inner_res > 0.0 && P.inner_res < 1.0) ? P.inner_res : 1e-4;
  // a slot is finished when its recurrence stopped (FP64) or its solve ended (mixed)
  int *fin_dev = mixed ? S->st.final : S->st.done;
  std::string fail;
  while (true) {
    std::vector<Request *> joined;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] {
        return (!S->ready.empty() && !S->free_sids.empty()) || !S->active.empty() ||
               (S->stop && S->pending.empty() && S->preparing == 0 && S->ready.empty());
      });
      if (S->active.empty() && S->ready.empty() && S->stop && S->pending.empty() &&
          S->preparing == 0)
        break;
      while (!S->ready.empty() && joined.size() < S->free_sids.size()) {
        joined.push_back(S->ready.front());
        S->ready.pop_front();
      }
    }
    hipError_t e = hipSuccess;
    if (!joined.empty()) {
      std::vector<char> fresh(S->active.size(), 0);
      for (Request *q : joined) {
        q->sid = S->free_sids.back();
        S->free_sids.pop_back();
        S->active.push_back(q);
        fresh.push_back(1);
      }
      e = upload_tables(S, fresh);
      if (e == hipSuccess)
        e = mixed ? serve_mixed_init(serve_view(S), S->st, P.res, delta, P.zero, s)
                  : serve_init(serve_view(S), S->st, P.zero, s);
    }
    const ServeDev V = serve_view(S);
    for (int k = 0; k < S->chunk && e == hipSuccess; k++)
      e = mixed ? serve_mixed_iteration(V, S->st, s)
                : serve_iteration(V, S->st, P.res, P.zero, P.ni, s);
    if (e == hipSuccess)
      e = mixed ? serve_mixed_refresh(V, S->st, P.res, P.ni, S->d_active, s)
                : serve_flags(V, S->st, S->d_active, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(S->h_done, fin_dev, S->max_slots * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    S->steps++;
    if (e != hipSuccess) {
      fail = std::string("ssf_server: ") + hipGetErrorString(e);
      // fail every resident request and stop iterating
      for (Request *q : S->active) {
        release_device(q, s);
        finish(S, q, SSF_EHIP, fail);
      }
      S->active.clear();
      continue;
    }
    // retire the finished slots
    std::vector<Request *> keep;
    bool changed = false;
    for (Request *q : S->active) {
      if (!S->h_done[q->sid]) {
        keep.push_back(q);
        continue;
      }
      changed = true;
      int status = 0;
      long long iter = 0, ani = 0;
      double ares = 0.0;
      q->x.resize(q->n);
      hipError_t e2 = hipMemcpy(&status, S->st.status + q->sid, sizeof(int), hipMemcpyDeviceToHost);
      if (e2 == hipSuccess) e2 = hipMemcpy(&iter, S->st.iter + q->sid, sizeof(long long), hipMemcpyDeviceToHost);
      if (e2 == hipSuccess) e2 = hipMemcpy(&ani, S->st.ani + q->sid, sizeof(long long), hipMemcpyDeviceToHost);
      if (e2 == hipSuccess) e2 = hipMemcpy(&ares, S->st.ares + q->sid, sizeof(double), hipMemcpyDeviceToHost);
      if (e2 == hipSuccess && q->n)
        e2 = hipMemcpy(q->x.data(), q->vec + q->n, q->n * sizeof(double), hipMemcpyDeviceToHost);
      if (mixed) {  // as the single-system mixed solve reports
        long long upd = 0;
        double rr = 0.0, yy = 0.0;
        if (e2 == hipSuccess) e2 = hipMemcpy(&status, S->st.mstatus + q->sid, sizeof(int), hipMemcpyDeviceToHost);
        if (e2 == hipSuccess) e2 = hipMemcpy(&upd, S->st.updates + q->sid, sizeof(long long), hipMemcpyDeviceToHost);
        if (e2 == hipSuccess) e2 = hipMemcpy(&rr, S->st.rr64 + q->sid, sizeof(double), hipMemcpyDeviceToHost);
        if (e2 == hipSuccess) e2 = hipMemcpy(&yy, S->st.nory64 + q->sid, sizeof(double), hipMemcpyDeviceToHost);
        q->info.status = status;
        q->info.outer = upd;
        q->info.iterations = status == 3 ? 0 : iter;
        q->info.ani = q->info.iterations;
        q->info.ares = yy > 0.0 ? rr / yy : 0.0;
        q->info.rel_residual = yy > 0.0 ? std::sqrt(rr / yy) : 0.0;
      } else {
        q->info.status = status;
        q->info.outer = 0;
        q->info.iterations = status == 3 ? 0 : (status == 1 || status == 2 ? ani + 1 : iter);
        q->info.ani = (status == 0) ? -1 : ani;
        q->info.ares = ares;
        q->info.rel_residual = -1.0;
      }
      S->free_sids.push_back(q->sid);
      release_device(q, s);
      finish(S, q, e2 == hipSuccess ? SSF_OK : SSF_EHIP,
             e2 == hipSuccess ? std::string() : std::string(hipGetErrorString(e2)));
    }
    if (changed) {
      S->active.swap(keep);
      std::vector<char> fresh(S->active.size(), 0);
      if (upload_tables(S, fresh) != hipSuccess) {
        for (Request *q : S->active) {
          release_device(q, s);
          finish(S, q, SSF_EHIP, "ssf_server: table upload failed");
        }
        S->active.clear();
      }
    }
  }
}

hipError_t alloc_state(ssf_server *S, int cap)
{
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  SlotState &t = S->st;
  c(salloc(&t.nory, cap)); c(salloc(&t.nom, 2 * (size_t)cap)); c(salloc(&t.ares, cap));
  c(salloc(&t.sc, 3 * (size_t)cap)); c(salloc(&t.sc4, 4 * (size_t)cap));
  c(salloc(&t.iter, cap)); c(salloc(&t.ani, cap)); c(salloc(&t.done, cap));
  c(salloc(&t.status, cap)); c(salloc(&t.pend, cap)); c(salloc(&t.pend_status, cap));
  c(salloc(&t.brk, cap));
  c(salloc(&t.nory64, cap)); c(salloc(&t.rr64, cap)); c(salloc(&t.best, cap));
  c(salloc(&t.scale, cap)); c(salloc(&t.delta2, cap)); c(salloc(&t.xx, cap));
  c(salloc(&t.bad, cap)); c(salloc(&t.since_best, cap)); c(salloc(&t.final, cap));
  c(salloc(&t.mstatus, cap)); c(salloc(&t.updates, cap));
  c(salloc(&S->d_active, 1));
  if (e == hipSuccess) c(hipHostMalloc((void **)&S->h_done, cap * sizeof(int), hipHostMallocDefault));
  // unused slots read as finished until a request takes them
  if (e == hipSuccess) c(hipMemset(t.done, 0xff, cap * sizeof(int)));
  if (e == hipSuccess) c(hipMemset(t.final, 0xff, cap * sizeof(int)));
  return e;
}

void free_all(ssf_server *S)
{
  SlotState &t = S->st;
  sfree(t.nory); sfree(t.nom); sfree(t.ares); sfree(t.sc); sfree(t.sc4);
  sfree(t.iter); sfree(t.ani); sfree(t.done);
  sfree(t.status); sfree(t.pend); sfree(t.pend_status); sfree(t.brk);
  sfree(t.nory64); sfree(t.rr64); sfree(t.best); sfree(t.scale); sfree(t.delta2); sfree(t.xx);
  sfree(t.bad); sfree(t.since_best); sfree(t.final); sfree(t.mstatus); sfree(t.updates);
  sfree(S->d_active);
  sfree(S->d_slots); sfree(S->d_blk_slot); sfree(S->d_blk_s0); sfree(S->d_blk_s1);
  sfree(S->d_part);
  if (S->h_done) (void)hipHostFree(S->h_done);
  S->h_done = nullptr;
}

}  // namespace

extern "C" int ssf_server_destroy(ssf_server *S)
{
  SSF_ENTER();
  if (!S) return SSF_OK;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_work.notify_all();
  }
  for (auto &t : S->workers)
    if (t.joinable()) t.join();
  if (S->loop.joinable()) S->loop.join();
  (void)hipSetDevice(S->device);
  for (auto &kv : S->all) release_device(kv.second.get(), S->ctx->stream);
  free_all(S);
  for (ssf_ctx *c : S->prep_ctx) ssf_ctx_destroy(c);
  if (S->ctx) ssf_ctx_destroy(S->ctx);
  delete S;
  return SSF_OK;
}

extern "C" int ssf_server_create(int device, int64_t budget_bytes, const ssf_solve_params *p,
                                 int32_t chunk, int32_t max_slots, int32_t prep_threads,
                                 ssf_server **out)
{
  SSF_ENTER();
  SSF_CHECK(p && out && budget_bytes > 0, SSF_EINVAL, "ssf_server_create: bad argument");
  SSF_CHECK(p->linsolver == SSF_SOLVER_CG && p->prec == SSF_PREC_DIAG, SSF_EUNSUP,
            "ssf_server_create: served solves are Jacobi-PCG (cg + diagprec)");
  SSF_CHECK((p->precision == SSF_PRECISION_FP64 || p->precision == SSF_PRECISION_MIXED) && !p->iv,
            SSF_EUNSUP, "ssf_server_create: served solves start from x = 0 (FP64 or mixed)");
  ssf_server *S = new ssf_server();
  S->device = device;
  S->params = *p;
  if (S->params.ni <= 0) S->params.ni = 100000;
  S->chunk = chunk > 0 ? chunk : 32;
  S->budget = budget_bytes;
  S->max_slots = max_slots > 0 ? max_slots : 1024;
  S->t0 = clk::now();
  int rc = ssf_ctx_create(device, &S->ctx);
  if (rc) {
    delete S;
    return rc;
  }
  const int nw = prep_threads > 0 ? std::min(prep_threads, 16) : 2;
  for (int w = 0; w < nw && rc == SSF_OK; w++) {
    ssf_ctx *c = nullptr;
    rc = ssf_ctx_create(device, &c);
    if (rc == SSF_OK) S->prep_ctx.push_back(c);
  }
  if (rc == SSF_OK && alloc_state(S, S->max_slots) != hipSuccess) {
    set_error("ssf_server_create: device allocation failed");
    rc = SSF_ENOMEM;
  }
  if (rc) {
    ssf_server_destroy(S);
    return rc;
  }
  for (int k = S->max_slots - 1; k >= 0; k--) S->free_sids.push_back(k);
  for (int w = 0; w < nw; w++) S->workers.emplace_back(prep_main, S, w);
  S->loop = std::thread(loop_main, S);
  *out = S;
  return SSF_OK;
}

extern "C" int ssf_server_submit(ssf_server *S, int kind, int64_t n, int64_t nn, const double *xyz,
                                 int64_t ne, const int64_t *conn, const int64_t *cn,
                                 const double *mat_a, const double *mat_b, const double *thick,
                                 const double *pv, int64_t npv, const double *rhs,
                                 int64_t *ticket)
{
  SSF_ENTER();
  SSF_CHECK(S && xyz && conn && cn && mat_a && rhs && ticket && n > 0 && ne > 0, SSF_EINVAL,
            "ssf_server_submit: bad argument");
  int nne, ndofn;
  if (kind == SSF_ELEM_HEX8_ELAST) { nne = 8; ndofn = 3; }
  else if (kind == SSF_ELEM_HEX8_HEAT) { nne = 8; ndofn = 1; }
  else if (kind == SSF_ELEM_QUAD4_PSTRESS) { nne = 4; ndofn = 2; }
  else { set_error("ssf_server_submit: unknown element kind"); return SSF_EINVAL; }
  SSF_CHECK(kind == SSF_ELEM_HEX8_HEAT || mat_b, SSF_EINVAL, "ssf_server_submit: nu missing");
  auto q = std::make_unique<Request>();
  q->kind = kind;
  q->n = n;
  q->nn = nn;
  q->ne = ne;
  q->nne = nne;
  q->ndofe = (int64_t)nne * ndofn;
  q->xyz.assign(xyz, xyz + nn * 3);
  q->conn.assign(conn, conn + ne * nne);
  q->cn.assign(cn, cn + ne * q->ndofe);
  q->mat_a.assign(mat_a, mat_a + ne);
  if (mat_b) q->mat_b.assign(mat_b, mat_b + ne);
  if (thick) q->thick.assign(thick, thick + ne);
  if (pv && npv > 0) q->pv.assign(pv, pv + npv);
  q->rhs.assign(rhs, rhs + n);
  std::unique_lock<std::mutex> lk(S->mu);
  SSF_CHECK(!S->stop, SSF_ESTATE, "ssf_server_submit: server is stopping");
  q->ticket = S->next_ticket++;
  q->t_submit = S->now();
  *ticket = q->ticket;
  S->pending.push_back(q.get());
  S->all[q->ticket] = std::move(q);
  S->cv_work.notify_all();
  return SSF_OK;
}

// Waits for a request, copies its solution (x[n], may be NULL) and result, and
// forgets it. times[4] (may be NULL): submit, admission (budget granted),
// ready (assembled), done, in seconds since the server started.
extern "C" int ssf_server_wait(ssf_server *S, int64_t ticket, double *x, ssf_solve_info *info,
                               double *times)
{
  SSF_ENTER();
  SSF_CHECK(S, SSF_EINVAL, "ssf_server_wait: NULL server");
  std::unique_ptr<Request> q;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    SSF_CHECK(it != S->all.end(), SSF_EINVAL, "ssf_server_wait: unknown ticket");
    Request *r = it->second.get();
    S->cv_done.wait(lk, [&] { return r->finished; });
    q = std::move(it->second);
    S->all.erase(it);
  }
  if (times) {
    times[0] = q->t_submit;
    times[1] = q->t_admit;
    times[2] = q->t_ready;
    times[3] = q->t_done;
  }
  if (q->rc != SSF_OK) {
    set_error(q->err);
    return q->rc;
  }
  if (x && q->n) std::memcpy(x, q->x.data(), q->n * sizeof(double));
  if (info) *info = q->info;
  return SSF_OK;
}

// stats[8]: completed requests, loop steps, peak reserved bytes, bytes now, the
// microseconds the prep threads spent on patterns, on matrix set-up (CSR, SELL,
// upload) and on device assembly, summed over requests, and the requests laid
// out in the half stencil form
extern "C" int ssf_server_stats(ssf_server *S, int64_t *stats)
{
  SSF_ENTER();
  SSF_CHECK(S && stats, SSF_EINVAL, "ssf_server_stats: bad argument");
  std::unique_lock<std::mutex> lk(S->mu);
  stats[0] = S->completed;
  stats[1] = S->steps;
  stats[2] = S->peak;
  stats[3] = S->used;
  stats[4] = (int64_t)(1e6 * S->t_pattern);
  stats[5] = (int64_t)(1e6 * S->t_create);
  stats[6] = (int64_t)(1e6 * S->t_assemble);
  stats[7] = S->half_requests;
  return SSF_OK;
}
