// Incomplete LDL^T preconditioning of symcomprow (GEFEL/scr.cpp:958-1060) and the
// cg_prec solve with tprec = incomdec (10) that gmatrix::solve_system runs for it
// (GEFEL/gmatrix.cpp:1234-1240), on gfx950.
//
// The factorization and both triangular sweeps are row recurrences: row i needs
// the finished rows of its pattern (factor, forward sweep) or of its transposed
// pattern (backward sweep). Rows are grouped into dependency levels once per
// pattern on the host; each level is one launch, one thread per row. A row's
// arithmetic is the reference's, operation for operation and in the same order
// (no contraction), so the factor and the preconditioned vector equal the
// reference's bit for bit:
//   factor   L_ij = (a_ij - sum_k L_ik L_jk D_k) / D_j over common k < j in column
//            order; D_i = a_ii - sum_j L_ij L_ij D_j in row order. gmatrix fixes
//            the weight of the dropped fill to 0 (GEFEL/gmatrix.cpp:24), so only
//            the pattern's entries are formed.
//   forward  z_i = y_i - sum_j L_ij z_j (row order), w_i = z_i / D_i
//   backward h_k = w_k - L_ik h_i for i > k in descending i, the order in which
//            prec_ildl_scr's row-by-row scatter reaches entry k.
// The sweeps of one application are captured once into a hipGraph (kernel nodes
// only) per pair of vectors and replayed; every level kernel returns at once
// after the solve has stopped.
#include <hip/hip_ext.h>

#include <cmath>

#include "ssf_device.hpp"

namespace ssf {

struct IldlPlan {
  int n = 0;
  int64_t negm = 0;
  int *adr = nullptr, *ci = nullptr;      // lower pattern, int32
  int *frows = nullptr, *brows = nullptr; // rows grouped by forward / backward level
  std::vector<int> flev, blev;            // level offsets (host)
  int *tptr = nullptr, *trow = nullptr, *tslot = nullptr;  // per k: (i, slot) of L_ik, i desc.
  double *tval = nullptr;                 // L_ik in that order (gathered after each factor)
  double *inc = nullptr;                  // the factor (reference layout)
  double *z = nullptr, *w = nullptr;      // forward sweep, before / after D^-1
  uint64_t version = ~0ULL;               // values_version of the factor
  hipGraphExec_t graph = nullptr;         // one application, for (gy, gh, gdone)
  const double *gy = nullptr;
  double *gh = nullptr;
  const int *gdone = nullptr;
};

namespace {

// one thread per row of a level: the row of the factor (scr.cpp:990-1022, weight 0)
__global__ __launch_bounds__(TPB) void k_ildl_factor(int cnt, const int *__restrict__ rows,
                                                     const int *__restrict__ adr,
                                                     const int *__restrict__ ci,
                                                     const double *__restrict__ a,
                                                     double *__restrict__ inc)
{
#pragma clang fp contract(off)
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const int b = adr[i], uj = adr[i + 1] - 1;
  for (int ii = b; ii < uj; ii++) {
    const int o = ci[ii];
    // rows_mult_ldl_scr (scr.cpp:958-975): merge of rows i and o below column o
    double s = 0.0;
    int p = b, q = adr[o];
    int cp = ci[p], cq = ci[q];
    while (cq < o) {
      if (cp == cq) {
        s += inc[p] * inc[q] * inc[adr[cq + 1] - 1];
        p++;
        q++;
      } else if (cp < cq) {
        p++;
      } else {
        q++;
      }
      cp = ci[p];
      cq = ci[q];
    }
    inc[ii] = (a[ii] - s) / inc[adr[o + 1] - 1];
  }
  double s = 0.0;
  for (int j = b; j < uj; j++) s += inc[j] * inc[j] * inc[adr[ci[j] + 1] - 1];
  inc[uj] = a[uj] - s;
}

// a row's entries are loaded ILDL_U at a time (indices, then values and the
// gathered vector entries, all in flight together) and summed in order: a level
// costs a couple of memory round trips instead of one per entry
constexpr int ILDL_U = 16;

// z_i = y_i - sum_j L_ij z_j, w_i = z_i / D_i (scr.cpp:1038-1049)
__global__ __launch_bounds__(TPB) void k_ildl_fwd(int cnt, const int *__restrict__ rows,
                                                  const int *__restrict__ adr,
                                                  const int *__restrict__ ci,
                                                  const double *__restrict__ inc,
                                                  const double *__restrict__ y,
                                                  double *__restrict__ z, double *__restrict__ w,
                                                  const int *done)
{
#pragma clang fp contract(off)
  if (done && *done) return;
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const int b = adr[i], uj = adr[i + 1] - 1;
  double s = 0.0;
  for (int j0 = b; j0 < uj; j0 += ILDL_U) {
    int c[ILDL_U];
    double l[ILDL_U], v[ILDL_U];
#pragma unroll
    for (int u = 0; u < ILDL_U; u++) c[u] = j0 + u < uj ? ci[j0 + u] : i;
#pragma unroll
    for (int u = 0; u < ILDL_U; u++) {
      l[u] = j0 + u < uj ? inc[j0 + u] : 0.0;
      v[u] = j0 + u < uj ? z[c[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < ILDL_U; u++)
      if (j0 + u < uj) s += l[u] * v[u];
  }
  const double zi = y[i] - s;
  z[i] = zi;
  w[i] = zi / inc[uj];
}

// h_k = w_k - L_ik h_i over i > k in descending i (scr.cpp:1051-1058); tval holds
// L_ik in that order
__global__ __launch_bounds__(TPB) void k_ildl_bwd(int cnt, const int *__restrict__ rows,
                                                  const int *__restrict__ tptr,
                                                  const int *__restrict__ trow,
                                                  const double *__restrict__ tval,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ h, const int *done)
{
#pragma clang fp contract(off)
  if (done && *done) return;
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= cnt) return;
  const int k = rows[t];
  const int e0 = tptr[k], e1 = tptr[k + 1];
  double v = w[k];
  for (int j0 = e0; j0 < e1; j0 += ILDL_U) {
    int c[ILDL_U];
    double l[ILDL_U], x[ILDL_U];
#pragma unroll
    for (int u = 0; u < ILDL_U; u++) c[u] = j0 + u < e1 ? trow[j0 + u] : k;
#pragma unroll
    for (int u = 0; u < ILDL_U; u++) {
      l[u] = j0 + u < e1 ? tval[j0 + u] : 0.0;
      x[u] = j0 + u < e1 ? h[c[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < ILDL_U; u++)
      if (j0 + u < e1) v -= l[u] * x[u];
  }
  h[k] = v;
}

__global__ __launch_bounds__(TPB) void k_ildl_tval(int64_t m, const int *__restrict__ tslot,
                                                   const double *__restrict__ inc,
                                                   double *__restrict__ tval)
{
  const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (e < m) tval[e] = inc[tslot[e]];
}

unsigned grid_of(int cnt) { return (unsigned)((cnt + TPB - 1) / TPB); }

}  // namespace

void ildl_free(IldlPlan *P, hipStream_t s)
{
  if (!P) return;
  if (P->graph) (void)hipGraphExecDestroy(P->graph);
  dfree(P->adr, s); dfree(P->ci, s); dfree(P->frows, s); dfree(P->brows, s);
  dfree(P->tptr, s); dfree(P->trow, s); dfree(P->tslot, s); dfree(P->tval, s);
  dfree(P->inc, s); dfree(P->z, s); dfree(P->w, s);
  delete P;
}

// the levels of matrix m's pattern and the factor of its current values
static int ildl_prepare(ssf_matrix *m)
{
  hipStream_t s = m->ctx->stream;
  if (!m->ildl) {
    const int64_t n = m->n, negm = m->negm;
    std::vector<int64_t> adr(n + 1), ci(negm);
    {
      int rc = ssf_matrix_pattern(m, adr.data(), ci.data());
      if (rc) return rc;
    }
    IldlPlan *P = new IldlPlan();
    P->n = (int)n;
    P->negm = negm;
    // forward levels: 1 + the deepest row of the pattern below the diagonal
    std::vector<int> lev(n, 0), a32(n + 1), c32(negm);
    int nf = 0;
    for (int64_t i = 0; i < n; i++) {
      int l = 0;
      for (int64_t j = adr[i]; j < adr[i + 1] - 1; j++) l = std::max(l, lev[ci[j]] + 1);
      lev[i] = l;
      nf = std::max(nf, l + 1);
    }
    auto group = [&](const std::vector<int> &lv, int nl, std::vector<int> &off,
                     std::vector<int> &rows) {
      off.assign(nl + 1, 0);
      for (int64_t i = 0; i < n; i++) off[lv[i] + 1]++;
      for (int l = 0; l < nl; l++) off[l + 1] += off[l];
      std::vector<int> fill(off.begin(), off.end() - 1);
      rows.resize(n);
      for (int64_t i = 0; i < n; i++) rows[fill[lv[i]]++] = (int)i;  // ascending within a level
    };
    std::vector<int> frows, brows;
    group(lev, nf, P->flev, frows);
    // transposed pattern: for column k the entries (i, slot) with i > k, descending i
    std::vector<int> tptr(n + 1, 0), trow(negm), tslot(negm);
    for (int64_t i = 0; i < n; i++)
      for (int64_t j = adr[i]; j < adr[i + 1] - 1; j++) tptr[ci[j] + 1]++;
    for (int64_t k = 0; k < n; k++) tptr[k + 1] += tptr[k];
    {
      std::vector<int> fill(tptr.begin(), tptr.end() - 1);
      for (int64_t i = n - 1; i >= 0; i--)
        for (int64_t j = adr[i]; j < adr[i + 1] - 1; j++) {
          const int q = fill[ci[j]]++;
          trow[q] = (int)i;
          tslot[q] = (int)j;
        }
    }
    // backward levels: 1 + the deepest later row that reaches k
    int nbk = 0;
    for (int64_t k = n - 1; k >= 0; k--) {
      int l = 0;
      for (int e = tptr[k]; e < tptr[k + 1]; e++) l = std::max(l, lev[trow[e]] + 1);
      lev[k] = l;  // rows > k already hold their backward level
      nbk = std::max(nbk, l + 1);
    }
    group(lev, nbk, P->blev, brows);
    for (int64_t i = 0; i <= n; i++) a32[i] = (int)adr[i];
    for (int64_t k = 0; k < negm; k++) c32[k] = (int)ci[k];
    hipError_t e = hipSuccess;
    auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    auto up = [&](int **dst, const std::vector<int> &v) {
      c(dalloc(dst, v.size(), s));
      if (e == hipSuccess && !v.empty())
        c(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, s));
    };
    up(&P->adr, a32);
    up(&P->ci, c32);
    up(&P->frows, frows);
    up(&P->brows, brows);
    up(&P->tptr, tptr);
    up(&P->trow, trow);
    up(&P->tslot, tslot);
    c(dalloc(&P->inc, negm, s));
    c(dalloc(&P->tval, std::max<int64_t>(1, negm - n), s));
    c(dalloc(&P->z, n, s));
    c(dalloc(&P->w, n, s));
    if (e == hipSuccess) c(hipStreamSynchronize(s));  // the host vectors go out of scope
    if (e != hipSuccess) {
      ildl_free(P, s);
      set_error(std::string("ildl: ") + hipGetErrorString(e));
      return SSF_EHIP;
    }
    m->ildl = P;
  }
  IldlPlan *P = m->ildl;
  if (P->version != m->values_version) {
    SSF_CHECK(m->a_ref, SSF_ESTATE, "ildl: the matrix keeps no reference-layout values");
    for (size_t l = 0; l + 1 < P->flev.size(); l++) {
      const int cnt = P->flev[l + 1] - P->flev[l];
      hipLaunchKernelGGL(k_ildl_factor, dim3(grid_of(cnt)), dim3(TPB), 0, s, cnt,
                         P->frows + P->flev[l], P->adr, P->ci, m->a_ref, P->inc);
    }
    const int64_t nl = P->negm - P->n;  // entries below the diagonal
    if (nl > 0)
      hipLaunchKernelGGL(k_ildl_tval, dim3((unsigned)((nl + TPB - 1) / TPB)), dim3(TPB), 0, s, nl,
                         P->tslot, P->inc, P->tval);
    SSF_HIP(hipGetLastError());
    P->version = m->values_version;
  }
  return SSF_OK;
}

// h = (L D L^T)^-1 y, level by level; skipped once *done is set
static hipError_t ildl_sweeps(const IldlPlan *P, const double *y, double *h, const int *done,
                              hipStream_t s)
{
  for (size_t l = 0; l + 1 < P->flev.size(); l++) {
    const int cnt = P->flev[l + 1] - P->flev[l];
    hipLaunchKernelGGL(k_ildl_fwd, dim3(grid_of(cnt)), dim3(TPB), 0, s, cnt,
                       P->frows + P->flev[l], P->adr, P->ci, P->inc, y, P->z, P->w, done);
  }
  for (size_t l = 0; l + 1 < P->blev.size(); l++) {
    const int cnt = P->blev[l + 1] - P->blev[l];
    hipLaunchKernelGGL(k_ildl_bwd, dim3(grid_of(cnt)), dim3(TPB), 0, s, cnt,
                       P->brows + P->blev[l], P->tptr, P->trow, P->tval, P->w, h, done);
  }
  return hipGetLastError();
}

// one application through the plan's graph (captured for these pointers)
static hipError_t ildl_apply(IldlPlan *P, const double *y, double *h, const int *done,
                             hipStream_t s)
{
  if (!P->graph || P->gy != y || P->gh != h || P->gdone != done) {
    if (P->graph) (void)hipGraphExecDestroy(P->graph);
    P->graph = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    e = ildl_sweeps(P, y, h, done, s);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&P->graph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      P->graph = nullptr;
      return e;
    }
    P->gy = y;
    P->gh = h;
    P->gdone = done;
  }
  return hipGraphLaunch(P->graph, s);
}

int ildl_factor_copy(ssf_matrix *m, double *out)
{
  int rc = ildl_prepare(m);
  if (rc) return rc;
  SSF_HIP(hipMemcpyAsync(out, m->ildl->inc, m->negm * sizeof(double), hipMemcpyDeviceToHost,
                         m->ctx->stream));
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  return SSF_OK;
}

int ildl_apply_dev(ssf_matrix *m, const double *y, double *h)
{
  int rc = ildl_prepare(m);
  if (rc) return rc;
  SSF_HIP(ildl_sweeps(m->ildl, y, h, nullptr, m->ctx->stream));
  return SSF_OK;
}

int ildl_bench(ssf_matrix *m, int reps, int graph, int64_t *nfwd, int64_t *nbwd, double *usec)
{
  int rc = ildl_prepare(m);
  if (rc) return rc;
  IldlPlan *P = m->ildl;
  hipStream_t s = m->ctx->stream;
  *nfwd = (int64_t)P->flev.size() - 1;
  *nbwd = (int64_t)P->blev.size() - 1;
  double *y = nullptr, *h = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = dalloc(&y, P->n, s);
  if (e == hipSuccess) e = dalloc(&h, P->n, s);
  if (e == hipSuccess) e = hipMemsetAsync(y, 0, P->n * sizeof(double), s);
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  auto once = [&]() { return graph ? ildl_apply(P, y, h, nullptr, s) : ildl_sweeps(P, y, h, nullptr, s); };
  if (e == hipSuccess) e = once();  // captures the graph
  if (e == hipSuccess) e = hipEventRecord(a, s);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = once();
  if (e == hipSuccess) e = hipEventRecord(b, s);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  *usec = reps > 0 ? 1000.0 * ms / reps : 0.0;
  // the graph holds these pointers: drop it with them
  if (P->graph) (void)hipGraphExecDestroy(P->graph);
  P->graph = nullptr;
  P->gy = nullptr;
  P->gh = nullptr;
  dfree(y, s);
  dfree(h, s);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  SSF_HIP(e);
  return SSF_OK;
}

int solve_ildl(ssf_matrix *m, const ssf_solve_params *p, const double *y, double *x,
               ssf_solve_info *info)
{
  int rc = ildl_prepare(m);
  if (rc) return rc;
  IldlPlan *P = m->ildl;
  hipStream_t s = m->ctx->stream;
  Workspace &w = m->ws;
  ssf_ctx *ctx = m->ctx;
  const SpmvTune &t = m->tune64;
  const bool havex = p->iv != 0;
  if (!havex) SSF_HIP(hipMemsetAsync(x, 0, m->n * sizeof(double), s));
  const int *done = &w.st->done;
  auto prec = [&]() { return ildl_apply(P, w.r, w.h, done, s); };
  // the first application runs before the state exists: no done flag for it
  auto prec0 = [&]() { return ildl_sweeps(P, w.r, w.h, nullptr, s); };
  SSF_HIP(pcg_init_prec(t, havex, m->csr(), m->val, x, y, w.r, w.h, w.d, w.part, w.st, p->zero,
                        s, prec0));
  // a few iterations per host check: after the stop test, every queued
  // iteration still dispatches its level kernels
  constexpr int CHUNK = 2;
  hipEvent_t ev[2];
  SSF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SSF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int64_t it = 0;
  int k = 0;
  bool pending = false;
  hipError_t err = hipSuccess;
  while (true) {
    bool launched_now = false;
    if (it < p->ni) {
      const int64_t c = std::min<int64_t>(CHUNK, p->ni - it);
      for (int64_t q = 0; q < c && err == hipSuccess; q++) {
        const bool smp = ctx->prof && q == 0;
        err = pcg_iteration_prec(t, m->csr(), m->val, x, w.r, w.d, w.p, w.h, w.part, w.st,
                                 p->res, p->zero, s, smp ? ctx->prof_ev[k & 1][0] : nullptr,
                                 smp ? ctx->prof_ev[k & 1][1] : nullptr, prec);
      }
      if (err != hipSuccess) break;
      it += c;
      (void)hipMemcpyAsync(&w.st_host[k & 1], w.st, sizeof(PcgState<double>),
                           hipMemcpyDeviceToHost, s);
      (void)hipEventRecord(ev[k & 1], s);
      launched_now = true;
    }
    if (pending) {
      (void)hipEventSynchronize(ev[(k - 1) & 1]);
      if (ctx->prof) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->prof_ev[(k - 1) & 1][0],
                                ctx->prof_ev[(k - 1) & 1][1]) == hipSuccess) {
          ctx->prof_us[0] += 1000.0 * ms;
          ctx->prof_n[0] += 1;
        }
      }
      if (w.st_host[(k - 1) & 1].done) break;
    }
    if (!launched_now) break;
    pending = true;
    k++;
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  SSF_HIP(err);
  PcgState<double> st;
  SSF_HIP(hipMemcpy(&st, w.st, sizeof(st), hipMemcpyDeviceToHost));
  info->status = st.status;
  info->outer = 0;
  if (st.status == 3) {  // zero right-hand side: x = 0, as the other cg_prec forms
    SSF_HIP(hipMemsetAsync(x, 0, m->n * sizeof(double), s));
    info->ani = 0;
    info->ares = st.ares;
    info->iterations = 0;
  } else if (st.status == 1 || st.status == 2) {
    info->ani = st.ani;
    info->ares = st.ares;
    info->iterations = st.ani + 1;
  } else {
    info->iterations = it;  // ni spent: cg_prec leaves ani / ares as they were
  }
  return SSF_OK;
}

}  // namespace ssf
