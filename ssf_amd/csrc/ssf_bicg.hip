// Biconjugate gradients of comprow::bicg (GEFEL/cr.cpp:1353-1419) on gfx950, for
// compressed_rows matrices.
//
// Both products of an iteration, p = A d and pp = A^T dd, come from ONE sweep
// over a SELL-64 layout of the union pattern of A and A^T whose slots carry
// a_ij and a_ji side by side: the column stream is read once for the two
// products and no transposed matrix or atomics are needed. Row i of A^T lists
// a_ji for ascending j, the order in which mtxv_cr (GEFEL/cr.cpp:843-858)
// accumulates c[i]; row i of A lists a_ij in the order of mxv_cr (:653-700).
// Slots of the union that A (or A^T) lacks, and the padding of a slice, hold 0.
//
// Iteration i (cr.cpp:1386-1411), three launches; the vectors of the recurrence
// are held in pairs (r, rr), (d, dd), (p, pp), so each row of a pair is one
// 16-byte access and the SpMV gathers d_j and dd_j with one load:
//   K1  p = A d, pp = A^T dd, partials dd.p               (k_dual_spmv, iter++)
//   K2  alpha = nom / dd.p; x += alpha d, r -= alpha p, rr -= alpha pp;
//       partials r.r, rr.rr, r.rr                          (k_bicg_update)
//   K3  normr = r.r, normrr = rr.rr; both / y.y < res -> stop; beta = r.rr / nom;
//       d = r + beta d, dd = rr + beta dd                  (k_bicg_dir)
// Every block re-reduces the partials in the same order, so all blocks take the
// same decision without inter-block synchronisation (as the PCG kernels do).
#include <hip/hip_ext.h>

#include <cmath>

#include "ssf_device.hpp"

namespace ssf {

struct BicgState {
  double normy;    // y.y
  double nom[2];   // r.rr of the current / next iteration (ping-pong by parity)
  double normr;    // r.r after the last completed iteration (r.r at the start)
  double ares;
  long long ani;
  long long iter;  // incremented by K1
  int status;      // 0 running / ni spent, 1 converged, 2 breakdown, 3 zero rhs
  int done;
};

struct DualSell {
  int n = 0, nslice = 0;
  int64_t tot = 0;
  int *off = nullptr, *col = nullptr;
  short *col16 = nullptr;              // col - 64 * slice, when every offset fits
  // stencil form of the union layout (then no column array): slot k of row r at
  // column (r - r % sten_b) + sten[k], sten padded with 0 to the slot count
  int *sten = nullptr;
  int sten_n = 0, sten_b = 1;
  int *src_a = nullptr, *src_t = nullptr;  // CSR slot of a_ij / a_ji, -1 if none
  double *va = nullptr, *vt = nullptr;
  uint64_t version = ~0ULL;            // values_version the values were taken at
  int nbs = 8, nbv = 8;                // blocks of K1 / of the vector kernels
  // iteration vectors in pairs, one 16-byte access per row: (r, rr), (d, dd),
  // (p, pp); the shadow quantities of BiCG sit beside their primal ones
  double2 *rv = nullptr, *dv = nullptr, *pv = nullptr;
  double *part = nullptr;
  BicgState *st = nullptr, *st_host = nullptr;  // st_host: 2 pinned copies
};

struct DualDev {
  int n, nslice;
  const int *off, *col;
  const short *col16;
  const double *va, *vt;
  const int *sten;
  int sten_b;
};

namespace {

// DUAL: (p, pp) = (A d, A^T dd) from the pairs dv = (d, dd), one 16-byte gather
// per entry, and the block partial of dd.p; otherwise q = A x (x, q plain).
// IDX: 0 int32 columns, 1 16-bit offsets from the slice's first row, 2 stencil
// form (the values alone; columns from the row's block and the slot's offset)
template <int IDX, bool DUAL>
__global__ __launch_bounds__(TPB) void k_dual_spmv(DualDev L, const double2 *__restrict__ dv,
                                                   double2 *__restrict__ pv,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ q,
                                                   double *__restrict__ part, const int *done,
                                                   long long *iter)
{
  typedef double tv __attribute__((ext_vector_type(SELL_V)));
  typedef int iv __attribute__((ext_vector_type(SELL_V)));
  typedef short sv __attribute__((ext_vector_type(SELL_V)));
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const int nb = gridDim.x;
  const int spb = (L.nslice + nb - 1) / nb;
  const int lb = xcd_block(blockIdx.x, nb);
  const int s0 = lb * spb, s1 = min(L.nslice, s0 + spb);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double acc = 0.0;
  for (int sl = s0 + wave; sl < s1; sl += TPB / 64) {
    const int off = L.off[sl];
    const int w = (L.off[sl + 1] - off) / 64;
    double sa = 0.0, st = 0.0;
    for (int q = 0; q < w; q += SELL_V) {
      const int k = off + q * 64 + lane * SELL_V;
      const tv va = ldnt((const tv *)(L.va + k));
      tv vt;
      if (DUAL) vt = ldnt((const tv *)(L.vt + k));
      int c[SELL_V];
      if (IDX == 2) {
        const int r = sl * 64 + lane;
        const int base = r - r % L.sten_b;
#pragma unroll
        for (int u = 0; u < SELL_V; u++) c[u] = min(max(base + L.sten[q + u], 0), L.n - 1);
      } else if (IDX == 1) {
        const sv cc = ldnt((const sv *)(L.col16 + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) c[u] = sl * 64 + cc[u];
      } else {
        const iv cc = ldnt((const iv *)(L.col + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) c[u] = cc[u];
      }
#pragma unroll
      for (int u = 0; u < SELL_V; u++) {
        if (DUAL) {
          const double2 g = dv[c[u]];
          sa += va[u] * g.x;
          st += vt[u] * g.y;
        } else {
          sa += va[u] * x[c[u]];
        }
      }
    }
    const int row = sl * 64 + lane;
    if (row < L.n) {
      if (DUAL) {
        pv[row] = make_double2(sa, st);
        acc += dv[row].y * sa;
      } else {
        q[row] = sa;
      }
    }
  }
  if (DUAL) {
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// r = y - q (q = A x) or y; rr = d = dd = r; partials y.y, r.rr
template <bool HAVEX>
__global__ __launch_bounds__(TPB) void k_bicg_init(int n, const double *__restrict__ y,
                                                   const double *__restrict__ q,
                                                   double2 *__restrict__ rv,
                                                   double2 *__restrict__ dv,
                                                   double *__restrict__ part)
{
  int r0, r1;
  chunk(n, r0, r1);
  double ayy = 0.0, arr = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double yi = y[i];
    const double ri = HAVEX ? yi - q[i] : yi;
    rv[i] = make_double2(ri, ri);
    dv[i] = make_double2(ri, ri);
    ayy += yi * yi;
    arr += ri * ri;
  }
  const double a = block_sum(ayy);
  const double c = block_sum(arr);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(TPB) void k_bicg_init2(const double *__restrict__ part, int nb,
                                                    BicgState *st, double zero)
{
  const double normy = reduce_partials(part, nb);
  const double nom = reduce_partials(part + nb, nb);
  if (threadIdx.x == 0) {
    st->normy = normy;
    st->nom[0] = nom;
    st->nom[1] = nom;
    st->normr = nom;
    st->ares = 0.0;
    st->ani = 0;
    st->iter = -1;
    st->status = 0;
    st->done = 0;
    if (normy < zero) {  // cr.cpp:1365-1369: return before x is touched
      st->status = 3;
      st->done = 1;
      st->ares = normy;
    }
  }
}

// iv = 0: x = 0 (cr.cpp:1371-1374), skipped when the right-hand side was zero
__global__ __launch_bounds__(TPB) void k_bicg_zero_x(int n, const BicgState *st,
                                                     double *__restrict__ x)
{
  if (st->status == 3) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

__global__ __launch_bounds__(TPB) void k_bicg_update(int n, BicgState *st,
                                                     const double *__restrict__ pden, int npden,
                                                     double *__restrict__ x,
                                                     double2 *__restrict__ rv,
                                                     const double2 *__restrict__ dv,
                                                     const double2 *__restrict__ pv,
                                                     double *__restrict__ prr)
{
  if (st->done) return;
  const int nb = gridDim.x;
  const double den = reduce_partials(pden, npden);
  const long long it = st->iter;
  const double alpha = st->nom[it & 1] / den;
  if (!isfinite(alpha)) {  // dd.p = 0: the recurrence breaks down
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = 2;
      st->ani = it;
      st->ares = st->normr;
      st->done = 1;
    }
    return;
  }
  int r0, r1;
  chunk(n, r0, r1);
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double2 di = dv[i], pi = pv[i];
    double2 ri = rv[i];
    x[i] += alpha * di.x;
    ri.x = ri.x - alpha * pi.x;
    ri.y = ri.y - alpha * pi.y;
    rv[i] = ri;
    const double si = ri.y;
    a1 += ri.x * ri.x;
    a2 += si * si;
    a3 += ri.x * si;
  }
  const double b1 = block_sum(a1);
  const double b2 = block_sum(a2);
  const double b3 = block_sum(a3);
  if (threadIdx.x == 0) {
    prr[blockIdx.x] = b1;
    prr[nb + blockIdx.x] = b2;
    prr[2 * nb + blockIdx.x] = b3;
  }
}

__global__ __launch_bounds__(TPB) void k_bicg_dir(int n, BicgState *st,
                                                  const double *__restrict__ prr, double res,
                                                  double2 *__restrict__ dv,
                                                  const double2 *__restrict__ rv)
{
  if (st->done) return;
  const int nb = gridDim.x;
  const double normr = reduce_partials(prr, nb);
  const double normrr = reduce_partials(prr + nb, nb);
  const double nomn = reduce_partials(prr + 2 * nb, nb);
  const long long it = st->iter;
  const double normy = st->normy;
  if (normr / normy < res && normrr / normy < res) {  // cr.cpp:1400
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->ani = it;
      st->ares = normr;
      st->normr = normr;
      st->status = 1;
      st->done = 1;
    }
    return;
  }
  const double beta = nomn / st->nom[it & 1];
  if (!isfinite(beta)) {  // r.rr of the previous iteration was 0
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->ani = it;
      st->ares = normr;
      st->normr = normr;
      st->status = 2;
      st->done = 1;
    }
    return;
  }
  int r0, r1;
  chunk(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double2 ri = rv[i];
    double2 di = dv[i];
    di.x = ri.x + beta * di.x;
    di.y = ri.y + beta * di.y;
    dv[i] = di;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->nom[(it + 1) & 1] = nomn;
    st->normr = normr;
  }
}

int vec_blocks_for(int n)
{
  int nb = std::min(1024, (n + TPB * 4 - 1) / (TPB * 4));
  nb = std::max(8, nb);
  return (nb + 7) / 8 * 8;
}

// union pattern of A (adr, ci: ascending columns) and A^T: per row, the merged
// ascending column list with the CSR slots of a_ij and a_ji (-1 if absent);
// false if it holds 2^31 entries or more
bool union_pattern(int64_t n, const std::vector<int64_t> &adr, const std::vector<int64_t> &ci,
                   std::vector<int> &urow, std::vector<int> &ucol, std::vector<int> &ua,
                   std::vector<int> &ut)
{
  // transpose: column j lists (row i, slot k) for ascending i
  std::vector<int> tadr(n + 1, 0), trow(adr[n]), tslot(adr[n]);
  for (int64_t k = 0; k < adr[n]; k++) tadr[ci[k] + 1]++;
  for (int64_t j = 0; j < n; j++) tadr[j + 1] += tadr[j];
  {
    std::vector<int> fill(tadr.begin(), tadr.end() - 1);
    for (int64_t i = 0; i < n; i++)
      for (int64_t k = adr[i]; k < adr[i + 1]; k++) {
        const int q = fill[ci[k]]++;
        trow[q] = (int)i;
        tslot[q] = (int)k;
      }
  }
  std::vector<int64_t> cnt(n + 1, 0);
  auto merge = [&](int64_t i, int *cols, int *sa, int *st) -> int {
    int64_t a = adr[i], ae = adr[i + 1], b = tadr[i], be = tadr[i + 1];
    int cnt = 0;
    while (a < ae || b < be) {
      const int64_t ca = a < ae ? ci[a] : INT64_MAX, cb = b < be ? trow[b] : INT64_MAX;
      const int64_t c = std::min(ca, cb);
      if (cols) {
        cols[cnt] = (int)c;
        sa[cnt] = ca == c ? (int)a : -1;
        st[cnt] = cb == c ? tslot[b] : -1;
      }
      if (ca == c) a++;
      if (cb == c) b++;
      cnt++;
    }
    return cnt;
  };
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) cnt[i + 1] = merge(i, nullptr, nullptr, nullptr);
  });
  for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  if (cnt[n] >= (1LL << 31)) return false;
  urow.assign(cnt.begin(), cnt.end());
  ucol.resize(urow[n]);
  ua.resize(urow[n]);
  ut.resize(urow[n]);
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++)
      merge(i, ucol.data() + urow[i], ua.data() + urow[i], ut.data() + urow[i]);
  });
  return true;
}

}  // namespace

void dual_free(DualSell *L, hipStream_t s)
{
  if (!L) return;
  dfree(L->off, s); dfree(L->col, s); dfree(L->col16, s); dfree(L->sten, s);
  dfree(L->src_a, s); dfree(L->src_t, s); dfree(L->va, s); dfree(L->vt, s);
  dfree(L->rv, s); dfree(L->dv, s); dfree(L->pv, s); dfree(L->part, s); dfree(L->st, s);
  if (L->st_host) (void)hipHostFree(L->st_host);
  delete L;
}

// the union layout, its values and the iteration vectors of matrix m (comprow)
static int dual_prepare(ssf_matrix *m)
{
  hipStream_t s = m->ctx->stream;
  const int64_t n = m->n;
  if (!m->dual) {
    std::vector<int64_t> adr(n + 1), ci(m->negm);
    {
      int rc = ssf_matrix_pattern(m, adr.data(), ci.data());
      if (rc) return rc;
    }
    std::vector<int> urow, ucol, ua, ut;
    std::vector<int> off, scol, src, sten;
    std::vector<short> scol16;
    int sten_b = 1;
    SSF_CHECK(union_pattern(n, adr, ci, urow, ucol, ua, ut) &&
                  build_sell(n, urow, ucol, off, scol, src, scol16, &sten, &sten_b),
              SSF_EUNSUP, "bicg: 2^31 entries or more in the union of A and A^T");
    std::vector<int> sten_pad(sten);
    sten_pad.resize((sten.size() + SELL_V - 1) / SELL_V * SELL_V, 0);
    const int64_t tot = off.back();
    std::vector<int> sa(tot), st(tot);
    parallel_for(tot, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; k++) {
        sa[k] = src[k] >= 0 ? ua[src[k]] : -1;
        st[k] = src[k] >= 0 ? ut[src[k]] : -1;
      }
    });
    DualSell *L = new DualSell();
    L->n = (int)n;
    L->nslice = (int)off.size() - 1;
    L->tot = tot;
    L->nbs = (int)std::max<int64_t>(8, std::min<int64_t>(2048, (L->nslice + 3) / 4));
    L->nbs = (L->nbs + 7) / 8 * 8;
    L->nbv = vec_blocks_for((int)n);
    hipError_t e = hipSuccess;
    auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    auto up = [&](auto **dst, const auto &v) {
      c(dalloc(dst, v.size(), s));
      if (e == hipSuccess && !v.empty())
        c(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, s));
    };
    up(&L->off, off);
    if (!sten.empty()) {  // stencil form: the values alone
      up(&L->sten, sten_pad);
      L->sten_n = (int)sten.size();
      L->sten_b = sten_b;
    } else if (!scol16.empty()) {
      up(&L->col16, scol16);
    } else {
      up(&L->col, scol);
    }
    up(&L->src_a, sa);
    up(&L->src_t, st);
    c(dalloc(&L->va, tot, s));
    c(dalloc(&L->vt, tot, s));
    for (double2 **v : {&L->rv, &L->dv, &L->pv}) c(dalloc(v, n, s));
    c(dalloc(&L->part, (size_t)L->nbs + 3 * (size_t)L->nbv, s));
    c(dalloc(&L->st, 1, s));
    if (e == hipSuccess) c(hipHostMalloc((void **)&L->st_host, 2 * sizeof(BicgState)));
    if (e == hipSuccess) c(hipStreamSynchronize(s));  // the host vectors go out of scope
    if (e != hipSuccess) {
      dual_free(L, s);
      set_error(std::string("bicg: ") + hipGetErrorString(e));
      return SSF_EHIP;
    }
    m->dual = L;
  }
  DualSell *L = m->dual;
  if (L->version != m->values_version) {
    SSF_CHECK(m->val, SSF_ESTATE, "bicg: the matrix keeps no CSR values");
    SSF_HIP(sell_gather(L->tot, L->src_a, m->val, L->va, s));
    SSF_HIP(sell_gather(L->tot, L->src_t, m->val, L->vt, s));
    L->version = m->values_version;
  }
  return SSF_OK;
}

// DUAL: (p, pp) into L->pv from L->dv; otherwise q = A x
template <bool DUAL>
static hipError_t dual_launch(const DualSell *L, const double *x, double *q, const int *done,
                              long long *iter, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
  const DualDev dd{L->n, L->nslice, L->off, L->col, L->col16, L->va, L->vt, L->sten, L->sten_b};
  dim3 grid(L->nbs), block(TPB);
#define SSF_DUAL(IDX)                                                                          \
  do {                                                                                         \
    if (ev0) hipExtLaunchKernelGGL((k_dual_spmv<IDX, DUAL>), grid, block, 0, s, ev0, ev1, 0, dd, \
                                   L->dv, L->pv, x, q, L->part, done, iter);                   \
    else hipLaunchKernelGGL((k_dual_spmv<IDX, DUAL>), grid, block, 0, s, dd, L->dv, L->pv, x, q, \
                            L->part, done, iter);                                              \
  } while (0)
  if (L->sten) SSF_DUAL(2);
  else if (L->col16) SSF_DUAL(1);
  else SSF_DUAL(0);
#undef SSF_DUAL
  return hipGetLastError();
}

static hipError_t bicg_iteration(const DualSell *L, double *x, double res, hipStream_t s,
                                 hipEvent_t ev0, hipEvent_t ev1)
{
  hipError_t e = dual_launch<true>(L, nullptr, nullptr, &L->st->done, &L->st->iter, s, ev0, ev1);
  if (e != hipSuccess) return e;
  double *prr = L->part + L->nbs;
  hipLaunchKernelGGL(k_bicg_update, dim3(L->nbv), dim3(TPB), 0, s, L->n, L->st, L->part, L->nbs,
                     x, L->rv, L->dv, L->pv, prr);
  hipLaunchKernelGGL(k_bicg_dir, dim3(L->nbv), dim3(TPB), 0, s, L->n, L->st, prr, res, L->dv,
                     L->rv);
  return hipGetLastError();
}

int solve_bicg(ssf_matrix *m, const ssf_solve_params *p, const double *y, double *x,
               ssf_solve_info *info)
{
  int rc = dual_prepare(m);
  if (rc) return rc;
  DualSell *L = m->dual;
  hipStream_t s = m->ctx->stream;
  ssf_ctx *ctx = m->ctx;
  const int n = L->n;
  const bool havex = p->iv != 0;
  // q = A x goes to the (p, pp) pairs' storage, read as n plain doubles
  double *q = reinterpret_cast<double *>(L->pv);
  if (havex) SSF_HIP(dual_launch<false>(L, x, q, nullptr, nullptr, s, nullptr, nullptr));
  if (havex)
    hipLaunchKernelGGL((k_bicg_init<true>), dim3(L->nbv), dim3(TPB), 0, s, n, y, q, L->rv, L->dv,
                       L->part);
  else
    hipLaunchKernelGGL((k_bicg_init<false>), dim3(L->nbv), dim3(TPB), 0, s, n, y, q, L->rv, L->dv,
                       L->part);
  hipLaunchKernelGGL(k_bicg_init2, dim3(1), dim3(TPB), 0, s, L->part, L->nbv, L->st, p->zero);
  if (!havex)
    hipLaunchKernelGGL(k_bicg_zero_x, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, s, n, L->st, x);
  SSF_HIP(hipGetLastError());
  // chunks of iterations between host checks; chunk k is enqueued before the
  // host waits for chunk k-1's state copy
  constexpr int CHUNK = 16;
  hipEvent_t ev[2];
  SSF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SSF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int64_t it = 0;
  int k = 0;
  bool pending = false;
  bool sample_valid[2] = {true, false};
  hipError_t err = hipSuccess;
  while (true) {
    bool launched_now = false;
    if (it < p->ni) {
      const int64_t c = std::min<int64_t>(CHUNK, p->ni - it);
      const int64_t qs = c > 1 ? 1 : 0;  // sampled: the chunk's second dual product
      for (int64_t q = 0; q < c && err == hipSuccess; q++) {
        const bool smp = ctx->prof && q == qs;
        err = bicg_iteration(L, x, p->res, s, smp ? ctx->prof_ev[k & 1][0] : nullptr,
                             smp ? ctx->prof_ev[k & 1][1] : nullptr);
      }
      if (err != hipSuccess) break;
      it += c;
      (void)hipMemcpyAsync(&L->st_host[k & 1], L->st, sizeof(BicgState), hipMemcpyDeviceToHost, s);
      (void)hipEventRecord(ev[k & 1], s);
      launched_now = true;
    }
    if (pending) {
      (void)hipEventSynchronize(ev[(k - 1) & 1]);
      if (ctx->prof && sample_valid[(k - 1) & 1]) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->prof_ev[(k - 1) & 1][0], ctx->prof_ev[(k - 1) & 1][1]) ==
            hipSuccess) {
          ctx->prof_us[0] += 1000.0 * ms;
          ctx->prof_n[0] += 1;
        }
      }
      sample_valid[k & 1] = !L->st_host[(k - 1) & 1].done;
      if (L->st_host[(k - 1) & 1].done) break;
    }
    if (!launched_now) break;
    pending = true;
    k++;
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  SSF_HIP(err);
  BicgState st;
  SSF_HIP(hipMemcpy(&st, L->st, sizeof(BicgState), hipMemcpyDeviceToHost));
  info->status = st.status;
  info->outer = 0;
  if (st.status == 3) {
    info->ani = 0;
    info->ares = st.ares;
    info->iterations = 0;
  } else if (st.status == 1 || st.status == 2) {
    info->ani = st.ani;
    info->ares = st.ares;
    info->iterations = st.ani + 1;
  } else {  // ni spent: ani = ni, ares = the last r.r (cr.cpp:1415)
    info->ani = p->ni;
    info->ares = st.normr;
    info->iterations = it;
  }
  return SSF_OK;
}

}  // namespace ssf
