// The allocation-free FP64 solve: ssf_solve_dev's cg recurrences with every
// working array (residual, direction, product, preconditioned residual,
// reduction partials, recurrence scalars) in one file-scope device array, for
// callers that cannot allocate device memory. The kernels and the tiling are the
// ones ssf_solve_dev launches, so the results are the same bit for bit.
#include <mutex>

#include "ssf_kernels.hpp"

namespace {

constexpr int64_t NA_MAX_N = SSF_NOALLOC_MAX_N;  // unknowns at most
constexpr int64_t NA_NB = 8192;                  // largest SpMV block count of a tiling
constexpr int64_t NA_PART = 4 * NA_NB;           // reduction partials (doubles)
constexpr int64_t NA_STATE = 16;                 // the recurrence scalars (doubles)
constexpr int64_t NA_DOUBLES = 4 * NA_MAX_N + NA_PART + NA_STATE;
static_assert(sizeof(ssf::PcgState<double>) <= NA_STATE * sizeof(double),
              "recurrence state does not fit its scratch slot");
static_assert((NA_MAX_N * sizeof(double)) % 256 == 0, "vector slots keep 256-byte alignment");

}  // namespace

// Held by every GPU the library is loaded on until the process ends. In host
// code this name is a stand-in, not a device address: kernels and copies take the
// per-device address from hipGetSymbolAddress (scratch_base below).
static __device__ __align__(256) double scratch[NA_DOUBLES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[NA_MAX_DEVICES] = {};

// the device address of `scratch` on `device`, looked up once per device with
// that device current
int scratch_base(int device, double **out)
{
  SSF_CHECK(device >= 0 && device < NA_MAX_DEVICES, SSF_EINVAL,
            "ssf_solve_dev_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ssf::set_error(std::string("ssf_solve_dev_noalloc: hipGetSymbolAddress of the scratch array "
                                 "failed: ") + hipGetErrorString(e));
      return SSF_EHIP;
    }
    g_scratch[device] = (double *)p;
  }
  *out = g_scratch[device];
  return SSF_OK;
}

constexpr int NA_CHUNK = 16;  // iterations enqueued between two state reads

}  // namespace

using namespace ssf;

extern "C" int ssf_solve_dev_noalloc(ssf_matrix *m, const ssf_solve_params *p,
                                     const double *rhs_dev, double *lhs_dev,
                                     ssf_solve_info *info)
{
  SSF_ENTER();
  SSF_CHECK(m && p && info && rhs_dev && lhs_dev, SSF_EINVAL,
            "ssf_solve_dev_noalloc: NULL argument");
  SSF_CHECK(m->n <= NA_MAX_N, SSF_EINVAL,
            "ssf_solve_dev_noalloc: more unknowns than the scratch array holds "
            "(SSF_NOALLOC_MAX_N = 2500000)");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_solve_dev_noalloc: matrix values were never set");
  SSF_CHECK(p->linsolver == SSF_SOLVER_CG && p->precision == SSF_PRECISION_FP64, SSF_EUNSUP,
            "ssf_solve_dev_noalloc: the FP64 cg solver only");
  SSF_CHECK(p->ni >= 0, SSF_EINVAL, "ssf_solve_dev_noalloc: negative ni");
  int mode;
  bool jac;
  if (m->storage == SSF_STORAGE_SYMM_COMP_ROWS && p->prec == SSF_PREC_DIAG) {
    mode = MODE_SCR_CG_PREC;  // symcomprow::cg_prec, GEFEL/scr.cpp:1093
    jac = true;
  } else if (m->storage == SSF_STORAGE_SYMM_COMP_ROWS && p->prec == SSF_PREC_NOPRECOND) {
    mode = MODE_SCR_CG;  // symcomprow::cg, GEFEL/scr.cpp:786
    jac = false;
  } else if (m->storage == SSF_STORAGE_COMPRESSED_ROWS && p->prec == SSF_PREC_NOPRECOND) {
    mode = MODE_CR_CG_PREC;  // comprow::cg_prec with noprecond, GEFEL/cr.cpp:1124
    jac = false;
  } else {
    set_error("ssf_solve_dev_noalloc: symm_comp_rows with diagprec or noprecond, or "
              "compressed_rows with noprecond");
    return SSF_EUNSUP;
  }
  SSF_CHECK(m->tune64.blocks <= NA_NB, SSF_EINVAL,
            "ssf_solve_dev_noalloc: the FP64 tiling has more blocks than the partials hold");
  SSF_HIP(hipSetDevice(m->ctx->device));
  double *base = nullptr;
  int rc = scratch_base(m->ctx->device, &base);
  if (rc) return rc;
  const int64_t n = m->n;
  hipStream_t s = m->ctx->stream;
  info->outer = 0;
  info->rel_residual = 0.0;
  if (n == 0) {
    info->status = 1;
    info->iterations = 0;
    return SSF_OK;
  }
  double *r = base, *d = base + NA_MAX_N, *pv = base + 2 * NA_MAX_N, *h = base + 3 * NA_MAX_N;
  double *part = base + 4 * NA_MAX_N;
  PcgState<double> *st = (PcgState<double> *)(part + NA_PART);
  const CsrDev A = m->csr();
  const SpmvTune &t = m->tune64;
  // x = 0 by a fill kernel (no memset node if the caller captures the stream)
  if (!p->iv) SSF_HIP(fill(n, 0.0, lhs_dev, s));
  SSF_HIP(pcg_init<double>(t, jac, p->iv != 0, A, m->val, m->diag, lhs_dev, rhs_dev, r, h, d,
                           part, st, p->zero, s));
  PcgState<double> hs{};
  int64_t it = 0;
  bool stopped = false;
  while (it < p->ni) {
    const int64_t c = std::min<int64_t>(NA_CHUNK, p->ni - it);
    for (int64_t q = 0; q < c; q++)
      SSF_HIP(pcg_iteration<double>(t, mode, jac, A, m->val, m->diag, lhs_dev, r, d, pv, h, part,
                                    st, p->res, p->zero, s));
    it += c;
    SSF_HIP(hipMemcpyAsync(&hs, st, sizeof hs, hipMemcpyDeviceToHost, s));
    SSF_HIP(hipStreamSynchronize(s));
    if (hs.done) {
      stopped = true;
      break;
    }
  }
  if (!stopped) {
    SSF_HIP(hipMemcpyAsync(&hs, st, sizeof hs, hipMemcpyDeviceToHost, s));
    SSF_HIP(hipStreamSynchronize(s));
  }
  info->status = hs.status;
  if (hs.status == 3) {
    // zero right-hand side: x = 0 (GEFEL/scr.cpp:821-830)
    if (mode != MODE_CR_CG_PREC) SSF_HIP(fill(n, 0.0, lhs_dev, s));
    info->ani = 0;
    info->ares = hs.ares;
    info->iterations = 0;
  } else if (hs.status == 1 || hs.status == 2) {
    info->ani = hs.ani;
    info->ares = hs.ares;
    info->iterations = hs.ani + 1;
  } else {
    // ni spent: cg_prec leaves ani/ares; the cg variants set ani = ni, ares = r.h
    if (mode != MODE_SCR_CG_PREC) {
      info->ani = p->ni;
      info->ares = hs.nom[p->ni & 1];
    }
    info->iterations = it;
  }
  // ||b - A x|| / ||b|| in FP64, in the slots of ssf_solve_dev's workspace layout
  SSF_HIP(spmv<double>(t, 2, A, m->val, lhs_dev, pv, rhs_dev, part, nullptr, nullptr, s));
  SSF_HIP(reduce_to_host_slot(part, t.blocks, part + 2 * NA_NB, s));
  SSF_HIP(dot_partial(n, rhs_dev, rhs_dev, part, 256, s));
  SSF_HIP(reduce_to_host_slot(part, 256, part + 2 * NA_NB + 1, s));
  double v[2];
  SSF_HIP(hipMemcpyAsync(v, part + 2 * NA_NB, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  info->rel_residual = v[1] > 0 ? std::sqrt(v[0] / v[1]) : 0.0;
  return SSF_OK;
}
