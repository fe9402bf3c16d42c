// Newton-Raphson load control of MEFEL (gnewton_raphson, MEFEL/SRC/newtonraph.cpp:
// 51-268) on device vectors, for linear elastic elements.
//
// The control flow (increments, acceptance, doubling and halving of dlambda, the
// cut at lambdar) runs on the host on two scalars per correction: the norms of
// the unbalanced forces and of the attained load. Everything of length n stays
// in HBM: the attained load fa, the displacements ra and their backup, the
// internal forces fi (the element pass of ssf_elements, loc_internal_forces),
// the unbalanced forces fb = fa - fi and the corrections dr from the solve of
// the stiffness matrix. For linear elastic elements the tangent stiffness is the
// assembled matrix at every state, so it is not assembled again per correction.
#include "ssf_kernels.hpp"

namespace ssf {
namespace {

struct NrWork {
  hipStream_t s = nullptr;
  int64_t n = 0;
  double *rb = nullptr, *fb = nullptr, *fi = nullptr, *dr = nullptr, *part = nullptr;
  ~NrWork()
  {
    dfree(rb, s); dfree(fb, s); dfree(fi, s); dfree(dr, s); dfree(part, s);
  }
};

constexpr int NR_NB = 256;  // blocks of the norm partials

// sqrt(a.a) (GEFEL/vector.cpp:2282 normv), reduced on the device, read back
int norm_dev(NrWork &w, const double *a, double *out)
{
  SSF_HIP(dot_partial(w.n, a, a, w.part, NR_NB, w.s));
  SSF_HIP(reduce_to_host_slot(w.part, NR_NB, w.part + NR_NB, w.s));
  double v = 0.0;
  SSF_HIP(hipMemcpyAsync(&v, w.part + NR_NB, sizeof(double), hipMemcpyDeviceToHost, w.s));
  SSF_HIP(hipStreamSynchronize(w.s));
  *out = std::sqrt(v);
  return SSF_OK;
}

}  // namespace
}  // namespace ssf

using namespace ssf;

extern "C" int ssf_newton_raphson(ssf_matrix *K, ssf_elements *el, const ssf_nr_params *nr,
                                  const ssf_solve_params *sp, const double *pv_dev, int64_t npv,
                                  const double *fc_dev, const double *fp_dev, double *fa_dev,
                                  double *ra_dev, double *lam_hist, double *dlam_hist,
                                  int64_t *nsol_hist, int32_t *acc_hist, ssf_nr_info *info)
{
  SSF_ENTER();
  SSF_CHECK(K && el && nr && sp && fc_dev && fp_dev && fa_dev && ra_dev && info, SSF_EINVAL,
            "ssf_newton_raphson: NULL argument");
  ssf_ctx *ectx = nullptr;
  int kind = 0;
  int64_t en = 0;
  elements_shape(el, &ectx, &kind, &en);
  SSF_CHECK(ectx == K->ctx && en == K->n, SSF_EINVAL,
            "ssf_newton_raphson: the element set and the matrix differ in context or unknowns");
  SSF_CHECK(kind == SSF_ELEM_HEX8_ELAST || kind == SSF_ELEM_QUAD4_PSTRESS, SSF_EUNSUP,
            "ssf_newton_raphson: mechanical element sets only (hex8 elasticity, quad4)");
  SSF_CHECK(nr->rnormtnr == 1 || nr->rnormtnr == 4, SSF_EUNSUP,
            "ssf_newton_raphson: residual norms rel_load_norm (1) and absol_norm (4) only");
  SSF_CHECK(nr->check_div == 0, SSF_EUNSUP,
            "ssf_newton_raphson: the divergence check of the inner loop is not on this path");
  SSF_CHECK(nr->li >= 0 && nr->ninr >= nr->li && nr->niilnr >= 0, SSF_EINVAL,
            "ssf_newton_raphson: bad increment counts");
  SSF_HIP(hipSetDevice(K->ctx->device));
  NrWork w;
  w.s = K->ctx->stream;
  w.n = K->n;
  const int64_t n = w.n;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&w.rb, n, w.s));
  c(dalloc(&w.fb, n, w.s));
  c(dalloc(&w.fi, n, w.s));
  c(dalloc(&w.dr, n, w.s));
  c(dalloc(&w.part, NR_NB + 1, w.s));
  if (e == hipSuccess) c(hipMemsetAsync(w.dr, 0, n * sizeof(double), w.s));  // memset in the reference
  SSF_HIP(e);
  ssf_solve_params q = *sp;
  int64_t solves = 0, rejected = 0, inner_total = 0;
  auto internal_forces = [&]() {
    return ssf_elements_results_dev(el, ra_dev, pv_dev, npv, nullptr, nullptr, w.fi);
  };
  auto unbalanced = [&]() -> int {  // fb = fa - fi (subv)
    int rc = internal_forces();
    if (rc) return rc;
    SSF_HIP(vec_sub(n, w.fb, fa_dev, w.fi, w.s));
    return SSF_OK;
  };
  auto correct = [&]() -> int {  // solve_system(K, dr, fb); ra += dr (addv)
    ssf_solve_info si{};
    int rc = ssf_solve_dev(K, &q, w.fb, w.dr, &si);
    if (rc) return rc;
    solves++;
    SSF_HIP(vals_add(n, ra_dev, 1.0, w.dr, w.s));
    return SSF_OK;
  };
  auto res_norm = [&](double *norf) -> int {  // compute_res_norm_nr (newtonraph.cpp:866)
    int rc = norm_dev(w, w.fb, norf);
    if (rc) return rc;
    if (nr->rnormtnr == 1) {
      double norfa = 0.0;
      rc = norm_dev(w, fa_dev, &norfa);
      if (rc) return rc;
      if (norfa > 1.0e-9) *norf /= norfa;
    }
    return SSF_OK;
  };
#define NR_TRY(x)              \
  do {                         \
    int rc_ = (x);             \
    if (rc_) return rc_;       \
  } while (0)
  double lambda = nr->ilambda, dlambda = nr->incrnr, norf = 0.0;
  bool check_rv = nr->check_lambdar != 0;
  int modif = 0;
  int32_t status = 1;
  if (lambda == 0.0) {  // assemble_attained_load_vector (MEFEL/SRC/nssolver.cpp:312)
    double aux = lambda;
    if (nr->check_lambdar && lambda > nr->lambdar) aux = nr->lambdar;
    SSF_HIP(axpy_rn(n, fa_dev, fc_dev, aux, fp_dev, w.s));
  }
  int64_t i;
  for (i = nr->li; i < nr->ninr; i++) {
    const int64_t k = i - nr->li, s0 = solves;
    SSF_HIP(hipMemcpyAsync(w.rb, ra_dev, n * sizeof(double), hipMemcpyDeviceToDevice, w.s));
    const double blambda = lambda;
    if (check_rv && dlambda + lambda > nr->lambdar) {
      check_rv = false;
      dlambda = nr->lambdar - lambda;
    }
    if (dlam_hist) dlam_hist[k] = dlambda;
    SSF_HIP(vals_add(n, fa_dev, dlambda, fp_dev, w.s));  // fa += fp dlambda
    NR_TRY(unbalanced());
    NR_TRY(correct());
    NR_TRY(unbalanced());
    NR_TRY(res_norm(&norf));
    if (norf < nr->errnr) {
      lambda += dlambda;
      if (lam_hist) lam_hist[k] = lambda;
      if (nsol_hist) nsol_hist[k] = solves - s0;
      if (acc_hist) acc_hist[k] = 1;
      modif++;
      if (modif > 1) {
        dlambda *= 2.0;
        if (dlambda > nr->maxincrnr) dlambda = nr->maxincrnr;
        modif = 0;
      }
      if (nr->check_lambdar && !check_rv) {
        i++;
        break;
      }
      continue;
    }
    int64_t j;
    for (j = 0; j < nr->niilnr; j++) {
      NR_TRY(correct());
      NR_TRY(unbalanced());
      NR_TRY(res_norm(&norf));
      inner_total++;
      if (norf < nr->errnr) break;
    }
    modif = 0;
    if (nsol_hist) nsol_hist[k] = solves - s0;
    if (j == nr->niilnr || norf > nr->errnr) {
      SSF_HIP(vals_add(n, fa_dev, -dlambda, fp_dev, w.s));  // fa += fp (-dlambda)
      SSF_HIP(hipMemcpyAsync(ra_dev, w.rb, n * sizeof(double), hipMemcpyDeviceToDevice, w.s));
      lambda = blambda;
      rejected++;
      if (lam_hist) lam_hist[k] = lambda;
      if (acc_hist) acc_hist[k] = 0;
      if (dlambda == nr->minincrnr) {
        status = 2;  // the increment cannot be decreased: no equilibrium reached
        i++;
        break;
      }
      dlambda /= 2.0;
      if (dlambda < nr->minincrnr) dlambda = nr->minincrnr;
    } else {
      lambda += dlambda;
      if (lam_hist) lam_hist[k] = lambda;
      if (acc_hist) acc_hist[k] = 1;
      if (nr->check_lambdar && !check_rv) {
        i++;
        break;
      }
    }
  }
#undef NR_TRY
  SSF_HIP(hipStreamSynchronize(w.s));
  info->lambda = lambda;
  info->norf = norf;
  info->increments = i - nr->li;
  info->solves = solves;
  info->rejected = rejected;
  info->inner = inner_total;
  info->status = status;
  return SSF_OK;
}
