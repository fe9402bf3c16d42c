// gmatrix_gpu members: each forwards to the C ABI of libssf_gpu (see the header).
#include "gmatrix_gpu.hpp"

#include <cstring>

namespace ssf_host {

gmatrix_gpu::gmatrix_gpu(int device)
{
  check(ssf_ctx_create(device, &ctx_), "ssf_ctx_create");
}

gmatrix_gpu::~gmatrix_gpu()
{
  if (m_) ssf_matrix_destroy(m_);
  if (ctx_) ssf_ctx_destroy(ctx_);
}

// the reference reports with print_err (GEFEL/iotools.cpp:25) and goes on
void gmatrix_gpu::check(int rc, const char *what)
{
  if (rc == SSF_OK) return;
  errors_++;
  std::fprintf(stderr, "\n\n ERROR in %s: %s (code %d)\n", what, ssf_last_error(), rc);
}

void gmatrix_gpu::initiate(int ts, long ndof, long ne, long ndofe, const long *cn)
{
  if (m_) ssf_matrix_destroy(m_);
  m_ = nullptr;
  n_ = ndof;
  check(ssf_matrix_create_elem(ctx_, ts, ndof, ne, ndofe, (const int64_t *)cn, &m_),
        "gmatrix_gpu::initiate");
}

void gmatrix_gpu::initiate_arrays(int ts, long n, const long *adr, const long *ci)
{
  if (m_) ssf_matrix_destroy(m_);
  m_ = nullptr;
  n_ = n;
  check(ssf_matrix_create_csr(ctx_, ts, n, (const int64_t *)adr, (const int64_t *)ci, &m_),
        "gmatrix_gpu::initiate_arrays");
}

void gmatrix_gpu::localized(const double *lm, const long *cn, long nc, long eid)
{
  (void)eid;  // the element order is the order of the calls, as in the reference
  if (nc_ && nc != nc_) prepmat();  // one element size per batch
  nc_ = nc;
  st_cn_.insert(st_cn_.end(), cn, cn + nc);
  st_lm_.insert(st_lm_.end(), lm, lm + nc * nc);
}

void gmatrix_gpu::nullmat()
{
  st_cn_.clear();
  st_lm_.clear();
  nc_ = 0;
  if (m_) check(ssf_matrix_zero(m_), "gmatrix_gpu::nullmat");
}

void gmatrix_gpu::prepmat()
{
  if (m_ && nc_ && !st_cn_.empty())
    check(ssf_matrix_localize_batch(m_, (int64_t)(st_cn_.size() / nc_), nc_, st_cn_.data(),
                                    st_lm_.data()),
          "gmatrix_gpu::prepmat");
  st_cn_.clear();
  st_lm_.clear();
  nc_ = 0;
}

void gmatrix_gpu::solve_system(double *lhs, const double *rhs)
{
  prepmat();
  ssf_solve_params p;
  std::memset(&p, 0, sizeof p);
  p.linsolver = tlinsol;
  p.prec = prec;
  p.ni = ni;
  p.res = res;
  p.zero = zero;
  p.iv = (int32_t)iv;
  p.precision = precision;
  ssf_solve_info info;
  std::memset(&info, 0, sizeof info);
  info.ani = ani;  // left unchanged by cg_prec when ni is spent
  info.ares = ares;
  check(ssf_solve(m_, &p, rhs, lhs, &info), "gmatrix_gpu::solve_system");
  ani = info.ani;
  ares = info.ares;
  status = info.status;
}

void gmatrix_gpu::gmxv(const double *a, double *b)
{
  check(ssf_gmxv(m_, a, b), "gmatrix_gpu::gmxv");
}

void gmatrix_gpu::give_values(double *a) const
{
  const_cast<gmatrix_gpu *>(this)->check(ssf_matrix_get_values(m_, a), "gmatrix_gpu::give_values");
}

long gmatrix_gpu::negm() const
{
  int64_t n = 0, negm = 0, nnz = 0;
  if (m_) ssf_matrix_sizes(m_, &n, &negm, &nnz);
  return (long)negm;
}

}  // namespace ssf_host
