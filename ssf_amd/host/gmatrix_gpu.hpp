// gmatrix_gpu: the symcomprow / comprow part of GEFEL's gmatrix (GEFEL/gmatrix.h:29-153)
// on libssf_gpu, for host C++ callers. The members keep the reference's names,
// argument meaning and error behaviour (print_err-style message, execution goes
// on); each forwards to one entry point of include/ssf_gpu.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "ssf_gpu.h"

namespace ssf_host {

// the layouts the C ABI fixes (bindings in other languages repeat them)
static_assert(sizeof(long) == sizeof(int64_t), "code numbers are the reference's 64-bit long");
static_assert(sizeof(ssf_solve_params) == 64, "ssf_solve_params layout");
static_assert(offsetof(ssf_solve_params, ni) == 8 && offsetof(ssf_solve_params, res) == 16 &&
                  offsetof(ssf_solve_params, zero) == 24 && offsetof(ssf_solve_params, iv) == 32 &&
                  offsetof(ssf_solve_params, precision) == 36 &&
                  offsetof(ssf_solve_params, inner_res) == 40 &&
                  offsetof(ssf_solve_params, inner_ni) == 48 &&
                  offsetof(ssf_solve_params, max_outer) == 56,
              "ssf_solve_params field offsets");
static_assert(sizeof(ssf_solve_info) == 48, "ssf_solve_info layout");
static_assert(offsetof(ssf_solve_info, ares) == 8 && offsetof(ssf_solve_info, status) == 16 &&
                  offsetof(ssf_solve_info, iterations) == 24 &&
                  offsetof(ssf_solve_info, outer) == 32 &&
                  offsetof(ssf_solve_info, rel_residual) == 40,
              "ssf_solve_info field offsets");

class gmatrix_gpu {
 public:
  // slesolv / precond fields the solve reads (slesolv::tlinsol, ni, res, iv;
  // precond::pt; gmatrix::zero = 1e-30, GEFEL/gmatrix.cpp:9)
  long ni = 1000;
  double res = 1e-20;
  long iv = 0;
  double zero = 1e-30;
  int tlinsol = SSF_SOLVER_CG;
  int prec = SSF_PREC_DIAG;
  int precision = SSF_PRECISION_FP64;
  // results of the last solve (slesolv::ani, ares)
  long ani = 0;
  double ares = 0.0;
  int status = 0;

  explicit gmatrix_gpu(int device = 0);
  ~gmatrix_gpu();
  gmatrix_gpu(const gmatrix_gpu &) = delete;
  gmatrix_gpu &operator=(const gmatrix_gpu &) = delete;

  // gmatrix::initiate (GEFEL/gmatrix.cpp:268) from element code numbers
  // cn[ne * ndofe] (gtopology::give_code_numbers, 1-based, <= 0 skipped);
  // ts = symm_comp_rows (11) or compressed_rows (10)
  void initiate(int ts, long ndof, long ne, long ndofe, const long *cn);
  // gmatrix::initiate from the reference's own arrays adr[n + 1], ci[adr[n]]
  void initiate_arrays(int ts, long n, const long *adr, const long *ci);
  // gmatrix::localized (GEFEL/gmatrix.cpp:581): lm[nc * nc] row-major of element
  // eid with code numbers cn[nc]; staged, added in element order at prepmat
  void localized(const double *lm, const long *cn, long nc, long eid);
  // gmatrix::nullmat
  void nullmat();
  // gmatrix::prepmat: the staged element matrices go to the device
  void prepmat();
  // gmatrix::solve_system (GEFEL/gmatrix.cpp:1049): lhs[n] (in/out), rhs[n]
  void solve_system(double *lhs, const double *rhs);
  // gmatrix::gmxv (GEFEL/gmatrix.cpp:1704): b = A a
  void gmxv(const double *a, double *b);
  // the values in the reference layout (symcomprow::a / comprow::a)
  void give_values(double *a) const;
  long n() const { return n_; }
  long negm() const;
  bool ok() const { return errors_ == 0; }

 private:
  void check(int rc, const char *what);
  ssf_ctx *ctx_ = nullptr;
  ssf_matrix *m_ = nullptr;
  long n_ = 0;
  int errors_ = 0;
  long nc_ = 0;                  // code numbers per staged element
  std::vector<int64_t> st_cn_;   // staged code numbers
  std::vector<double> st_lm_;    // staged element matrices
};

}  // namespace ssf_host
