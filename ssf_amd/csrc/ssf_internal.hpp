// Internal structures of libssf_gpu (MI355X / gfx950). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ssf_gpu.h"

namespace ssf {

void set_error(const std::string &msg);

#define SSF_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::ssf::set_error(std::string("HIP error '") + hipGetErrorString(e_) + "' at " + \
                       __FILE__ + ":" + std::to_string(__LINE__) + " in " + #call); \
      return SSF_EHIP;                                                             \
    }                                                                              \
  } while (0)

// clears an error left by an earlier, unrelated runtime call on this thread so
// that launch checks report only their own launch
#define SSF_ENTER() (void)hipGetLastError()

#define SSF_CHECK(cond, code, msg)  \
  do {                              \
    if (!(cond)) {                  \
      ::ssf::set_error(msg);        \
      return code;                  \
    }                               \
  } while (0)

// Device scalars of one Krylov recurrence. Written by the first thread of block 0
// of the kernel that owns each field; read by later kernels on the same stream.
template <typename T>
struct PcgState {
  T nory;          // ||y||^2
  T nom[2];        // r.h of the current / next iteration (ping-pong by parity)
  T ares;          // attained residual measure (reference meaning per mode)
  T scale;         // mixed: scaling of the inner right-hand side
  long long ani;   // iteration index at exit
  long long iter;  // iteration counter (incremented by the SpMV kernel)
  int status;      // 0 running / ni spent, 1 converged, 2 zero denominator, 3 zero rhs
  int done;        // set once the recurrence has stopped
};

// Krylov recurrence variants of the reference.
enum PcgMode : int {
  MODE_SCR_CG_PREC = 0,  // symcomprow::cg_prec, Jacobi, squared test   GEFEL/scr.cpp:1093
  MODE_SCR_CG = 1,       // symcomprow::cg, no preconditioner          GEFEL/scr.cpp:786
  MODE_CR_CG_PREC = 2,   // comprow::cg_prec with noprecond            GEFEL/cr.cpp:1124
};

struct SpmvTune {
  int lanes = 8;     // lanes (of a 64-wide wavefront) per row
  int blocks = 1024; // blocks of 256 threads (rows split in contiguous chunks)
};

struct Workspace {  // vectors and reduction scratch of one solve shape
  int64_t n = 0;
  int nb = 0;
  double *x = nullptr, *r = nullptr, *d = nullptr, *p = nullptr, *h = nullptr, *y = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *d32 = nullptr, *p32 = nullptr, *h32 = nullptr,
        *y32 = nullptr;
  double *part = nullptr;   // 4 * nb doubles
  float *part32 = nullptr;  // 4 * nb floats
  PcgState<double> *st = nullptr;
  PcgState<float> *st32 = nullptr;
  PcgState<double> *st_host = nullptr;  // pinned
  PcgState<float> *st32_host = nullptr; // pinned
};

}  // namespace ssf

struct ssf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 256;
  int live_matrices = 0;  // a context outlives every matrix created on it
};

struct ssf_matrix {
  ssf_ctx *ctx = nullptr;
  int storage = SSF_STORAGE_SYMM_COMP_ROWS;
  int64_t n = 0;
  int64_t negm = 0;      // entries in the reference layout
  int64_t nnz = 0;       // entries of the full CSR on the device
  std::vector<int64_t> adr, ci;  // reference-layout pattern (host)
  // device full CSR
  int *rowptr = nullptr;
  int *col = nullptr;
  double *val = nullptr;
  float *val32 = nullptr;
  double *diag = nullptr;
  float *diag32 = nullptr;
  // reference layout values on the device and the scatter into the full CSR
  double *a_ref = nullptr;   // == val for comprow
  int *map_lo = nullptr;     // ref entry -> full slot (lower / same)
  int *map_up = nullptr;     // ref entry -> mirrored full slot or -1
  bool values_set = false;
  bool val32_valid = false;
  // element assembly structures (built at the first assemble / localize_batch)
  int64_t asm_ne = 0, asm_ndofe = 0;
  int *slot = nullptr;       // [asm_ne * ndofe * ndofe] reference-layout slot or -1
  int *order = nullptr;      // element ids sorted by colour
  std::vector<int64_t> colour_start;
  std::vector<int64_t> asm_cn_host;  // for identity checks
  ssf::SpmvTune tune64, tune32;
  bool tuned64 = false, tuned32 = false;
  ssf::Workspace ws;
};
