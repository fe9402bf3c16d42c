// Internal structures of libssf_gpu (MI355X / gfx950). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "ssf_gpu.h"

namespace ssf {

void set_error(const std::string &msg);

inline int host_threads()
{
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  return (int)std::min(16u, h);
}

// f(b, e) over contiguous ranges of [0, n) on up to 16 host threads
template <typename F>
void parallel_for(int64_t n, F f)
{
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / 4096 + 1));
  if (nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    const int64_t b = n * t / nt, e = n * (t + 1) / nt;
    th.emplace_back([=]() { f(b, e); });
  }
  for (auto &x : th) x.join();
}

// Device memory of matrices and workspaces is stream-ordered (hipMallocAsync /
// hipFreeAsync on the owning context's stream): releasing it never waits for
// work on other streams, e.g. a served request's solve while another request
// is being assembled.
template <typename T>
hipError_t dalloc(T **p, size_t count, hipStream_t s)
{
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMallocAsync((void **)p, count * sizeof(T), s);
}

template <typename T>
void dfree(T *&p, hipStream_t s)
{
  if (p) (void)hipFreeAsync(p, s);
  p = nullptr;
}

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

// Device state of the mixed-precision (reliable-update) solve.
struct MixState {
  double nory64;   // ||b||^2
  double rr64;     // ||b - A x||^2 at the last refresh
  double best;     // smallest rr64 that was a 4x improvement
  double scale;    // FP32 quantities are the FP64 ones times scale = 1/||r|| at the last refresh
  double delta2;   // the FP32 recurrence refreshes after reducing its residual by delta^2
  double norm_a;   // ||A||_inf (FP64 values), for the rounding floor of b - A x
  long long updates;
  int bad;         // consecutive refreshes whose FP32 reduction the FP64 residual did not follow
  int since_best;  // refreshes since rr64 last improved by 10 %
  int final;       // the solve has ended (status set)
  int status;      // 1 converged, 4 FP64 floor reached above res
};

// The mixed solve's two device states side by side: one copy brings both to the
// host.
struct MixProbe {
  MixState mix;
  PcgState<float> st;
};

// Krylov recurrence variants of the reference.
enum PcgMode : int {
  MODE_SCR_CG_PREC = 0,  // symcomprow::cg_prec, Jacobi, squared test   GEFEL/scr.cpp:1093
  MODE_SCR_CG = 1,       // symcomprow::cg, no preconditioner          GEFEL/scr.cpp:786
  MODE_CR_CG_PREC = 2,   // comprow::cg_prec with noprecond            GEFEL/cr.cpp:1124
  MODE_MIXED_INNER = 3,  // FP32 recurrence of the mixed solve: the direction is
                         // updated before the test rr < nory stops it
};

struct SpmvTune {
  int format = 0;    // 0: row-vector kernel, 1: LDS-tiled stream kernel, 2: SELL-64 slices
  int lanes = 8;     // lanes (of a 64-wide wavefront) per row
  int blocks = 1024; // blocks of 256 threads (rows / tiles split in contiguous chunks)
  int idx16 = 0;     // column index form: 0 int32; 1 16-bit (stream tile- / SELL
                     // slice-relative); 2 SELL row patterns (a pattern id per row
                     // and a shared dictionary of column offsets); 3 SELL stencil form;
                     // 4 symmetric half of the stencil form (lanes: 256-thread units
                     // per block, 1 or 2)
  int map = 1;       // SELL slices to waves: 0 a contiguous chunk per block; 1 XCD
                     // interleave (each XCD a contiguous eighth, its waves in turn)
};

// Device view of a full CSR matrix and its stream tiling.
constexpr int SELL_STEN_MAX = 32;    // slots of the stencil form held in the kernel arguments
constexpr int SELL_STEN_MAX_B = 96;  // slots of the block stencil form (offsets in device memory)

struct CsrDev {
  int n = 0;
  const int *rowptr = nullptr;
  const int *col = nullptr;
  const short *col16 = nullptr;  // col - first row of the tile (null if it does not fit)
  const int *tiles = nullptr;    // first row of each tile, ntiles + 1 entries
  int ntiles = 0;
  // SELL-64: slice s = rows 64s..64s+63, stored column-major in groups of
  // SELL_V entries per row: entry q of row 64s + l at
  // sell_off[s] + (q / SELL_V) * 64 * SELL_V + l * SELL_V + q % SELL_V
  int nslice = 0;
  const int *sell_off = nullptr;  // nslice + 1, multiples of 64 * SELL_V
  const int *sell_col = nullptr;  // padding: column of a row of the slice, value 0
  const short *sell_col16 = nullptr;  // sell_col - 64 s (null if an offset does not fit)
  // row patterns: row r of the layout (r < 64 nslice) reads its entries' columns
  // as min(r, n-1) + sell_dict[sell_dict_off[sell_pid[r]] + q]
  const unsigned short *sell_pid = nullptr;
  const int *sell_dict = nullptr;
  const int *sell_dict_off = nullptr;
  // stencil form: every row of the layout has the same slots, slot k at column
  // base + sten[k] with base = row - row % sten_b (sten ascending; a slot the row
  // lacks holds value 0), so the columns need no stream and no dictionary.
  // sten_b = 1: offsets from the row itself, at most SELL_STEN_MAX of them, in
  // sten; sten_b = 2, 3: offsets from the first row of the row's block of
  // sten_b rows (the unknowns of one node), read from sten_d
  int sten_n = 0;
  int sten_b = 1;
  int sten[SELL_STEN_MAX] = {};
  const int *sten_d = nullptr;
  const double *sell_v64 = nullptr;
  const float *sell_v32 = nullptr;
  // symmetric half of the stencil form (sten_b = 1, sten[k] = -sten[sten_n - 1 - k],
  // sten[half_nl] = 0): the half_nl + 1 lower slots of every row, diagonal-major,
  // half_v[k * half_pad + r] = a(r, r + sten[k]), 0 for rows r >= n; half_pad
  // covers 64 nslice - sten[0] rows, so the mirrored slot a(r, r - sten[k]) =
  // half_v[k * half_pad + r - sten[k]] needs no bounds test
  int half_nl = 0;
  int64_t half_pad = 0;
  const double *half_v64 = nullptr;
  const float *half_v32 = nullptr;
  // block variant (half_b = 2, 3: the unknowns of a node numbered together, the
  // block stencil form): node q holds nv = half_nl half_b^2 + half_b (half_b + 1) / 2
  // entries, the half_b x half_b blocks towards its half_nl lower neighbour nodes
  // q + half_m[j] (half_m ascending, negative), entry (t, e) of block j as entry
  // i = (j half_b + t) half_b + e, then the lower triangle of its diagonal block,
  // entry (t, e <= t) as i = half_nl half_b^2 + t (t + 1) / 2 + e. Slice-major:
  // entry i of node 64 s + l at half_v[(s nv + i) 64 + l]; 0 for nodes beyond
  // n / half_b. half_pad (a multiple of 64) covers 64 ceil(n / half_b / 64) -
  // half_m[0] nodes.
  int half_b = 1;
  int half_m[16] = {};
  int slice_map = 1;  // SpmvTune::map of the launch
};

constexpr int SELL_V = 4;  // entries per lane and load (16-byte FP32, 32-byte FP64 loads)


struct Workspace {  // vectors and reduction scratch of one solve shape
  int64_t n = 0;
  int nb = 0;
  double *x = nullptr, *r = nullptr, *d = nullptr, *p = nullptr, *h = nullptr, *y = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *d32 = nullptr, *p32 = nullptr, *h32 = nullptr,
        *y32 = nullptr;
  double *part = nullptr;   // 4 * nb doubles
  float *part32 = nullptr;  // 4 * nb floats
  PcgState<double> *st = nullptr;
  PcgState<float> *st32 = nullptr;      // &probe->st (mixed solves)
  PcgState<double> *st_host = nullptr;  // pinned
  MixProbe *probe = nullptr;            // holds *mix and *st32
  MixProbe *probe_host = nullptr;       // pinned, 2 entries
  MixState *mix = nullptr;              // &probe->mix
};

struct DualSell;  // union layout of A and A^T of a comprow matrix (ssf_bicg.hip)
struct IldlPlan;  // incomplete LDL^T levels and factor of a symcomprow matrix (ssf_ildl.hip)

}  // namespace ssf

struct ssf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 256;
  // a context outlives every matrix created on it; matrices of one context may
  // be created and destroyed on different host threads (served solves)
  std::atomic<int> live_matrices{0};
  // kernel sampling inside solves: the second SpMV of every iteration chunk is
  // bracketed by HIP events on the solve stream; [0] FP64, [1] FP32 operator
  bool prof = false;
  double prof_us[2] = {0.0, 0.0};
  long long prof_n[2] = {0, 0};
  hipEvent_t prof_ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [parity][start/end]
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
  double *isd = nullptr;     // 1/sqrt|a_ii|: the scaling of the FP32 operator (mixed solves)
  double *norm_a = nullptr;  // ||A||_inf, with the FP32 copy (mixed solves)
  // reference layout values on the device and the scatter into the full CSR
  double *a_ref = nullptr;   // == val for comprow
  int *map_lo = nullptr;     // ref entry -> full slot (lower / same)
  int *map_up = nullptr;     // ref entry -> mirrored full slot or -1
  bool values_set = false;
  bool val32_valid = false;
  // element assembly plan (built at the first assemble / localize_batch of a
  // code-number table): the contributions e nd^2 + i nd + j of the element
  // matrices to entry s, asm_cidx[asm_cstart[s] .. asm_cstart[s + 1]), in the
  // reference's localize order
  int64_t asm_ne = 0, asm_ndofe = 0, asm_nc = 0;
  int *asm_cstart = nullptr;  // [negm + 1]
  int *asm_cidx = nullptr;    // [asm_nc]
  std::vector<int64_t> asm_cn_host;  // for identity checks
  // stream tiling: whole rows, at most STREAM_CAP entries per tile
  int *tiles = nullptr;
  int ntiles = 0;
  short *col16 = nullptr;
  // SELL-64 layout of the full CSR (values gathered from val through sell_src)
  int nslice = 0;
  int64_t sell_tot = 0;
  int *sell_off = nullptr, *sell_col = nullptr, *sell_src = nullptr;
  short *sell_col16 = nullptr;
  unsigned short *sell_pid = nullptr;  // row patterns (null when not built)
  int *sell_dict = nullptr, *sell_dict_off = nullptr;
  int sell_npat = 0;
  std::vector<int> sten;  // stencil form of the SELL layout (empty: not stencil-aligned)
  int sten_b = 1;         // its block: offsets from row - row % sten_b
  int *sten_d = nullptr;  // device copy of sten
  double *sell_v64 = nullptr;
  float *sell_v32 = nullptr;
  // symmetric half of the stencil form (CsrDev::half_*), rebuilt from the SELL
  // values whenever they change; half_nl = 0: the matrix has no such form
  int half_nl = 0;
  int64_t half_pad = 0;
  double *half_v64 = nullptr;
  float *half_v32 = nullptr;
  int half_b = 1;
  std::vector<int> half_m;   // block variant: lower node offsets, ascending
  int *half_m_d = nullptr;   // their device copy
  int *half_slot = nullptr;  // block variant: (SELL slot, row in block) of every stored entry
  // SELL-only matrices (laid out on the device, no full CSR): SELL positions of
  // the reference-layout entries and the device copy of the pattern
  bool sell_only = false;
  int *sell_lo = nullptr, *sell_up = nullptr, *d_adr32 = nullptr, *d_ci32 = nullptr;
  ssf::SpmvTune tune64, tune32;
  bool tuned64 = false, tuned32 = false;
  // fingerprint of the reference-layout pattern (adr, ci), for operations that
  // combine two matrices' values; counts every change of the values
  uint64_t pattern_fp = 0;
  uint64_t values_version = 0;
  // biconjugate gradients: union layout of A and A^T with both values (built at
  // the first bicg solve)
  ssf::DualSell *dual = nullptr;
  // incomplete LDL^T: dependency levels of the pattern, the factor of the values
  // (built at the first incomdec solve, factored again after a values change)
  ssf::IldlPlan *ildl = nullptr;
  // as the system matrix of a nonstationary step: the inputs it was formed from
  const ssf_matrix *np_k = nullptr, *np_c = nullptr;
  uint64_t np_kv = 0, np_cv = 0;
  double np_s = 0.0;
  ssf::CsrDev csr() const
  {
    ssf::CsrDev c;
    c.n = (int)n;
    c.rowptr = rowptr;
    c.col = col;
    c.col16 = col16;
    c.tiles = tiles;
    c.ntiles = ntiles;
    c.nslice = nslice;
    c.sell_off = sell_off;
    c.sell_col = sell_col;
    c.sell_col16 = sell_col16;
    c.sell_pid = sell_pid;
    c.sell_dict = sell_dict;
    c.sell_dict_off = sell_dict_off;
    c.sten_n = (int)sten.size();
    c.sten_b = sten_b;
    for (size_t k = 0; k < sten.size() && k < (size_t)ssf::SELL_STEN_MAX; k++) c.sten[k] = sten[k];
    c.sten_d = sten_d;
    c.sell_v64 = sell_v64;
    c.sell_v32 = sell_v32;
    c.half_nl = half_nl;
    c.half_pad = half_pad;
    c.half_v64 = half_v64;
    c.half_v32 = half_v32;
    c.half_b = half_b;
    for (size_t k = 0; k < half_m.size() && k < 16; k++) c.half_m[k] = half_m[k];
    return c;
  }
  ssf::Workspace ws;
};

namespace ssf {
// pattern checksum: sum over adr and ci of (value + 1) * mix64(index ^ salt),
// wrapping; the same on the host and the device (order-independent sums)
__host__ __device__ inline unsigned long long mix64(unsigned long long x)
{
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
constexpr unsigned long long kSaltAdr = 0x5AD5AD5AD5ULL, kSaltCi = 0xC1C1C1C1C1ULL;
hipError_t pattern_checksum_dev(const int *adr, int64_t n, const int *ci, int64_t negm,
                                hipStream_t s, uint64_t *out);
// symcomprow pattern built on the device from element code numbers cn[ne*nd]
// (int32, device); adr[n+1], ci[negm] are allocated stream-ordered on s
int pattern_build_dev(int64_t n, int64_t ne, int nd, const int *cn, hipStream_t s, int **adr,
                      int **ci, int64_t *negm);
int64_t pattern_build_bytes(int64_t ne, int nd);  // its transient device bytes
// entries of the SELL-64 layout of the full matrix of a device symcomprow pattern
hipError_t sell_entries_dev(int64_t n, int64_t negm, const int *adr, const int *ci, hipStream_t s,
                            int64_t *tot);
// matrix_create_sell from a device pattern (ownership of adr / ci passes to the matrix);
// half_form: also keep the symmetric half of a stencil layout for the solves
int matrix_create_sell_dev(ssf_ctx *ctx, int64_t n, int *adr, int *ci, int64_t negm,
                           ssf_matrix **out, bool half_form = true);
// symbolic pattern of symcomprow / comprow from element code numbers (host)
int pattern_from_elements(int storage, int64_t n, int64_t ne, int64_t ndofe, const int64_t *cn,
                          std::vector<int64_t> &adr, std::vector<int64_t> &ci);
// device storage of a reference-layout pattern (ssf_matrix_create_csr); the
// host pattern vectors are moved in when given, the stream tiling is optional
int matrix_create(ssf_ctx *ctx, int storage, int64_t n, const int64_t *adr, const int64_t *ci,
                  std::vector<int64_t> *adr_own, std::vector<int64_t> *ci_own, bool stream_tiles,
                  ssf_matrix **out);
// a symcomprow matrix held in the SELL-64 layout only, laid out on the device
// from the pattern (moved in); assembled, multiplied and solved in FP64
int matrix_create_sell(ssf_ctx *ctx, int64_t n, std::vector<int64_t> &adr,
                       std::vector<int64_t> &ci, ssf_matrix **out);
// (re)builds the half stencil form of the FP64 (precision 0) or FP32 SELL values
// when the matrix has one (half_nl > 0)
int matrix_half_build(ssf_matrix *m, int precision);
// values of the half form per precision (0 when the matrix has none)
int64_t matrix_half_entries(const ssf_matrix *m);
// releases the element-assembly plan of a matrix
// and, if host_pattern, its host copy of the reference-layout pattern
void matrix_drop_assembly(ssf_matrix *m, bool host_pattern);
// keeps only what the SELL-64 solve reads (slice offsets, columns of the
// narrowest width, SELL values, diagonal): the CSR and the FP32 copies go
void matrix_keep_sell_only(ssf_matrix *m);
// SELL-64 layout of a full CSR (rows of 64-row slices padded to the slice's
// longest row, rounded up to SELL_V; padding: the row's own column, value 0);
// src maps a SELL entry to its CSR entry or -1. false beyond 2^31 entries.
bool build_sell(int64_t n, const std::vector<int> &rowptr, const std::vector<int> &col,
                std::vector<int> &off, std::vector<int> &scol, std::vector<int> &src,
                std::vector<short> &scol16,
                std::vector<int> *sten = nullptr, int *sten_b = nullptr);
// row patterns of a SELL-64 layout (off, scol from build_sell): per row of the
// layout its column offsets from min(row, n-1), deduplicated into dict (each
// pattern's entries start at a multiple of SELL_V, dict_off[p]); pid[r] = the
// pattern of layout row r. false when there are more than 65535 patterns or
// max_dict dictionary entries (the layout keeps its columns then).
bool build_sell_patterns(int64_t n, const std::vector<int> &off, const std::vector<int> &scol,
                         std::vector<unsigned short> &pid, std::vector<int> &dict,
                         std::vector<int> &dict_off, int64_t max_dict);
// comprow::bicg (GEFEL/cr.cpp:1353) of a compressed_rows matrix on device vectors;
// fills info except rel_residual
int solve_bicg(ssf_matrix *m, const ssf_solve_params *p, const double *y, double *x,
               ssf_solve_info *info);
void dual_free(DualSell *L, hipStream_t s);
// cg_prec with tprec = incomdec (GEFEL/scr.cpp:1093, :990, :1031) of a symcomprow
// matrix on device vectors; fills info except rel_residual
int solve_ildl(ssf_matrix *m, const ssf_solve_params *p, const double *y, double *x,
               ssf_solve_info *info);
int ildl_factor_copy(ssf_matrix *m, double *out);                 // factor -> host
int ildl_apply_dev(ssf_matrix *m, const double *y, double *h);    // h = (LDL^T)^-1 y
int ildl_bench(ssf_matrix *m, int reps, int graph, int64_t *nfwd, int64_t *nbwd, double *usec);
void ildl_free(IldlPlan *P, hipStream_t s);
// context, element kind and number of unknowns of an element set (ssf_post.hip)
void elements_shape(const ssf_elements *el, ssf_ctx **ctx, int *kind, int64_t *n);
}  // namespace ssf
