// Launchers of the gfx950 kernels (defined in ssf_kernels.hip / ssf_elast.hip).
#pragma once

#include <functional>
#include <vector>

#include "ssf_internal.hpp"

namespace ssf {

constexpr int TPB = 256;          // threads per block for every streaming kernel
constexpr int STREAM_KPT = 8;     // matrix entries per thread and tile in the stream SpMV
constexpr int STREAM_CAP = TPB * STREAM_KPT;  // entries per stream tile

// y = A x (kind 0); y = A x with partial[b] = sum x_i y_i (kind 1);
// y = b - A x with partial[b] = sum y_i^2 (kind 2). done (may be null) skips the
// launch body once set; iter (may be null) is incremented by block 0. ev0/ev1
// (may be null) receive the kernel's own start and end times.
template <typename T>
hipError_t spmv(const SpmvTune &t, int kind, const CsrDev &A, const T *val, const T *x, T *y,
                const T *b, T *partial, const int *done, long long *iter, hipStream_t s,
                hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// the symmetric half of a stencil-form SELL layout (CsrDev::half_*): hv[k * half_pad
// + r] = slot k of row r of sval for k <= A.half_nl, 0 for r >= n
template <typename T>
hipError_t half_from_sell(const CsrDev &A, const T *sval, T *hv, hipStream_t s);
// the block variant (A.half_b > 1): stored entry i of every node from SELL slot
// slot[2 i] of row half_b q + slot[2 i + 1]; nv entries per node
template <typename T>
hipError_t half_from_sell_block(const CsrDev &A, const int *slot, int nv, const T *sval, T *hv,
                                hipStream_t s);

// PCG pieces (see ssf_kernels.hip for the recurrence).
template <typename T>
hipError_t pcg_init(const SpmvTune &t, bool jac, bool havex, const CsrDev &A, const T *val,
                    const T *diag, const T *x, const T *y, T *r, T *h, T *d, T *part,
                    PcgState<T> *st, T zero, hipStream_t s);
template <typename T>
hipError_t pcg_iteration(const SpmvTune &t, int mode, bool jac, const CsrDev &A, const T *val,
                         const T *diag, T *x, T *r, T *d, T *p, T *h, T *part,
                         PcgState<T> *st, T res, T zero, hipStream_t s,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// FP64 PCG with a preconditioner enqueued by prec (h = M^-1 r, on stream s, reading
// r and writing h; it may skip its work once st->done is set), in the cg_prec
// recurrence and stopping test (GEFEL/scr.cpp:1093-1203). Partial layout as
// pcg_iteration's, r.h in [P1 + nv, P1 + 2 nv).
hipError_t pcg_init_prec(const SpmvTune &t, bool havex, const CsrDev &A, const double *val,
                         const double *x, const double *y, double *r, double *h, double *d,
                         double *part, PcgState<double> *st, double zero, hipStream_t s,
                         const std::function<hipError_t()> &prec);
hipError_t pcg_iteration_prec(const SpmvTune &t, const CsrDev &A, const double *val, double *x,
                              double *r, double *d, double *p, double *h, double *part,
                              PcgState<double> *st, double res, double zero, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1,
                              const std::function<hipError_t()> &prec);

// mixed precision with reliable updates (ssf_solver.hip). The FP32 operator is
// float(S A S) with S = diag(1/sqrt|a_ii|): isd = diag(S) (n doubles), val32 from
// the full CSR values (if val32 and A.rowptr), sval32 from the SELL values (if
// sval32 and A.nslice)
hipError_t mixed_operator(const CsrDev &A, const double *diag, const double *val, float *val32,
                          const double *sval, float *sval32, double *isd, hipStream_t s);
// isd = 1/sqrt|diag| (1 where diag = 0); val32 = float(val_ij isd_i isd_j) of a CSR
hipError_t inv_sqrt_diag(int n, const double *diag, double *isd, hipStream_t s);
hipError_t csr_scale32(int n, const int *rowptr, const int *col, const double *val,
                       const double *isd, float *val32, hipStream_t s);
// r32 = float(r64 * scale * isd)
hipError_t mixed_rhs32(int n, const double *r64, const double *isd, float *r32, double scale,
                       hipStream_t s);
// after pcg_init of the FP32 recurrence (CG, no preconditioner)
hipError_t mixed_init(MixState *ms, PcgState<float> *st, const double *slots,
                      const double *norm_a, double res, double delta, hipStream_t s);
// px: vec_blocks(t32, n) doubles of scratch (partials of ||x64||^2)
hipError_t mixed_update(const SpmvTune &t64, const SpmvTune &t32, const CsrDev &A,
                        const double *val64, const double *isd, const double *b, double *x64,
                        double *r64, double *part64, double *px, float *x32, float *r32,
                        float *d32, float *part32, PcgState<float> *st, MixState *ms,
                        double res, hipStream_t s);
// ||A||_inf into *out (device); partial: nb doubles of scratch
hipError_t row_absmax(int n, const int *rowptr, const double *val, double *partial, int nb,
                      double *out, hipStream_t s);
// the same from the SELL-64 layout (sell_off: nslice + 1 offsets, v: its values)
hipError_t sell_row_absmax(int n, const int *sell_off, const double *v, double *partial, int nb,
                           double *out, hipStream_t s);
hipError_t copy_kernel(int64_t n, const double *a, double *b, hipStream_t s);
// SELL values from the full CSR values: out[k] = src[k] >= 0 ? val[src[k]] : 0
hipError_t sell_gather(int64_t tot, const int *src, const double *val, double *out, hipStream_t s);

// SELL-64 layout of the full matrix of a symcomprow pattern, built on the device
// (ssf_layout.hip). lo[k] / up[k]: SELL positions of scr entry k in its row and,
// off the diagonal, mirrored in its column's row (-1 on the diagonal). Exactly
// one of col (32-bit) and col16 (16-bit offsets from the slice's first row) is
// returned, col16 when every offset fits.
struct SellLayout {
  int nslice = 0;
  int64_t tot = 0;
  int *off = nullptr, *col = nullptr, *lo = nullptr, *up = nullptr;
  short *col16 = nullptr;
  // stencil form (no column arrays): slot k of row r at column
  // (r - r % sten_b) + sten[k]; sten on the device padded with 0 to the slot
  // count, sten_h the list on the host
  int *sten = nullptr;
  int sten_b = 1;
  std::vector<int> sten_h;
};
hipError_t sell_layout_build(int n, int64_t negm, const int *adr, const int *ci, hipStream_t s,
                             SellLayout *out);
// SELL values and diagonal from the reference-layout values a_ref
hipError_t sell_expand(int n, int64_t negm, const int *adr, const int *lo, const int *up,
                       const double *a_ref, double *sv, double *diag, hipStream_t s);

// element-wise helpers
hipError_t expand_ref(int64_t negm, const int *map_lo, const int *map_up, const double *a,
                      double *val, hipStream_t s);
hipError_t extract_diag(int n, const int *rowptr, const int *col, const double *val,
                        double *diag, float *diag32, hipStream_t s);
hipError_t to_f32(int64_t n, const double *a, float *b, double scale, hipStream_t s);
hipError_t axpy_f32(int64_t n, double alpha, const float *e, double *x, hipStream_t s);
hipError_t dot_partial(int64_t n, const double *a, const double *b, double *part, int nb,
                       hipStream_t s);
hipError_t fill(int64_t n, double v, double *x, hipStream_t s);
hipError_t reduce_to_host_slot(const double *part, int nb, double *out, hipStream_t s);

// assembly by entry (ssf_assemble.hip): the plan of a pattern and code-number table
// (cn[ne*nd] int32 on the device, the reference-layout pattern adr / ci int32 on
// the device): for every entry its contributions e nd^2 + i nd + j in the
// reference's localize order, cidx[cstart[s] .. cstart[s+1])
int asm_plan_build(int storage, int64_t n, int64_t negm, int64_t ne, int nd, const int *cn,
                   const int *adr, const int *ci, hipStream_t s, int **cstart, int **cidx,
                   int64_t *nc);
// The element-matrix buffer: P values per element (the full matrix, P = nd^2, or
// its lower triangle, P = nd (nd + 1) / 2 with (i, j), j <= i, at i (i + 1) / 2 + j)
// in groups of W = 2^wshift elements: value t of element e at
// (e / W) W P + t W + e % W. The one-thread-per-element kernels write W = 64 (their
// stores coalesce), the wavefront-per-element elasticity kernel and the caller's
// matrices of localize_batch W = 1.
inline int asm_wshift(int kind) { return kind == SSF_ELEM_HEX8_ELAST ? 0 : 6; }
inline int64_t asm_kbuf_count(int64_t ne, int nd, int wshift)
{
  const int64_t w = int64_t(1) << wshift;
  return (ne + w - 1) / w * w * ((int64_t)nd * (nd + 1) / 2);
}
// a[s] += its contributions from kbuf (lower triangles when sym, else full matrices)
hipError_t asm_gather(int64_t negm, int nd, bool sym, int wshift, const int *cstart,
                      const int *cidx, const double *kbuf, double *a, hipStream_t s);
// lift[d] += fbuf[lidx[k]] for k in [lstart[d], lstart[d+1])
hipError_t lift_gather(int64_t n, const int *lstart, const int *lidx, const double *fbuf,
                       double *lift, hipStream_t s);
void lift_plan_host(int64_t n, int64_t ne, int nd, const int64_t *cn, std::vector<int> &pidx,
                    std::vector<int> &lstart, std::vector<int> &lidx, int64_t *npe);
// device bytes of building an assembly plan and assembling once (peak)
inline int64_t asm_bytes(int64_t ne, int nd, int64_t negm)
{
  const int64_t ncmax = ne * nd * nd;
  return 8 * ne * nd + 20 * ncmax + 4 * (negm + 1) + 8 * asm_kbuf_count(ne, nd, 6) + 4 * ne * nd;
}
// value and vector updates rounded as the reference's loops (no contraction)
hipError_t vals_scale(int64_t n, double *a, double c, hipStream_t s);             // a *= c
hipError_t vals_add(int64_t n, double *a, double c, const double *b, hipStream_t s);  // a += c b
hipError_t vals_combine(int64_t n, double *out, const double *k, double sc, const double *cm,
                        hipStream_t s);                                            // out = k sc + 1.0 cm
hipError_t axpy_rn(int64_t n, double *y, const double *x, double c, const double *v,
                   hipStream_t s);                                                 // y = x + c v
hipError_t vec_sub(int64_t n, double *f, const double *a, const double *b, hipStream_t s);
hipError_t vec_sc2_add(int64_t n, double *f, const double *rhs, double alpha, double dt,
                       const double *p, hipStream_t s);                           // f = rhs alpha dt + p
hipError_t vec_rate(int64_t n, double *v, const double *x, const double *d, double dt,
                    double alpha, hipStream_t s);                                 // v = (x - d) / dt / alpha

// element matrices of every element as lower triangles into kbuf, and the lift
// values of the elements with prescribed values (pidx[e] >= 0) into fbuf (fbuf
// may be null)
hipError_t assemble_elements(int kind, int64_t ne, const int *conn, const int *cn,
                             const double *xyz, const double *mat_a, const double *mat_b,
                             const double *thick, double *kbuf, const int *pidx,
                             const double *pv, double *fbuf, hipStream_t s);

}  // namespace ssf
