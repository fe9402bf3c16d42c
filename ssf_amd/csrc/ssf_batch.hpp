// Batched (many systems x R right-hand sides) Jacobi-PCG on gfx950: device views.
#pragma once

#include "ssf_device.hpp"

namespace ssf {

constexpr int PEND_ADVANCE = 1;  // iteration completed: iter += 1
constexpr int PEND_STOP = 2;     // recurrence stops (status in pend_status)

// block-diagonal matrix of the batch and its block table: block b covers rows
// [blk_r0[b], blk_r1[b]) of system blk_sys[b]; the blocks of system s are
// [sys_blk0[s], sys_blk0[s+1])
struct BatchDev {
  int nsys = 0;
  int R = 1;
  int nblk = 0;
  int64_t nrows = 0;
  const int *rowptr = nullptr;  // global rows, global (block-diagonal) columns
  const int *col = nullptr;
  const int *blk_sys = nullptr, *blk_r0 = nullptr, *blk_r1 = nullptr, *sys_blk0 = nullptr;
  // SELL-64 slices of the systems, concatenated: slice g covers global rows
  // slice_row0[g] .. +63 of one system; block b covers slices [blk_s0[b], blk_s1[b])
  const int *soff = nullptr;        // entry offsets, total slices + 1
  const int *scol = nullptr;        // global 32-bit columns (when scol16 is null)
  const short *scol16 = nullptr;    // offsets from the slice's first row
  const int *slice_row0 = nullptr;
  const int *blk_s0 = nullptr, *blk_s1 = nullptr;
  // work list of the persistent kernels: the blocks of the systems still
  // running, wl[0, *nwl) (rebuilt once per iteration chunk); null: every block.
  // Block-table kernels launch min(nblk, pgrid) workgroups that stride over it.
  const int *wl = nullptr;
  const int *nwl = nullptr;
  int pgrid = 2048;
};

struct BatchTune {
  int format = 2;  // 0: CSR row-vector SpMM, 2: SELL-64 (one lane per row, R sums)
  int lanes = 8;   // format 0: lanes per row (R * L)
};

// per (system, rhs) pair q = s * R + r
template <typename T>
struct BatchState {
  T *nory = nullptr;
  T *nom = nullptr;  // 2 per pair, ping-pong by iteration parity
  long long *iter = nullptr;
  int *done = nullptr, *status = nullptr, *pend = nullptr, *pend_status = nullptr,
      *brk = nullptr;
  long long *ani = nullptr;
  T *ares = nullptr;
  int *sys_done = nullptr;  // per system
};

template <typename T>
struct BatchVec {  // interleaved [row * R + r]
  const T *val = nullptr, *diag = nullptr;
  const T *sval = nullptr;  // SELL values
  T *x = nullptr, *r = nullptr, *d = nullptr, *p = nullptr, *h = nullptr, *y = nullptr;
  T *part = nullptr;  // 3 * nblk * R (FP64), 4 * nblk * R (FP32)
};

struct BatchMixDev {
  double *nory64 = nullptr, *rr64 = nullptr, *best = nullptr, *scale = nullptr,
         *delta2 = nullptr;
  double *norm_a = nullptr;  // per system: ||A||_inf
  int *bad = nullptr, *since_best = nullptr, *final = nullptr, *status = nullptr;
  long long *updates = nullptr;
  int *sys_final = nullptr;
  int *active = nullptr;
  // per row: 1/sqrt|a_ii|, the scaling S of the FP32 operator float(S A S)
  double *isd = nullptr;
};

// Y = A X (kind 0), with per-(block, rhs) partials of d.p (kind 1), or
// Y = B - A X with partials of |y|^2 (kind 2); val: CSR values (format 0),
// sval: SELL values (format 2)
template <typename T>
hipError_t bspmm(const BatchTune &t, int kind, const BatchDev &B, const T *val, const T *sval,
                 const T *x, T *y, const T *b, T *partial, const int *sys_done, hipStream_t s);
template <typename T>
hipError_t batch_init(const BatchDev &B, const BatchVec<T> &v, BatchState<T> st, T zero,
                      hipStream_t s);
template <typename T>
hipError_t batch_iteration(const BatchTune &t, const BatchDev &B, const BatchVec<T> &v,
                           BatchState<T> st, T res, T zero, bool mixed, hipStream_t s);
template <typename T>
hipError_t batch_flags(const BatchDev &B, BatchState<T> st, int *active, hipStream_t s);
// per system ||A||_inf of the FP64 values into norm_a[nsys]; blkmax: nblk doubles
hipError_t batch_norm_inf(const BatchDev &B, const double *val, double *blkmax, double *norm_a,
                          hipStream_t s);
hipError_t offset_copy(int64_t n, const int *src, int *dst, int add, hipStream_t s);
// work list of the blocks of the systems with skip[s] == 0 into wl / *nwl; with
// hist, also hist[*ctr] = *active (the unfinished pairs counted before) and
// *ctr += 1, the per-chunk record of a solve replayed from one graph
hipError_t batch_compact(const BatchDev &B, const int *skip, int *wl, int *nwl, const int *active,
                         int *hist, int *ctr, int hist_cap, hipStream_t s);
hipError_t zero_int(int *p, hipStream_t s);
// the FP32 SELL values float(v_ij isd_i isd_j) of the nslice concatenated slices
hipError_t batch_sell_scale32(const BatchDev &B, int nslice, const double *v, const double *isd,
                              float *v32, hipStream_t s);
hipError_t batch_mixed_start(const BatchTune &t64, const BatchTune &t32, const BatchDev &B,
                             const BatchVec<double> &v64, const BatchVec<float> &v32,
                             BatchState<float> st, const BatchMixDev &md, double res,
                             double delta, hipStream_t s);
hipError_t batch_mixed_refresh(const BatchTune &t64, const BatchDev &B,
                               const BatchVec<double> &v64, const BatchVec<float> &v32,
                               BatchState<float> st, const BatchMixDev &md, double res,
                               hipStream_t s);

}  // namespace ssf
