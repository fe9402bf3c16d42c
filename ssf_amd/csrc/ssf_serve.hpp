// Served solves on gfx950: device views of the resident requests ("slots").
#pragma once

#include "ssf_device.hpp"

namespace ssf {

// One resident request: its own matrix (the SELL-64 slices of the gmatrix's full
// rows) and its own solve vectors, wherever they were allocated. Kernels reach
// them through this descriptor, so admitting or retiring a request moves no
// matrix data.
struct SlotDesc {
  const int *soff;      // slice offsets, nslice + 1
  const int *scol;      // 32-bit columns (used when scol16 and sten are null)
  const short *scol16;  // 16-bit offsets from the slice's first row
  // stencil form (when non-null): slot k of row r at column
  // (r - r % sten_b) + sten[k], sten padded with 0 to the slot count
  const int *sten;
  int sten_n, sten_b;
  const double *sval;
  // symmetric half of the stencil form (CsrDev::half_*; 13 lower slots from the
  // row or 13 lower neighbour nodes of 3 unknowns), or null: then the solve
  // streams the slices above. half_off: the lower offsets (stencil offsets from
  // the row, or node offsets), ascending
  const double *half_v;
  const float *half_v32;
  const int *half_off;
  int64_t half_pad;
  int half_b;
  const double *diag;
  const double *y;
  double *x, *r, *d, *p, *h;  // FP64 recurrence; the mixed form uses x, r only
  // mixed form: the FP32 operator float(S A S) with S = diag(isd),
  // isd = 1/sqrt|a_ii|, its CG recurrence (h = r), ||A||_inf of the FP64 values
  const float *sval32;
  const double *isd;
  float *x32, *r32, *d32, *p32;
  const double *norm_a;
  int n;
  int sid;         // index of the slot's scalars in SlotState (stable while resident)
  int blk0, nblk;  // the slot's blocks in the block table
  int fresh;       // admitted since the last step: initialise before iterating
};

// Per-slot scalars of the cg_prec recurrence (GEFEL/scr.cpp:1093-1203).
struct SlotState {
  double *nory, *nom, *ares;  // nom: 2 per slot (iteration parity)
  double *sc;                 // 3 per slot: d.p, r.r, r.h of the running iteration
  double *sc4;                // 4 per slot: d.p, r.r, r.p, p.p (the mixed form's fused step)
  long long *iter, *ani;
  int *done, *status, *pend, *pend_status, *brk;
  // mixed form (reliable updates, as the single-system solve): ||b||^2, ||b - A x||^2
  // at the last refresh and its best, scale of the FP32 quantities, delta^2,
  // ||x||^2 of the refresh; final: the solve has ended with mstatus
  double *nory64, *rr64, *best, *scale, *delta2, *xx;
  int *bad, *since_best, *final, *mstatus;
  long long *updates;
};

struct ServeDev {
  int nslot = 0, nblk = 0;
  const SlotDesc *slots = nullptr;
  const int *blk_slot = nullptr, *blk_s0 = nullptr, *blk_s1 = nullptr;  // local slices
  double *part = nullptr;  // 4 per block (2 for most kernels)
};

// 64-row slices of a slot per block (12 for a slot in the half form of 3-unknown
// nodes): the partition depends only on the slot
constexpr int SERVE_SLICES_PER_BLOCK = 16;

hipError_t serve_init(const ServeDev &S, SlotState st, double zero, hipStream_t s);
// ni: a slot stops with status 0 after its ni-th iteration
hipError_t serve_iteration(const ServeDev &S, SlotState st, double res, double zero,
                           long long ni, hipStream_t s);
// applies pending state; counts unfinished slots into *active (zeroed first)
hipError_t serve_flags(const ServeDev &S, SlotState st, int *active, hipStream_t s);

// mixed form: fresh slots start from x = 0 (FP64 residual y, FP32 recurrence on
// y / ||y||); an iteration advances the FP32 recurrences; a refresh folds the FP32
// corrections into the FP64 solutions, recomputes b - A x in FP64, tests it and
// restarts the FP32 recurrences on it (ni: FP32 iterations at most)
hipError_t serve_mixed_init(const ServeDev &S, SlotState st, double res, double delta,
                            double zero, hipStream_t s);
hipError_t serve_mixed_iteration(const ServeDev &S, SlotState st, hipStream_t s);
hipError_t serve_mixed_refresh(const ServeDev &S, SlotState st, double res, long long ni,
                               int *active, hipStream_t s);

}  // namespace ssf
