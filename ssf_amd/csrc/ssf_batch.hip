// Batched Jacobi-PCG over many independent systems with R right-hand sides each
// (gfx950). The systems are concatenated block-diagonally; vectors are stored
// interleaved, v[row * R + r], so the SpMM gathers one contiguous R-vector per
// column. Every block owns a contiguous row range of exactly one system, so every
// scalar of a system is reduced from that system's blocks only, in a fixed order:
// a system's result does not depend on the other systems in the batch.
//
// Per iteration (for each active (system, rhs) pair, the recurrence of
// GEFEL/scr.cpp:1142-1197):
//   B1  P = A D, partials d.p              (k_bspmm, kind 1)
//   B2  x += a d, r -= a p, h = r / diag, partials r.r, r.h   (k_bupdate)
//   B3  test, d = h + b d                   (k_bdir)
// Mixed precision runs the FP32 recurrence as CG on the symmetrically scaled
// operator float(S A S), S = diag(1/sqrt|a_ii|) (h = r, no diagonal), and adds,
// once per chunk, the reliable update of ssf_solver.hip per pair:
// x64 += S x32 / s, r64 = b - A x64 in FP64, FP32 residual refresh. On SELL
// slices its iteration is two passes: the SpMM with the partials of d.p, r.r,
// r.p and p.p (k_bsell, kind 3), then x, r and d in one step (k_bstep).
#include "ssf_batch.hpp"

namespace ssf {

// ---------------------------------------------------------- work list

// Block-table kernels are persistent: min(nblk, pgrid) workgroups stride over
// the work list (the blocks of the systems still running), so systems that have
// finished cost nothing and a launch covers any number of systems.
__device__ __forceinline__ int wl_count(const BatchDev &B) { return B.nwl ? *B.nwl : B.nblk; }
__device__ __forceinline__ int wl_block(const BatchDev &B, int i) { return B.nwl ? B.wl[i] : i; }
static inline dim3 wl_grid(const BatchDev &B) { return dim3((unsigned)std::max(1, std::min(B.nblk, B.pgrid))); }

// ------------------------------------------------------------ reductions

// SELL-64 SpMM: one wavefront per slice, one lane per row holding the R sums of
// its row; each entry is loaded once (SELL_V per load) and its column's R
// values of x are read contiguously. Rows past the block's end are padding.
template <typename T, int R, bool I16, int KIND>
__global__ __launch_bounds__(TPB) void k_bsell(BatchDev B, const T *__restrict__ val,
                                               const T *__restrict__ x, T *__restrict__ y,
                                               const T *__restrict__ b,
                                               T *__restrict__ partial,
                                               const int *__restrict__ sys_done)
{
  typedef T tv __attribute__((ext_vector_type(SELL_V)));
  typedef int iv __attribute__((ext_vector_type(SELL_V)));
  typedef short sv __attribute__((ext_vector_type(SELL_V)));
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (sys_done && sys_done[s]) continue;
  const int rend = B.blk_r1[blk];
  // KIND 3 (fused FP32 step, b = r): d.p, r.r, r.p, p.p per right-hand side
  constexpr int NA = KIND == 3 ? 4 : 1;
  T acc[NA][R];
#pragma unroll
  for (int k = 0; k < NA; k++)
#pragma unroll
    for (int r = 0; r < R; r++) acc[k][r] = T(0);
  for (int sl = B.blk_s0[blk] + wave; sl < B.blk_s1[blk]; sl += TPB / 64) {
    const int off = B.soff[sl];
    const int w = (B.soff[sl + 1] - off) / 64;
    const int row0 = B.slice_row0[sl];
    T sum[R];
#pragma unroll
    for (int r = 0; r < R; r++) sum[r] = T(0);
    for (int q = 0; q < w; q += SELL_V) {
      const int k = off + q * 64 + lane * SELL_V;
      const tv vv = __builtin_nontemporal_load((const tv *)(val + k));
      int cc[SELL_V];
      if (I16) {
        const sv c16 = __builtin_nontemporal_load((const sv *)(B.scol16 + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) cc[u] = row0 + c16[u];
      } else {
        const iv c32 = __builtin_nontemporal_load((const iv *)(B.scol + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) cc[u] = c32[u];
      }
#pragma unroll
      for (int u = 0; u < SELL_V; u++) {
        const T *__restrict__ xr = x + (int64_t)cc[u] * R;
#pragma unroll
        for (int r = 0; r < R; r++) sum[r] += vv[u] * xr[r];
      }
    }
    const int row = row0 + lane;
    if (row < rend) {
      const int64_t q0 = (int64_t)row * R;
#pragma unroll
      for (int r = 0; r < R; r++) {
        T v = sum[r];
        if (KIND == 3) {
          const T rv = b[q0 + r];
          acc[0][r] += x[q0 + r] * v;
          acc[1][r] += rv * rv;
          acc[2][r] += rv * v;
          acc[3][r] += v * v;
        } else if (KIND == 2) {
          v = b[q0 + r] - v;
          acc[0][r] += v * v;
        } else if (KIND == 1) {
          acc[0][r] += x[q0 + r] * v;
        }
        y[q0 + r] = v;
      }
    }
  }
  if (KIND == 3) {
    // the 4 R sums in one pass: wave sums, then wave 0's lanes over the waves
    __shared__ T red[TPB / 64][4 * R];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int r = 0; r < R; r++) {
        const T t = wave_sum(acc[k][r]);
        if (lane == 0) red[wave][k * R + r] = t;
      }
    __syncthreads();
    if (threadIdx.x < 4 * R) {
      T t = red[0][threadIdx.x];
#pragma unroll
      for (int w = 1; w < TPB / 64; w++) t += red[w][threadIdx.x];
      const int k = threadIdx.x / R, r = threadIdx.x % R;
      partial[((int64_t)k * B.nblk + blk) * R + r] = t;
    }
    __syncthreads();  // red is rewritten for the next block
  } else if (KIND != 0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const T t = block_sum(acc[0][r]);
      if (threadIdx.x == 0) partial[(int64_t)blk * R + r] = t;
    }
  }
  }
}

// sum of a per-system quantity over the blocks of system s (fixed order)
template <typename T, int R>
__device__ __forceinline__ T sys_sum(const BatchDev &B, const T *__restrict__ part, int s, int r)
{
  const int b0 = B.sys_blk0[s], b1 = B.sys_blk0[s + 1];
  T t = T(0);
  for (int b = b0; b < b1; b++) t += part[(int64_t)b * R + r];
  return t;
}

// block-wide per-rhs reduction: thread t holds a value for rhs t % R; a fixed
// tree over LDS (pairs t, t + w share their rhs), result in out_lds[0, R)
template <typename T, int R>
__device__ __forceinline__ void block_sum_r(T v, T *out_lds)
{
  __shared__ T red[TPB];
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = TPB / 2; w >= R; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < R) out_lds[threadIdx.x] = red[threadIdx.x];
  __syncthreads();
}

// Sum over the blocks of system s of part[b * R + r], for every r, by the whole
// block: thread (j, r) adds the system's blocks j, j + TPB/R, ... (counted from
// the system's first block), then the fixed tree. The order depends only on the
// system's own block count, not on its place in the batch.
template <typename T, int R>
__device__ __forceinline__ void sys_sum_block(const BatchDev &B, const T *__restrict__ part,
                                              int s, T *out_lds)
{
  const int b0 = B.sys_blk0[s], b1 = B.sys_blk0[s + 1];
  constexpr int NJ = TPB / R;
  const int r = threadIdx.x % R;
  T t = T(0);
  for (int b = b0 + (int)threadIdx.x / R; b < b1; b += NJ) t += part[(int64_t)b * R + r];
  block_sum_r<T, R>(t, out_lds);
}

// ----------------------------------------------------------------- SpMM

// G = R * L lanes per row: lane (l, r) sums entries l, l + L, ... of the row for
// right-hand side r; the L partial sums are combined with shuffles (fixed order).
template <typename T, int R, int L, int KIND>
__global__ __launch_bounds__(TPB) void k_bspmm(BatchDev B, const T *__restrict__ val,
                                               const T *__restrict__ x, T *__restrict__ y,
                                               const T *__restrict__ b,
                                               T *__restrict__ partial,
                                               const int *__restrict__ sys_done)
{
  constexpr int G = R * L;
  constexpr int NG = TPB / G;
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  const int r = lane % R, l = lane / R;
  __shared__ T outp[R];
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (sys_done && sys_done[s]) continue;
  const int r0 = B.blk_r0[blk], r1 = B.blk_r1[blk];
  T acc = T(0);
  for (int row = r0 + grp; row < r1; row += NG) {
    const int beg = B.rowptr[row], end = B.rowptr[row + 1];
    T sum = T(0);
    for (int k = beg + l; k < end; k += L) sum += val[k] * x[(int64_t)B.col[k] * R + r];
#pragma unroll
    for (int off = G / 2; off >= R; off >>= 1) sum += __shfl_xor(sum, off, G);
    if (l == 0) {
      const int64_t q = (int64_t)row * R + r;
      if (KIND == 2) {
        sum = b[q] - sum;
        acc += sum * sum;
      } else if (KIND == 1) {
        acc += x[q] * sum;
      }
      y[q] = sum;
    }
  }
  if (KIND != 0) {
    // per right-hand side: thread t holds a partial sum for rhs t % R
    block_sum_r<T, R>(acc, outp);
    if (threadIdx.x < R) partial[(int64_t)blk * R + threadIdx.x] = outp[threadIdx.x];
  }
  }
}

// ------------------------------------------------------------- B2: update

template <typename T, int R, bool JAC>
__global__ __launch_bounds__(TPB) void k_bupdate(BatchDev B, BatchState<T> st,
                                                 const T *__restrict__ pden, T zero,
                                                 T *__restrict__ x, T *__restrict__ rr_,
                                                 const T *__restrict__ d,
                                                 const T *__restrict__ p,
                                                 const T *__restrict__ diag,
                                                 T *__restrict__ h, T *__restrict__ prr)
{
  __shared__ T alpha[R];
  __shared__ T outa[R], outb[R], dens[R];
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (st.sys_done[s]) continue;
  sys_sum_block<T, R>(B, pden, s, dens);
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    const int q = s * R + r;
    T a = T(0);
    if (!st.done[q]) {
      const T den = dens[r];
      const T nom = st.nom[2 * q + (st.iter[q] & 1)];
      if (fabs(den) < zero) {
        a = T(0);
        if (blk == B.sys_blk0[s]) {  // first block records the stop, applied by k_badvance
          st.brk[q] = 1;
          st.ani[q] = st.iter[q];
          st.ares[q] = nom;
        }
      } else {
        a = nom / den;
      }
    }
    alpha[r] = a;
  }
  __syncthreads();
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  const int r = threadIdx.x % R;  // TPB is a multiple of R
  const T a = alpha[r];
  T srr = T(0), srh = T(0);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) {
    if (a != T(0)) {
      x[e] += a * d[e];
      rr_[e] -= a * p[e];
    }
    const T ri = rr_[e];
    srr += ri * ri;
    if (JAC) {
      const T hi = ri / diag[e / R];
      h[e] = hi;
      srh += ri * hi;
    }
  }
  block_sum_r<T, R>(srr, outa);
  if (JAC) block_sum_r<T, R>(srh, outb);
  else if (threadIdx.x < R) outb[threadIdx.x] = outa[threadIdx.x];
  if (threadIdx.x < R) {
    prr[(int64_t)blk * R + threadIdx.x] = outa[threadIdx.x];
    prr[(int64_t)(B.nblk + blk) * R + threadIdx.x] = outb[threadIdx.x];
  }
  }
}

// ------------------------------------------- B2 + B3 fused (mixed, FP32)

// One step of the FP32 recurrence after P = A D with the four dots (k_bsell,
// KIND 3): alpha = r.r / d.p, x += alpha d, r -= alpha p, and with ||r - alpha
// p||^2 = r.r - 2 alpha r.p + alpha^2 p.p, beta and d = r + beta d in the same
// pass, as k_cg_fused_step does for one system. The new r.r < nory pauses the
// pair until the refresh; state changes are left pending for k_badvance.
template <typename T, int R>
__global__ __launch_bounds__(TPB) void k_bstep(BatchDev B, BatchState<T> st,
                                               const T *__restrict__ part, T zero,
                                               T *__restrict__ x, T *__restrict__ rv,
                                               T *__restrict__ d, const T *__restrict__ p)
{
  __shared__ T dots[4][R];
  __shared__ T alpha[R], beta[R];
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (st.sys_done[s]) continue;
  const bool first = blk == B.sys_blk0[s];
#pragma unroll
  for (int k = 0; k < 4; k++) sys_sum_block<T, R>(B, part + (int64_t)k * B.nblk * R, s, dots[k]);
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    const int q = s * R + r;
    T a = T(0), bt = T(0);
    if (!st.done[q] && !st.brk[q]) {
      const T dp = dots[0][r], rr = dots[1][r], rp = dots[2][r], pp = dots[3][r];
      const long long it = st.iter[q];
      if (fabs(dp) < zero) {
        if (first) {  // |d.Ap| < zero: the pair stops, applied by k_badvance
          st.brk[q] = 1;
          st.ani[q] = it;
          st.ares[q] = rr;
        }
      } else {
        a = rr / dp;
        const T rrn = rr - T(2) * a * rp + a * a * pp;
        bt = rrn / rr;
        if (first) {
          st.nom[2 * q + ((it + 1) & 1)] = rrn;
          st.pend[q] = PEND_ADVANCE | (rrn < st.nory[q] ? PEND_STOP : 0);
          st.pend_status[q] = 1;
        }
      }
    }
    alpha[r] = a;
    beta[r] = bt;
  }
  __syncthreads();
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  const int r = threadIdx.x % R;  // TPB is a multiple of R
  const T a = alpha[r], bt = beta[r];
  if (a != T(0))
    for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) {
      x[e] += a * d[e];
      const T ri = rv[e] - a * p[e];
      rv[e] = ri;
      d[e] = ri + bt * d[e];
    }
  __syncthreads();  // alpha / beta are rewritten for the next block
  }
}

// -------------------------------------------------------------- B3: direction

template <typename T, int R, bool MIXED>
__global__ __launch_bounds__(TPB) void k_bdir(BatchDev B, BatchState<T> st,
                                              const T *__restrict__ prr, T res,
                                              T *__restrict__ d, const T *__restrict__ h)
{
  __shared__ T beta[R];
  __shared__ int upd[R];
  __shared__ T rrs[R], noms[R];
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (st.sys_done[s]) continue;
  const bool first = blk == B.sys_blk0[s];
  sys_sum_block<T, R>(B, prr, s, rrs);
  sys_sum_block<T, R>(B, prr + (int64_t)B.nblk * R, s, noms);
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    const int q = s * R + r;
    int u = 0;
    T bt = T(0);
    // state changes are left pending for k_badvance: every block of the system
    // reads the same iter / done / nom while this kernel runs
    if (!st.done[q] && !st.brk[q]) {
      const T rr = rrs[r];
      const T nomn = noms[r];
      const long long it = st.iter[q];
      const T nom = st.nom[2 * q + (it & 1)];
      bt = nomn / nom;
      if (MIXED) {
        u = 1;  // direction always updated; rr < nory pauses until the refresh
        if (first) {
          st.nom[2 * q + ((it + 1) & 1)] = nomn;
          st.pend[q] = PEND_ADVANCE | (rr < st.nory[q] ? PEND_STOP : 0);
          st.pend_status[q] = 1;
        }
      } else {
        const bool stop = fabs(rr / st.nory[q]) < res;  // GEFEL/scr.cpp:1188
        u = stop ? 0 : 1;
        if (first) {
          st.nom[2 * q + ((it + 1) & 1)] = nomn;
          if (stop) {
            st.ani[q] = it;
            st.ares[q] = nomn / st.nory[q];
            st.pend[q] = PEND_STOP;
            st.pend_status[q] = 1;
          } else {
            st.pend[q] = PEND_ADVANCE;
          }
        }
      }
    }
    beta[r] = bt;
    upd[r] = u;
  }
  __syncthreads();
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  const int r = threadIdx.x % R;
  if (upd[r]) {
    const T bt = beta[r];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) d[e] = bt * d[e] + h[e];
  }
  __syncthreads();  // beta / upd are rewritten for the next block
  }
}

// applies the pending state changes of the iteration, then: a system is done when
// all its right-hand sides are; counts the active pairs for the host
template <typename T, int R>
__global__ void k_badvance(BatchDev B, BatchState<T> st, int *active)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.nsys) return;
  int a = 0;
  for (int r = 0; r < R; r++) {
    const int q = s * R + r;
    if (st.brk[q]) {  // |d.Ap| < zero in k_bupdate (GEFEL/scr.cpp:1148-1152)
      st.done[q] = 1;
      st.status[q] = 2;
      st.brk[q] = 0;
      st.pend[q] = 0;
    }
    const int pd = st.pend[q];
    if (pd) {
      if (pd & PEND_ADVANCE) st.iter[q] += 1;
      if (pd & PEND_STOP) {
        st.done[q] = 1;
        st.status[q] = st.pend_status[q];
      }
      st.pend[q] = 0;
    }
    a += st.done[q] ? 0 : 1;
  }
  st.sys_done[s] = a == 0;
  if (a && active) atomicAdd(active, a);
}

// --------------------------------------------------------------- init

// r = y (x = 0), h = r / diag (JAC) or r, d = h, partials y.y, r.h
template <typename T, int R, bool JAC>
__global__ __launch_bounds__(TPB) void k_binit(BatchDev B, const T *__restrict__ y,
                                               const T *__restrict__ diag, T *__restrict__ r_,
                                               T *__restrict__ h, T *__restrict__ d,
                                               T *__restrict__ prr)
{
  __shared__ T outa[R], outb[R];
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  T syy = T(0), srh = T(0);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) {
    const T yi = y[e];
    const T hi = JAC ? yi / diag[e / R] : yi;
    r_[e] = yi;
    if (JAC) h[e] = hi;
    d[e] = hi;
    syy += yi * yi;
    srh += yi * hi;
  }
  block_sum_r<T, R>(syy, outa);
  block_sum_r<T, R>(srh, outb);
  if (threadIdx.x < R) {
    prr[(int64_t)blk * R + threadIdx.x] = outa[threadIdx.x];
    prr[(int64_t)(B.nblk + blk) * R + threadIdx.x] = outb[threadIdx.x];
  }
  }
}

template <typename T, int R>
__global__ void k_binit2(BatchDev B, BatchState<T> st, const T *__restrict__ prr, T zero)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.nsys * R) return;
  const int s = q / R, r = q % R;
  const T nory = sys_sum<T, R>(B, prr, s, r);
  const T nom = sys_sum<T, R>(B, prr + (int64_t)B.nblk * R, s, r);
  st.nory[q] = nory;
  st.nom[2 * q] = nom;
  st.nom[2 * q + 1] = nom;
  st.iter[q] = 0;
  st.status[q] = 0;
  st.done[q] = 0;
  st.pend[q] = 0;
  st.brk[q] = 0;
  st.ani[q] = -1;
  st.ares[q] = T(0);
  if (nory < zero) {  // zero right-hand side: x = 0
    st.status[q] = 3;
    st.done[q] = 1;
    st.ani[q] = 0;
    st.ares[q] = nory;
  }
  if (r == 0) st.sys_done[s] = 0;
}

// ------------------------------------------------- mixed: reliable updates

struct BatchMix {
  double *nory64, *rr64, *best, *scale, *delta2, *norm_a;
  int *bad, *since_best, *final, *status;
  long long *updates;
};

// x64 += S x32 / s, x32 = 0 for the unfinished pairs; block partials of ||x64||^2
// per rhs into px[blk * R + r]
template <int R>
__global__ __launch_bounds__(TPB) void k_bmx_xupdate(BatchDev B, BatchMix mx,
                                                     const int *__restrict__ sys_final,
                                                     const double *__restrict__ isd,
                                                     double *__restrict__ x64,
                                                     float *__restrict__ x32,
                                                     double *__restrict__ px)
{
  __shared__ double outa[R];
  const int r = threadIdx.x % R;
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (sys_final[s]) continue;
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  const int q = s * R + r;
  const bool fin = mx.final[q];
  const double inv = 1.0 / mx.scale[q];
  double xx = 0.0;
  if (!fin) {
    for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) {
      const double xi = x64[e] + (double)x32[e] * inv * isd[e / R];
      x64[e] = xi;
      x32[e] = 0.0f;
      xx += xi * xi;
    }
  }
  block_sum_r<double, R>(xx, outa);
  if (threadIdx.x < R) px[(int64_t)blk * R + threadIdx.x] = outa[threadIdx.x];
  }
}

template <int R>
__global__ __launch_bounds__(TPB) void k_bmx_refresh(BatchDev B, BatchMix mx,
                                                     const int *__restrict__ sys_final,
                                                     const double *__restrict__ part64,
                                                     const double *__restrict__ r64,
                                                     const double *__restrict__ isd,
                                                     float *__restrict__ r32,
                                                     float *__restrict__ d32,
                                                     float *__restrict__ prr)
{
  __shared__ double scl[R];
  __shared__ float cfac[R];
  __shared__ int rst[R], fin[R];
  __shared__ float outa[R];
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
  const int blk = wl_block(B, wi);
  const int s = B.blk_sys[blk];
  if (sys_final[s]) continue;
  __syncthreads();  // the previous block's reads of scl / cfac / rst / fin are done
  if (threadIdx.x < R) {
    const int q = s * R + threadIdx.x;
    const double rr64 = sys_sum<double, R>(B, part64, s, threadIdx.x);
    const double old = mx.scale[q];
    const double sc = rr64 > 0.0 ? 1.0 / sqrt(rr64) : old;
    scl[threadIdx.x] = sc;
    cfac[threadIdx.x] = (float)(sc / old);
    rst[threadIdx.x] = rr64 > 1e4 * mx.best[q];
    fin[threadIdx.x] = mx.final[q];
  }
  __syncthreads();
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  const int r = threadIdx.x % R;
  float srr = 0.f;
  if (!fin[r]) {
    const double sc = scl[r];
    const float c = cfac[r];
    const bool restart = rst[r];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) {
      const float ri = (float)(r64[e] * sc * isd[e / R]);
      r32[e] = ri;
      d32[e] = restart ? ri : d32[e] * c;
      srr += ri * ri;
    }
  }
  block_sum_r<float, R>(srr, outa);
  if (threadIdx.x < R) {
    prr[(int64_t)blk * R + threadIdx.x] = outa[threadIdx.x];
    prr[(int64_t)(B.nblk + blk) * R + threadIdx.x] = outa[threadIdx.x];
  }
  }
}

// below this the FP64 residual is within the rounding error of its own
// computation: at most m eps (|b| + |A||x|) per row for m entries per row,
// typically sqrt(m) eps; 64 eps (||b|| + ||A||_inf ||x||) in 2-norm
constexpr double BMIX_FLOOR = 64.0 * 2.220446049250313e-16;

template <int R>
__global__ void k_bmx_check(BatchDev B, BatchMix mx, BatchState<float> st,
                            const double *__restrict__ part64, const float *__restrict__ prr,
                            const double *__restrict__ px, double res)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.nsys * R) return;
  if (mx.final[q]) return;
  const int s = q / R, r = q % R;
  const double rr64 = sys_sum<double, R>(B, part64, s, r);
  const float nomn = sys_sum<float, R>(B, prr + (int64_t)B.nblk * R, s, r);
  const double prev = mx.rr64[q];
  mx.rr64[q] = rr64;
  mx.updates[q] += 1;
  if (rr64 < res * mx.nory64[q]) {
    mx.final[q] = 1;
    mx.status[q] = 1;
    st.done[q] = 1;
    return;
  }
  const bool met = st.done[q] && st.status[q] == 1;
  mx.bad[q] = (met && rr64 > 0.25 * prev) ? mx.bad[q] + 1 : 0;
  if (rr64 < 0.81 * mx.best[q]) mx.since_best[q] = 0;
  else mx.since_best[q] += 1;
  if (rr64 < mx.best[q]) mx.best[q] = rr64;
  // stuck at the FP64 floor: within the rounding floor and no longer improving
  // (far above it the CG residual norm may stall or grow for many refreshes)
  const double xx = sys_sum<double, R>(B, px, s, r);
  const double fl = BMIX_FLOOR * (sqrt(mx.nory64[q]) + mx.norm_a[s] * sqrt(xx));
  if (rr64 < fl * fl && (mx.bad[q] >= 3 || mx.since_best[q] >= 10)) {
    mx.final[q] = 1;
    mx.status[q] = 4;
    st.done[q] = 1;
    return;
  }
  mx.scale[q] = rr64 > 0.0 ? 1.0 / sqrt(rr64) : mx.scale[q];
  // next refresh once the FP32 residual falls by delta from its value now, or
  // to the final target if that is closer
  st.nory[q] = (float)((double)nomn * fmax(mx.delta2[q], res * mx.nory64[q] / rr64));
  st.nom[2 * q + (st.iter[q] & 1)] = nomn;
  st.done[q] = 0;
  st.status[q] = 0;
  st.pend[q] = 0;
  st.brk[q] = 0;
}

// system-level final flags for the mixed solve; counts the unfinished pairs
template <int R>
__global__ void k_bmx_flags(BatchDev B, BatchMix mx, BatchState<float> st, int *sys_final,
                            int *active)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.nsys) return;
  int a = 0, run = 0;
  for (int r = 0; r < R; r++) {
    a += mx.final[s * R + r] ? 0 : 1;
    run += (mx.final[s * R + r] || st.done[s * R + r]) ? 0 : 1;
  }
  if (a == 0) sys_final[s] = 1;
  st.sys_done[s] = run == 0;  // the FP32 kernels skip a system with nothing to iterate
  if (a) atomicAdd(active, a);
}

template <int R>
__global__ void k_bmx_init(BatchDev B, BatchMix mx, BatchState<float> st,
                           const double *__restrict__ part64, const double *__restrict__ pyy,
                           double res, double delta, int *sys_final)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.nsys * R) return;
  const int s = q / R, r = q % R;
  const double rr = sys_sum<double, R>(B, part64, s, r);
  const double yy = sys_sum<double, R>(B, pyy, s, r);
  mx.nory64[q] = yy;
  mx.rr64[q] = rr;
  mx.best[q] = rr;
  mx.scale[q] = rr > 0.0 ? 1.0 / sqrt(rr) : 1.0;
  mx.delta2[q] = delta * delta;
  mx.bad[q] = 0;
  mx.since_best[q] = 0;
  mx.updates[q] = 0;
  mx.final[q] = 0;
  mx.status[q] = 0;
  if (r == 0) sys_final[s] = 0;
  if (yy < 1e-300) {  // zero right-hand side: x = 0 stays
    mx.final[q] = 1;
    mx.status[q] = 3;
  } else if (rr < res * yy) {
    mx.final[q] = 1;
    mx.status[q] = 1;
  }
}

// residual of the FP64 refresh in the units of the FP32 recurrence: s S r64
template <int R>
__global__ __launch_bounds__(TPB) void k_bmx_scale_rhs(BatchDev B, BatchMix mx,
                                                       const double *__restrict__ r64,
                                                       const double *__restrict__ isd,
                                                       float *__restrict__ y32)
{
  const int r = threadIdx.x % R;
  const int wn = wl_count(B);
  for (int wi = blockIdx.x; wi < wn; wi += gridDim.x) {
    const int blk = wl_block(B, wi);
    const int s = B.blk_sys[blk];
    const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
    const double sc = mx.scale[s * R + r];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) y32[e] = (float)(r64[e] * sc * isd[e / R]);
  }
}

// after k_binit2 of the FP32 recurrence (st.nory = y32.y32 of its start)
template <int R>
__global__ void k_bmx_init32(BatchDev B, BatchMix mx, BatchState<float> st, double res)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.nsys * R) return;
  const double rr = mx.rr64[q];
  st.nory[q] = rr > 0.0 ? (float)((double)st.nory[q] * fmax(mx.delta2[q], res * mx.nory64[q] / rr))
                        : 0.f;
  if (mx.final[q]) st.done[q] = 1;
  st.status[q] = 0;
}

// the scaled FP32 SELL values of the mixed solve: v32 = float(v isd_row isd_col),
// one thread per layout row (rows past a system's end hold padding, value 0)
__global__ void k_bsell_scale32(BatchDev B, int nslice, const double *__restrict__ v,
                                const double *__restrict__ isd, float *__restrict__ v32)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslice * 64) return;
  const int g = t / 64, lane = t % 64;
  const int off = B.soff[g];
  const int w = (B.soff[g + 1] - off) / 64;
  const int64_t last = B.nrows - 1;
  const int row0 = B.slice_row0[g];
  const double si = isd[min((int64_t)row0 + lane, last)];
  for (int q = 0; q < w; q++) {
    const int k = off + (q / SELL_V) * 64 * SELL_V + lane * SELL_V + q % SELL_V;
    int64_t c = B.scol16 ? (int64_t)row0 + B.scol16[k] : (int64_t)B.scol[k];
    c = c < 0 ? 0 : (c > last ? last : c);
    v32[k] = (float)(v[k] * si * isd[c]);
  }
}

hipError_t batch_sell_scale32(const BatchDev &B, int nslice, const double *v, const double *isd,
                              float *v32, hipStream_t s)
{
  if (nslice <= 0 || B.nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bsell_scale32, dim3((nslice * 64 + TPB - 1) / TPB), dim3(TPB), 0, s, B,
                     nslice, v, isd, v32);
  return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_bdot_yy(BatchDev B, int R, const double *__restrict__ y,
                                                 double *__restrict__ part)
{
  // per block and rhs: sum y^2 (R <= 8)
  __shared__ double red[TPB];
  const int blk = blockIdx.x;
  const int64_t e0 = (int64_t)B.blk_r0[blk] * R, e1 = (int64_t)B.blk_r1[blk] * R;
  double a = 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += TPB) a += y[e] * y[e];
  red[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x < R) {
    double t = 0.0;
    for (int i = threadIdx.x; i < TPB; i += R) t += red[i];
    part[(int64_t)blk * R + threadIdx.x] = t;
  }
}

// ----------------------------------------------------------- dispatch

template <typename T, int R>
static hipError_t bspmm_r(const BatchTune &t, int kind, const BatchDev &B, const T *val,
                          const T *sval, const T *x, T *y, const T *b, T *partial,
                          const int *sys_done, hipStream_t s)
{
  dim3 grid = wl_grid(B), block(TPB);
  if (t.format == 2 && B.soff && sval) {
#define SSF_BSELL(I, K) hipLaunchKernelGGL((k_bsell<T, R, I, K>), grid, block, 0, s, B, sval, x, y, b, partial, sys_done)
    const bool i16 = B.scol16 != nullptr;
    if (kind == 0) { if (i16) SSF_BSELL(true, 0); else SSF_BSELL(false, 0); }
    else if (kind == 1) { if (i16) SSF_BSELL(true, 1); else SSF_BSELL(false, 1); }
    else if (kind == 2) { if (i16) SSF_BSELL(true, 2); else SSF_BSELL(false, 2); }
    else if (i16) SSF_BSELL(true, 3);
    else SSF_BSELL(false, 3);
#undef SSF_BSELL
    return hipGetLastError();
  }
#define SSF_BS(L, K) hipLaunchKernelGGL((k_bspmm<T, R, L, K>), grid, block, 0, s, B, val, x, y, b, partial, sys_done)
#define SSF_BSK(L)              \
  do {                          \
    if (kind == 0) SSF_BS(L, 0); \
    else if (kind == 1) SSF_BS(L, 1); \
    else SSF_BS(L, 2);          \
  } while (0)
  const int L = t.lanes / R;
  if (R * 16 <= 64 && L >= 16) SSF_BSK((R * 16 <= 64 ? 16 : 1));
  else if (R * 8 <= 64 && L >= 8) SSF_BSK((R * 8 <= 64 ? 8 : 1));
  else if (R * 4 <= 64 && L >= 4) SSF_BSK((R * 4 <= 64 ? 4 : 1));
  else if (R * 2 <= 64 && L >= 2) SSF_BSK((R * 2 <= 64 ? 2 : 1));
  else SSF_BSK(1);
#undef SSF_BSK
#undef SSF_BS
  return hipGetLastError();
}

template <typename T>
hipError_t bspmm(const BatchTune &t, int kind, const BatchDev &B, const T *val, const T *sval,
                 const T *x, T *y, const T *b, T *partial, const int *sys_done, hipStream_t s)
{
  switch (B.R) {
    case 1: return bspmm_r<T, 1>(t, kind, B, val, sval, x, y, b, partial, sys_done, s);
    case 2: return bspmm_r<T, 2>(t, kind, B, val, sval, x, y, b, partial, sys_done, s);
    case 4: return bspmm_r<T, 4>(t, kind, B, val, sval, x, y, b, partial, sys_done, s);
    default: return bspmm_r<T, 8>(t, kind, B, val, sval, x, y, b, partial, sys_done, s);
  }
}
template hipError_t bspmm<double>(const BatchTune &, int, const BatchDev &, const double *,
                                  const double *, const double *, double *, const double *,
                                  double *, const int *, hipStream_t);
template hipError_t bspmm<float>(const BatchTune &, int, const BatchDev &, const float *,
                                 const float *, const float *, float *, const float *, float *,
                                 const int *, hipStream_t);

template <typename T, int R>
static hipError_t biter_r(const BatchTune &t, const BatchDev &B, const BatchVec<T> &v,
                          BatchState<T> st, T res, T zero, bool mixed, hipStream_t s)
{
  dim3 grid = wl_grid(B), block(TPB);
  if (mixed && t.format == 2 && B.soff && v.sval) {
    // the FP32 recurrence in two passes: SpMM with the four dots, fused step
    hipError_t e = bspmm_r<T, R>(t, 3, B, v.val, v.sval, v.d, v.p, v.r, v.part, st.sys_done, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_bstep<T, R>), grid, block, 0, s, B, st, v.part, zero, v.x, v.r, v.d,
                       v.p);
    hipLaunchKernelGGL((k_badvance<T, R>), dim3((B.nsys + TPB - 1) / TPB), dim3(TPB), 0, s, B,
                       st, nullptr);
    return hipGetLastError();
  }
  hipError_t e = bspmm_r<T, R>(t, 1, B, v.val, v.sval, v.d, v.p, nullptr, v.part, st.sys_done, s);
  if (e != hipSuccess) return e;
  T *prr = v.part + (int64_t)B.nblk * R;
  if (mixed) {  // the FP32 recurrence on the scaled operator: h = r
    hipLaunchKernelGGL((k_bupdate<T, R, false>), grid, block, 0, s, B, st, v.part, zero, v.x, v.r,
                       v.d, v.p, nullptr, nullptr, prr);
    hipLaunchKernelGGL((k_bdir<T, R, true>), grid, block, 0, s, B, st, prr, res, v.d, v.r);
  } else {
    hipLaunchKernelGGL((k_bupdate<T, R, true>), grid, block, 0, s, B, st, v.part, zero, v.x, v.r,
                       v.d, v.p, v.diag, v.h, prr);
    hipLaunchKernelGGL((k_bdir<T, R, false>), grid, block, 0, s, B, st, prr, res, v.d, v.h);
  }
  hipLaunchKernelGGL((k_badvance<T, R>), dim3((B.nsys + TPB - 1) / TPB), dim3(TPB), 0, s, B, st,
                     nullptr);
  return hipGetLastError();
}

template <typename T>
hipError_t batch_iteration(const BatchTune &t, const BatchDev &B, const BatchVec<T> &v,
                           BatchState<T> st, T res, T zero, bool mixed, hipStream_t s)
{
  switch (B.R) {
    case 1: return biter_r<T, 1>(t, B, v, st, res, zero, mixed, s);
    case 2: return biter_r<T, 2>(t, B, v, st, res, zero, mixed, s);
    case 4: return biter_r<T, 4>(t, B, v, st, res, zero, mixed, s);
    default: return biter_r<T, 8>(t, B, v, st, res, zero, mixed, s);
  }
}
template hipError_t batch_iteration<double>(const BatchTune &, const BatchDev &,
                                            const BatchVec<double> &, BatchState<double>, double,
                                            double, bool, hipStream_t);
template hipError_t batch_iteration<float>(const BatchTune &, const BatchDev &,
                                           const BatchVec<float> &, BatchState<float>, float,
                                           float, bool, hipStream_t);

template <typename T, int R>
static hipError_t binit_r(const BatchDev &B, const BatchVec<T> &v, BatchState<T> st, T zero,
                          hipStream_t s, bool jac = true)
{
  T *prr = v.part + (int64_t)B.nblk * R;
  if (jac)
    hipLaunchKernelGGL((k_binit<T, R, true>), wl_grid(B), dim3(TPB), 0, s, B, v.y, v.diag, v.r,
                       v.h, v.d, prr);
  else
    hipLaunchKernelGGL((k_binit<T, R, false>), wl_grid(B), dim3(TPB), 0, s, B, v.y, nullptr, v.r,
                       nullptr, v.d, prr);
  const int nq = B.nsys * R;
  hipLaunchKernelGGL((k_binit2<T, R>), dim3((nq + TPB - 1) / TPB), dim3(TPB), 0, s, B, st, prr,
                     zero);
  return hipGetLastError();
}

template <typename T>
hipError_t batch_init(const BatchDev &B, const BatchVec<T> &v, BatchState<T> st, T zero,
                      hipStream_t s)
{
  switch (B.R) {
    case 1: return binit_r<T, 1>(B, v, st, zero, s);
    case 2: return binit_r<T, 2>(B, v, st, zero, s);
    case 4: return binit_r<T, 4>(B, v, st, zero, s);
    default: return binit_r<T, 8>(B, v, st, zero, s);
  }
}
template hipError_t batch_init<double>(const BatchDev &, const BatchVec<double> &,
                                       BatchState<double>, double, hipStream_t);
template hipError_t batch_init<float>(const BatchDev &, const BatchVec<float> &,
                                      BatchState<float>, float, hipStream_t);

template <typename T>
hipError_t batch_flags(const BatchDev &B, BatchState<T> st, int *active, hipStream_t s)
{
  const unsigned g = (unsigned)((B.nsys + TPB - 1) / TPB);
  switch (B.R) {
    case 1: hipLaunchKernelGGL((k_badvance<T, 1>), dim3(g), dim3(TPB), 0, s, B, st, active); break;
    case 2: hipLaunchKernelGGL((k_badvance<T, 2>), dim3(g), dim3(TPB), 0, s, B, st, active); break;
    case 4: hipLaunchKernelGGL((k_badvance<T, 4>), dim3(g), dim3(TPB), 0, s, B, st, active); break;
    default: hipLaunchKernelGGL((k_badvance<T, 8>), dim3(g), dim3(TPB), 0, s, B, st, active); break;
  }
  return hipGetLastError();
}
template hipError_t batch_flags<double>(const BatchDev &, BatchState<double>, int *, hipStream_t);
template hipError_t batch_flags<float>(const BatchDev &, BatchState<float>, int *, hipStream_t);

// ---------------------------------------------------- mixed-precision steps

__global__ void k_zero_int(int *p) { *p = 0; }

template <int R>
static BatchMix mixview(const BatchMixDev &m)
{
  BatchMix x;
  x.nory64 = m.nory64; x.rr64 = m.rr64; x.best = m.best; x.scale = m.scale; x.delta2 = m.delta2;
  x.norm_a = m.norm_a;
  x.bad = m.bad; x.since_best = m.since_best; x.final = m.final; x.status = m.status;
  x.updates = m.updates;
  return x;
}

template <int R>
static hipError_t bmx_start_r(const BatchTune &t64, const BatchTune &t32, const BatchDev &B,
                              const BatchVec<double> &v64, const BatchVec<float> &v32,
                              BatchState<float> st, const BatchMixDev &md, double res,
                              double delta, hipStream_t s)
{
  BatchMix mx = mixview<R>(md);
  dim3 grid = wl_grid(B), block(TPB);
  const int nq = B.nsys * R;
  const unsigned gq = (unsigned)((nq + TPB - 1) / TPB);
  // y.y and the initial FP64 residual r = b - A x
  double *pyy = v64.part + 2 * (int64_t)B.nblk * R;
  hipLaunchKernelGGL(k_bdot_yy, dim3(B.nblk), block, 0, s, B, R, v64.y, pyy);
  hipError_t e = bspmm_r<double, R>(t64, 2, B, v64.val, v64.sval, v64.x, v64.r, v64.y, v64.part,
                                    nullptr, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_bmx_init<R>), dim3(gq), dim3(TPB), 0, s, B, mx, st, v64.part, pyy, res,
                     delta, md.sys_final);
  // FP32 recurrence on the scaled residual (x32 = 0 set by the caller)
  hipLaunchKernelGGL((k_bmx_scale_rhs<R>), grid, block, 0, s, B, mx, v64.r, md.isd, v32.y);
  e = binit_r<float, R>(B, v32, st, 1e-30f, s, false);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_bmx_init32<R>), dim3(gq), dim3(TPB), 0, s, B, mx, st, res);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, md.active);
  hipLaunchKernelGGL((k_bmx_flags<R>), dim3((B.nsys + TPB - 1) / TPB), dim3(TPB), 0, s, B, mx, st,
                     md.sys_final, md.active);
  return hipGetLastError();
}

template <int R>
static hipError_t bmx_refresh_r(const BatchTune &t64, const BatchDev &B,
                                const BatchVec<double> &v64, const BatchVec<float> &v32,
                                BatchState<float> st, const BatchMixDev &md, double res,
                                hipStream_t s)
{
  BatchMix mx = mixview<R>(md);
  dim3 grid = wl_grid(B), block(TPB);
  const int nq = B.nsys * R;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, md.active);
  double *px = v64.part + (int64_t)B.nblk * R;  // free in the mixed solve
  hipLaunchKernelGGL((k_bmx_xupdate<R>), grid, block, 0, s, B, mx, md.sys_final, md.isd, v64.x,
                     v32.x, px);
  hipError_t e = bspmm_r<double, R>(t64, 2, B, v64.val, v64.sval, v64.x, v64.r, v64.y,
                                    v64.part, md.sys_final, s);
  if (e != hipSuccess) return e;
  float *prr = v32.part + (int64_t)B.nblk * R;
  hipLaunchKernelGGL((k_bmx_refresh<R>), grid, block, 0, s, B, mx, md.sys_final, v64.part, v64.r,
                     md.isd, v32.r, v32.d, prr);
  hipLaunchKernelGGL((k_bmx_check<R>), dim3((nq + TPB - 1) / TPB), dim3(TPB), 0, s, B, mx, st,
                     v64.part, prr, px, res);
  hipLaunchKernelGGL((k_bmx_flags<R>), dim3((B.nsys + TPB - 1) / TPB), dim3(TPB), 0, s, B, mx, st,
                     md.sys_final, md.active);
  return hipGetLastError();
}

// Work list of the blocks of the running systems (skip[s] == 0; every system when
// skip is null), in system order: one workgroup scans the per-system block
// counts 1024 systems at a time. With hist: hist[*ctr] = *active, *ctr += 1.
constexpr int CTPB = 1024;
__global__ __launch_bounds__(CTPB) void k_bcompact(BatchDev B, const int *__restrict__ skip,
                                                   int *__restrict__ wl, int *__restrict__ nwl,
                                                   const int *__restrict__ active,
                                                   int *__restrict__ hist, int *__restrict__ ctr,
                                                   int hist_cap)
{
  __shared__ int scan[CTPB];
  int base = 0;
  for (int s0 = 0; s0 < B.nsys; s0 += CTPB) {
    const int sy = s0 + threadIdx.x;
    int cnt = 0;
    if (sy < B.nsys && !(skip && skip[sy])) cnt = B.sys_blk0[sy + 1] - B.sys_blk0[sy];
    scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < CTPB; off <<= 1) {
      const int v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int start = base + scan[threadIdx.x] - cnt;
    for (int k = 0; k < cnt; k++) wl[start + k] = B.sys_blk0[sy] + k;
    base += scan[CTPB - 1];
    __syncthreads();  // scan is rewritten by the next group of systems
  }
  if (threadIdx.x == 0) {
    *nwl = base;
    if (hist) {
      const int c = *ctr;
      if (c < hist_cap) hist[c] = *active;
      *ctr = c + 1;
    }
  }
}

hipError_t batch_compact(const BatchDev &B, const int *skip, int *wl, int *nwl, const int *active,
                         int *hist, int *ctr, int hist_cap, hipStream_t s)
{
  hipLaunchKernelGGL(k_bcompact, dim3(1), dim3(CTPB), 0, s, B, skip, wl, nwl, active, hist, ctr,
                     hist_cap);
  return hipGetLastError();
}

__global__ void k_offset_copy(int64_t n, const int *__restrict__ src, int *__restrict__ dst,
                              int add)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] + add;
}

hipError_t offset_copy(int64_t n, const int *src, int *dst, int add, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_offset_copy, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n,
                     src, dst, add);
  return hipGetLastError();
}

// ||A||_inf per system: block maxima of the absolute row sums, then the maximum
// over the blocks of each system (order-independent)
__global__ __launch_bounds__(TPB) void k_bnorm_blk(BatchDev B, const double *__restrict__ val,
                                                   double *__restrict__ blkmax)
{
  __shared__ double lds[TPB];
  const int blk = blockIdx.x;
  double m = 0.0;
  for (int i = B.blk_r0[blk] + threadIdx.x; i < B.blk_r1[blk]; i += TPB) {
    double t = 0.0;
    for (int k = B.rowptr[i]; k < B.rowptr[i + 1]; k++) t += fabs(val[k]);
    m = fmax(m, t);
  }
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blkmax[blk] = lds[0];
}

__global__ void k_bnorm_sys(BatchDev B, const double *__restrict__ blkmax,
                            double *__restrict__ norm_a)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.nsys) return;
  double m = 0.0;
  for (int b = B.sys_blk0[s]; b < B.sys_blk0[s + 1]; b++) m = fmax(m, blkmax[b]);
  norm_a[s] = m;
}

hipError_t batch_norm_inf(const BatchDev &B, const double *val, double *blkmax, double *norm_a,
                          hipStream_t s)
{
  hipLaunchKernelGGL(k_bnorm_blk, dim3(B.nblk), dim3(TPB), 0, s, B, val, blkmax);
  hipLaunchKernelGGL(k_bnorm_sys, dim3((B.nsys + TPB - 1) / TPB), dim3(TPB), 0, s, B, blkmax,
                     norm_a);
  return hipGetLastError();
}

hipError_t zero_int(int *p, hipStream_t s)
{
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, p);
  return hipGetLastError();
}

hipError_t batch_mixed_start(const BatchTune &t64, const BatchTune &t32, const BatchDev &B,
                             const BatchVec<double> &v64, const BatchVec<float> &v32,
                             BatchState<float> st, const BatchMixDev &md, double res,
                             double delta, hipStream_t s)
{
  switch (B.R) {
    case 1: return bmx_start_r<1>(t64, t32, B, v64, v32, st, md, res, delta, s);
    case 2: return bmx_start_r<2>(t64, t32, B, v64, v32, st, md, res, delta, s);
    case 4: return bmx_start_r<4>(t64, t32, B, v64, v32, st, md, res, delta, s);
    default: return bmx_start_r<8>(t64, t32, B, v64, v32, st, md, res, delta, s);
  }
}

hipError_t batch_mixed_refresh(const BatchTune &t64, const BatchDev &B,
                               const BatchVec<double> &v64, const BatchVec<float> &v32,
                               BatchState<float> st, const BatchMixDev &md, double res,
                               hipStream_t s)
{
  switch (B.R) {
    case 1: return bmx_refresh_r<1>(t64, B, v64, v32, st, md, res, s);
    case 2: return bmx_refresh_r<2>(t64, B, v64, v32, st, md, res, s);
    case 4: return bmx_refresh_r<4>(t64, B, v64, v32, st, md, res, s);
    default: return bmx_refresh_r<8>(t64, B, v64, v32, st, md, res, s);
  }
}

}  // namespace ssf
