// CSR SpMV (row-vector and LDS-tiled stream kernels) and the PCG recurrences of
// the SIFEL solve path on gfx950.
//
// SpMV kernels, both deterministic for a fixed tiling:
//  * row-vector: G lanes of a wavefront per row, strided over the row's entries,
//    shuffle reduction; rows split in contiguous chunks per block.
//  * stream: the rows are cut into tiles of whole rows holding at most
//    STREAM_CAP entries. Each thread of a block issues STREAM_KPT independent,
//    coalesced loads of values and column indices (32- or 16-bit, the latter
//    relative to the tile's first row), multiplies by the gathered x and leaves
//    the products in LDS; then G lanes per row sum their row's products in a
//    fixed order. Tiles are split in contiguous ranges per block.
//
// PCG per iteration i (GEFEL/scr.cpp:1142-1197):
//   K1  p = A d, partial d.p                       (SpMV kind 1, increments iter)
//   K2  den = d.p; |den| < zero -> stop (status 2); x += a d, r -= a p, h = M^-1 r,
//       partials r.r and r.h                        (k_pcg_update)
//   K3  rr = r.r, nom' = r.h, stopping test of the mode, else d = b d + h
//                                                   (k_pcg_dir)
// Scalars are re-reduced from the per-block partials by every consumer block in
// the same order, so all blocks take the same decision without inter-block
// synchronisation.
#include <hip/hip_ext.h>

#include <type_traits>

#include "ssf_device.hpp"

namespace ssf {

// ------------------------------------------------------------ row-vector SpMV

template <typename T, int G, int KIND>
__device__ __forceinline__ void finish_row(int row, T s, const T *__restrict__ x,
                                           T *__restrict__ y, const T *__restrict__ b, T &acc)
{
  if (KIND == 2) {
    s = b[row] - s;
    acc += s * s;
  } else if (KIND == 1) {
    acc += x[row] * s;
  }
  y[row] = s;
}

template <typename T, int G, int KIND>
__global__ __launch_bounds__(TPB) void k_spmv_vec(CsrDev A, const T *__restrict__ val,
                                                  const T *__restrict__ x, T *__restrict__ y,
                                                  const T *__restrict__ b,
                                                  T *__restrict__ partial, const int *done,
                                                  long long *iter)
{
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  int r0, r1;
  chunk(A.n, r0, r1);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  constexpr int NG = TPB / G;
  T acc = T(0);
  for (int row = r0 + grp; row < r1; row += NG) {
    const T s = row_dot<T, G>(A.rowptr, A.col, val, x, row, lane);
    if (lane == 0) finish_row<T, G, KIND>(row, s, x, y, b, acc);
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------- stream SpMV

template <typename T, bool I16, int G, int KIND>
__global__ __launch_bounds__(TPB) void k_spmv_stream(CsrDev A, const T *__restrict__ val,
                                                     const T *__restrict__ x,
                                                     T *__restrict__ y, const T *__restrict__ b,
                                                     T *__restrict__ partial, const int *done,
                                                     long long *iter)
{
  __shared__ T prod[STREAM_CAP];
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const int nb = gridDim.x;
  const int tpb = (A.ntiles + nb - 1) / nb;
  const int lb = xcd_block(blockIdx.x, nb);
  const int t0 = lb * tpb, t1 = min(A.ntiles, t0 + tpb);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  constexpr int NG = TPB / G;
  T acc = T(0);
  // software pipeline: the values and columns of tile t+1 are loaded into
  // registers while tile t is reduced from LDS
  T v[STREAM_KPT];
  int c[STREAM_KPT];
  auto load_tile = [&](int tile) {
    const int r0 = A.tiles[tile];
    const int s = A.rowptr[r0], e = A.rowptr[A.tiles[tile + 1]];
#pragma unroll
    for (int j = 0; j < STREAM_KPT; j++) {
      const int k = s + j * TPB + threadIdx.x;
      if (k < e) {
        v[j] = ldnt(val + k);
        c[j] = I16 ? r0 + (int)ldnt(A.col16 + k) : ldnt(A.col + k);
      } else {
        v[j] = T(0);
        c[j] = -1;
      }
    }
  };
  if (t0 < t1) load_tile(t0);
  for (int tile = t0; tile < t1; tile++) {
    const int r0 = A.tiles[tile], r1 = A.tiles[tile + 1];
    const int s = A.rowptr[r0];
    // phase 1: products of the streamed entries with the gathered x
#pragma unroll
    for (int j = 0; j < STREAM_KPT; j++)
      if (c[j] >= 0) prod[j * TPB + threadIdx.x] = v[j] * x[c[j]];
    if (tile + 1 < t1) load_tile(tile + 1);
    __syncthreads();
    // phase 2: G lanes per row, fixed order
    const int nrows = r1 - r0;
    for (int rr = grp; rr < nrows; rr += NG) {
      const int row = r0 + rr;
      const int k0 = A.rowptr[row] - s, k1 = A.rowptr[row + 1] - s;
      T sum = T(0);
      for (int k = k0 + lane; k < k1; k += G) sum += prod[k];
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, G);
      if (lane == 0) finish_row<T, G, KIND>(row, sum, x, y, b, acc);
    }
    __syncthreads();
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------- SELL SpMV

// The 64-row slices a wave streams: first, first + step, ... below last.
// map 0: each block a contiguous range of slices, its waves in turn (blocks in
// XCD-contiguous order). map 1 (XCD interleave): workgroups are dispatched
// round-robin to the 8 XCDs; XCD c keeps the contiguous eighth c of the slices
// and all of its waves take them in turn, so the concurrent waves of an XCD
// stream one window of the value array (HBM pages and prefetch stay together)
// while the x entries of neighbouring rows stay in that XCD's L2.
__device__ __forceinline__ void wave_slices(int nslice, int map, int wave, int &first, int &last,
                                            int &step, int W = TPB / 64)
{
  const int nb = gridDim.x;
  if (map == 1 && (nb & 7) == 0) {
    const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nbx = nb >> 3;
    const int spx = (nslice + 7) / 8;
    const int x0 = min(nslice, xcd * spx);
    first = x0 + lb * W + wave;
    last = min(nslice, x0 + spx);
    step = nbx * W;
  } else {
    const int spb = (nslice + nb - 1) / nb;
    const int s0 = xcd_block(blockIdx.x, nb) * spb;
    first = s0 + wave;
    last = min(nslice, s0 + spb);
    step = W;
  }
}

template <typename T>
__device__ __forceinline__ const T *sell_values(const CsrDev &A);
template <>
__device__ __forceinline__ const double *sell_values<double>(const CsrDev &A)
{
  return A.sell_v64;
}
template <>
__device__ __forceinline__ const float *sell_values<float>(const CsrDev &A)
{
  return A.sell_v32;
}

// One wavefront per 64-row slice, one lane per row: the lane streams its row's
// entries SELL_V at a time (one 16/32-byte value load and the columns), gathers x
// and sums in column order. No LDS, no barrier. A block's waves take the slices
// of its contiguous slice range round-robin. Column forms (IDX): 0 int32, 1 16-bit
// offsets from the slice's first row.
template <typename T, int IDX, int KIND>
__global__ __launch_bounds__(TPB) void k_spmv_sell(CsrDev A, const T *__restrict__ x,
                                                   T *__restrict__ y, const T *__restrict__ b,
                                                   T *__restrict__ partial, const int *done,
                                                   long long *iter)
{
  typedef T tv __attribute__((ext_vector_type(SELL_V)));
  typedef int iv __attribute__((ext_vector_type(SELL_V)));
  typedef short sv __attribute__((ext_vector_type(SELL_V)));
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const T *__restrict__ sval = sell_values<T>(A);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  int sfirst, slast, sstep;
  wave_slices(A.nslice, A.slice_map, wave, sfirst, slast, sstep);
  T acc = T(0);
  for (int sl = sfirst; sl < slast; sl += sstep) {
    const int off = A.sell_off[sl];
    const int w = (A.sell_off[sl + 1] - off) / 64;  // a multiple of SELL_V
    T sum = T(0);
    for (int q = 0; q < w; q += SELL_V) {
      const int k = off + q * 64 + lane * SELL_V;
      const tv vv = ldnt((const tv *)(sval + k));
      if (IDX == 1) {  // 16-bit offsets from the slice's first row
        const sv cc = ldnt((const sv *)(A.sell_col16 + k));
        const T *__restrict__ xs = x + sl * 64;
#pragma unroll
        for (int u = 0; u < SELL_V; u++) sum += vv[u] * xs[cc[u]];
      } else {
        const iv cc = ldnt((const iv *)(A.sell_col + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) sum += vv[u] * x[cc[u]];
      }
    }
    const int row = sl * 64 + lane;
    if (row < A.n) finish_row<T, 1, KIND>(row, sum, x, y, b, acc);
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}


// NG consecutive groups of SELL_V entries of one row in the row-pattern form: all
// value loads, then all pattern-offset loads, then all gathers are issued before
// the first product, so the lane waits once per call; the row sum grows in
// column order. vb: byte offset of the first group's values in sval; ob: byte
// offset of its pattern offsets in dict; rb: the row.
template <typename T, int NG>
__device__ __forceinline__ T sell_pat_groups(const T *__restrict__ sval, unsigned vb,
                                             const int *__restrict__ dict, unsigned ob,
                                             const T *__restrict__ x, int rb, T sum)
{
  typedef T tv __attribute__((ext_vector_type(SELL_V)));
  typedef int iv __attribute__((ext_vector_type(SELL_V)));
  tv v[NG];
  iv o[NG];
  T g[NG][SELL_V];
#pragma unroll
  for (int k = 0; k < NG; k++) v[k] = ldnt(at_byte<tv>(sval, vb + k * 64 * SELL_V * sizeof(T)));
#pragma unroll
  for (int k = 0; k < NG; k++) o[k] = *at_byte<iv>(dict, ob + k * SELL_V * sizeof(int));
#pragma unroll
  for (int k = 0; k < NG; k++)
#pragma unroll
    for (int u = 0; u < SELL_V; u++) g[k][u] = *at_byte<T>(x, (unsigned)(rb + o[k][u]) * sizeof(T));
#pragma unroll
  for (int k = 0; k < NG; k++)
#pragma unroll
    for (int u = 0; u < SELL_V; u++) sum += v[k][u] * g[k][u];
  return sum;
}

// the last ng < D groups of a row (ng is the same for the whole wavefront)
template <typename T, int D>
__device__ __forceinline__ T sell_pat_tail(int ng, const T *__restrict__ sval, unsigned vb,
                                           const int *__restrict__ dict, unsigned ob,
                                           const T *__restrict__ x, int rb, T sum)
{
  if constexpr (D > 1) {
    if (ng == D - 1) return sell_pat_groups<T, D - 1>(sval, vb, dict, ob, x, rb, sum);
    return sell_pat_tail<T, D - 1>(ng, sval, vb, dict, ob, x, rb, sum);
  }
  return sum;
}

// SELL-64 with row patterns: the columns of a row are the row plus the offsets of
// its pattern, kept once in a dictionary read through the caches, so the HBM
// stream is the values and a 2-byte pattern id per row. The slice bounds and the
// pattern offsets of a wave's next slice are loaded while the current slice
// streams, and a row is taken D groups (D * SELL_V entries) per round, each round
// with all of its loads in flight before the first product. The row is summed in
// column order, as in the column forms, so the products are equal bit for bit.
// Values, dictionary and x are addressed by 32-bit byte offsets: the layout is
// built with row patterns only when the values and x of either precision fit in
// 4 GiB.
template <typename T, int KIND, int D>
__global__ __launch_bounds__(TPB) void k_spmv_sell_pat(CsrDev A, const T *__restrict__ x,
                                                       T *__restrict__ y, const T *__restrict__ b,
                                                       T *__restrict__ partial, const int *done,
                                                       long long *iter)
{
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const T *__restrict__ sval = sell_values<T>(A);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  int s0w, s1, ST;
  wave_slices(A.nslice, A.slice_map, wave, s0w, s1, ST);
  T acc = T(0);
  // slice bounds one slice ahead, pattern offsets one and pattern ids two
  // slices ahead of the one being summed
  int sl = s0w;
  int offc = 0, endc = 0, dofc = 0, pid1 = 0;
  if (sl < s1) {
    offc = A.sell_off[sl];
    endc = A.sell_off[sl + 1];
    dofc = A.sell_dict_off[A.sell_pid[sl * 64 + lane]];
    if (sl + ST < s1) pid1 = A.sell_pid[(sl + ST) * 64 + lane];
  }
  for (; sl < s1; sl += ST) {
    const int off = __builtin_amdgcn_readfirstlane(offc);
    const int w = __builtin_amdgcn_readfirstlane(endc - offc) / 64;  // a multiple of SELL_V
    const unsigned ob = (unsigned)dofc * sizeof(int);
    const int rb = min(sl * 64 + lane, A.n - 1);
    if (sl + ST < s1) {
      offc = A.sell_off[sl + ST];
      endc = A.sell_off[sl + ST + 1];
      dofc = A.sell_dict_off[pid1];
      if (sl + 2 * ST < s1) pid1 = A.sell_pid[(sl + 2 * ST) * 64 + lane];
    }
    const unsigned vb = (unsigned)(off + lane * SELL_V) * sizeof(T);
    T sum = T(0);
    int q0 = 0;
    for (; q0 + D * SELL_V <= w; q0 += D * SELL_V)
      sum = sell_pat_groups<T, D>(sval, vb + q0 * 64 * sizeof(T), A.sell_dict, ob + q0 * sizeof(int),
                                  x, rb, sum);
    if (q0 < w)
      sum = sell_pat_tail<T, D>((w - q0) / SELL_V, sval, vb + q0 * 64 * sizeof(T), A.sell_dict,
                                ob + q0 * sizeof(int), x, rb, sum);
    const int row = sl * 64 + lane;
    if (row < A.n) finish_row<T, 1, KIND>(row, sum, x, y, b, acc);
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// SELL-64 in the stencil form: every row has the same NGRP * SELL_V slots and slot
// k sits at column row + A.sten[k] (a slot the row lacks holds value 0 and reads
// an x entry clamped into range), so the value stream is the only per-entry
// traffic and the offsets are the same for every lane. A row is one round: all
// value loads, then all gathers, then the products in slot (= column) order, so
// the sums equal the column forms' bit for bit, absent slots adding exact zeros.
template <typename T, int KIND, int NGRP>
__global__ __launch_bounds__(TPB) void k_spmv_sell_sten(CsrDev A, const T *__restrict__ x,
                                                        T *__restrict__ y, const T *__restrict__ b,
                                                        T *__restrict__ partial, const int *done,
                                                        long long *iter)
{
  typedef T tv __attribute__((ext_vector_type(SELL_V)));
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const T *__restrict__ sval = sell_values<T>(A);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  int sfirst, slast, sstep;
  wave_slices(A.nslice, A.slice_map, wave, sfirst, slast, sstep);
  const int nm1 = A.n - 1;
  T acc = T(0);
  for (int sl = sfirst; sl < slast; sl += sstep) {
    const unsigned vb = (unsigned)(sl * 64 * NGRP * SELL_V + lane * SELL_V) * sizeof(T);
    const int rb = sl * 64 + lane;
    tv v[NGRP];
    T g[NGRP][SELL_V];
#pragma unroll
    for (int k = 0; k < NGRP; k++) v[k] = ldnt(at_byte<tv>(sval, vb + k * 64 * SELL_V * sizeof(T)));
#pragma unroll
    for (int k = 0; k < NGRP; k++)
#pragma unroll
      for (int u = 0; u < SELL_V; u++) {
        const int c = min(max(rb + A.sten[k * SELL_V + u], 0), nm1);
        g[k][u] = *at_byte<T>(x, (unsigned)c * sizeof(T));
      }
    T sum = T(0);
#pragma unroll
    for (int k = 0; k < NGRP; k++)
#pragma unroll
      for (int u = 0; u < SELL_V; u++) sum += v[k][u] * g[k][u];
    if (rb < A.n) finish_row<T, 1, KIND>(rb, sum, x, y, b, acc);
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// SELL-64 in the block stencil form: slot k of row r sits at column
// (r - r % B) + A.sten_d[k], the offsets counted from the first row of the row's
// block of B rows. With B = 3 and the unknowns of a node numbered together (3D
// elasticity), every row's columns are its node's 27 neighbour nodes times 3
// unknowns: 81 slots, one list for all rows, where the offsets from the row
// itself would differ with r % 3. A row takes D groups per round; the sum is in
// slot order, which is column order, so the products equal the column forms'
// bit for bit (absent slots hold 0 and add exact zeros for finite x).
template <typename T, int KIND, int D, int B>
__global__ __launch_bounds__(TPB) void k_spmv_sell_bsten(CsrDev A, const T *__restrict__ x,
                                                         T *__restrict__ y, const T *__restrict__ b,
                                                         T *__restrict__ partial, const int *done,
                                                         long long *iter)
{
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const T *__restrict__ sval = sell_values<T>(A);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  int sfirst, slast, sstep;
  wave_slices(A.nslice, A.slice_map, wave, sfirst, slast, sstep);
  const int nm1 = A.n - 1;
  const int ng = (A.sten_n + SELL_V - 1) / SELL_V;  // groups per row
  const int *__restrict__ so = A.sten_d;
  T acc = T(0);
  for (int sl = sfirst; sl < slast; sl += sstep) {
    const unsigned vb = (unsigned)(sl * 64 * ng * SELL_V + lane * SELL_V) * sizeof(T);
    const int rb = sl * 64 + lane;
    const int base = B == 1 ? rb : rb - rb % B;
    T sum = T(0);
    int g0 = 0;
    for (; g0 + D <= ng; g0 += D)
      sum = sten_groups<T, D>(sval, vb + g0 * 64 * SELL_V * sizeof(T), so + g0 * SELL_V, x, base,
                              nm1, sum);
    for (; g0 < ng; g0++)
      sum = sten_groups<T, 1>(sval, vb + g0 * 64 * SELL_V * sizeof(T), so + g0 * SELL_V, x, base,
                              nm1, sum);
    if (rb < A.n) finish_row<T, 1, KIND>(rb, sum, x, y, b, acc);
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

template <typename T>
__device__ __forceinline__ const T *half_values(const CsrDev &A);
template <>
__device__ __forceinline__ const double *half_values<double>(const CsrDev &A)
{
  return A.half_v64;
}
template <>
__device__ __forceinline__ const float *half_values<float>(const CsrDev &A)
{
  return A.half_v32;
}

// KIND as the other SpMV kernels, and 3: p = A d with the block partials of d.p,
// r.r, r.p and p.p (b = r) at partial[j * gridDim.x + block], j = 0..3, for the
// fused step of the FP32 recurrence (k_cg_fused_step).
//
// Symmetric half of the stencil form (symcomprow matrices whose stencil list is
// symmetric, sten[NL] = 0). Each row holds its NL + 1 lower slots, diagonal-major
// (hv[k * pad + r] = a(r, r + sten[k])); the upper entry a(r, r - sten[k]) is
// a(r - sten[k], r), slot k of the row |sten[k]| further on, read from the same
// diagonal. The wavefronts of one XCD stream neighbouring slices together, so the
// mirrored reads are L2 hits and HBM carries each stored value once: 14 values
// per heat row instead of 28. Blocks of 256 or 512 threads, one wavefront per
// slice. The row sums in column order (lower slots ascending, the diagonal, the
// upper ones ascending), so the products equal the other forms' bit for bit.
template <typename T, int KIND, int NL>
__global__ __launch_bounds__(512) void k_spmv_half(CsrDev A, const T *__restrict__ x,
                                                   T *__restrict__ y, const T *__restrict__ b,
                                                   T *__restrict__ partial, const int *done,
                                                   long long *iter)
{
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const T *__restrict__ hv = half_values<T>(A);
  const int W = blockDim.x / 64;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  int sfirst, slast, sstep;
  wave_slices(A.nslice, A.slice_map, wave, sfirst, slast, sstep, W);
  const int nm1 = A.n - 1;
  T acc = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  for (int sl = sfirst; sl < slast; sl += sstep) {
    const int r = sl * 64 + lane;
    T gd;
    T sum = half_row<T, NL>(hv, A.half_pad, [&](int k) { return A.sten[k]; }, x, r, nm1, gd);
    if (r < A.n) {
      if (KIND == 3) {  // the dots of the fused CG step; b holds the residual
        const T rv = b[r];
        acc += gd * sum;
        a1 += rv * rv;
        a2 += rv * sum;
        a3 += sum * sum;
      } else if (KIND == 2) {
        sum = b[r] - sum;
        acc += sum * sum;
      } else if (KIND == 1) {
        acc += gd * sum;
      }
      y[r] = sum;
    }
  }
  if (KIND == 3) {
    __shared__ T lds4[4][8];
    acc = wave_sum(acc);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    a3 = wave_sum(a3);
    if (lane == 0) {
      lds4[0][wave] = acc;
      lds4[1][wave] = a1;
      lds4[2][wave] = a2;
      lds4[3][wave] = a3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      T t = lds4[threadIdx.x][0];
      for (int i = 1; i < W; i++) t += lds4[threadIdx.x][i];
      partial[threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
  } else if (KIND != 0) {
    __shared__ T lds[8];
    acc = wave_sum(acc);
    if (lane == 0) lds[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      T t = lds[0];
      for (int i = 1; i < W; i++) t += lds[i];
      partial[blockIdx.x] = t;
    }
  }
}

// Block variant of the half form (A.half_b = B = 2, 3; CsrDev::half_*): one lane
// per node, its B rows summed in column order (lower neighbour nodes ascending,
// the diagonal block, the upper nodes ascending; within a node its unknowns in
// order), which is the block stencil form's slot order, so the products equal
// the other forms' bit for bit. The entries of a slice of 64 nodes lie together
// (entry i of the slice's nodes in 64 consecutive values), so a wavefront streams
// one contiguous run per slice and every load reads 256 contiguous bytes; 16-byte
// groups per lane were measured as fast with 30 % more traffic, the mirrored
// blocks then pulling whole groups. The upper block towards node q - m_j is the
// transpose of block j of that node, read from the same place of its slice.
template <typename T, int KIND, int B, int NLN>
__global__ __launch_bounds__(512) void k_spmv_halfb(CsrDev A, const T *__restrict__ x,
                                                    T *__restrict__ y, const T *__restrict__ b,
                                                    T *__restrict__ partial, const int *done,
                                                    long long *iter)
{
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  const T *__restrict__ hv = half_values<T>(A);
  const int W = blockDim.x / 64;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  const int nn = A.n / B;
  int sfirst, slast, sstep;
  wave_slices((nn + 63) / 64, A.slice_map, wave, sfirst, slast, sstep, W);
  T acc = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  for (int sl = sfirst; sl < slast; sl += sstep) {
    const int q = sl * 64 + lane;
    T sum[B], xd[B];
    half_node<T, B, NLN>(hv, [&](int j) { return A.half_m[j]; }, x, q, nn, sum, xd);
    if (q < nn) {
#pragma unroll
      for (int t = 0; t < B; t++) {
        const int r = B * q + t;
        T st = sum[t];
        if (KIND == 3) {
          const T rv = b[r];
          acc += xd[t] * st;
          a1 += rv * rv;
          a2 += rv * st;
          a3 += st * st;
        } else if (KIND == 2) {
          st = b[r] - st;
          acc += st * st;
        } else if (KIND == 1) {
          acc += xd[t] * st;
        }
        y[r] = st;
      }
    }
  }
  if (KIND == 3) {
    __shared__ T lds4[4][8];
    acc = wave_sum(acc);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    a3 = wave_sum(a3);
    if (lane == 0) {
      lds4[0][wave] = acc;
      lds4[1][wave] = a1;
      lds4[2][wave] = a2;
      lds4[3][wave] = a3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      T t = lds4[threadIdx.x][0];
      for (int i = 1; i < W; i++) t += lds4[threadIdx.x][i];
      partial[threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
  } else if (KIND != 0) {
    __shared__ T lds[8];
    acc = wave_sum(acc);
    if (lane == 0) lds[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      T t = lds[0];
      for (int i = 1; i < W; i++) t += lds[i];
      partial[blockIdx.x] = t;
    }
  }
}

// entry idx of node q (q < half_pad) from SELL slot slot[2 idx] of row
// half_b q + slot[2 idx + 1], in the slice-major order of k_spmv_halfb
template <typename T>
__global__ void k_half_from_sell_b(CsrDev A, const int *__restrict__ slot, int nv,
                                   const T *__restrict__ sval, T *__restrict__ hv)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)nv * A.half_pad) return;
  const int64_t sl = i / (64 * (int64_t)nv);
  const int idx = (int)(i / 64 % nv), q = (int)(sl * 64 + i % 64);
  T v = T(0);
  if (q < A.n / A.half_b) {
    const int k = slot[2 * idx], r = A.half_b * q + slot[2 * idx + 1];
    const int64_t ng = (A.sten_n + SELL_V - 1) / SELL_V;
    v = sval[(r / 64) * 64 * ng * SELL_V + (k / SELL_V) * 64 * SELL_V + (r % 64) * SELL_V +
             k % SELL_V];
  }
  hv[i] = v;
}

template <typename T>
hipError_t half_from_sell_block(const CsrDev &A, const int *slot, int nv, const T *sval, T *hv,
                                hipStream_t s)
{
  const int64_t tot = (int64_t)nv * A.half_pad;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_half_from_sell_b<T>, dim3((unsigned)((tot + TPB - 1) / TPB)), dim3(TPB), 0,
                     s, A, slot, nv, sval, hv);
  return hipGetLastError();
}
template hipError_t half_from_sell_block<double>(const CsrDev &, const int *, int, const double *,
                                                 double *, hipStream_t);
template hipError_t half_from_sell_block<float>(const CsrDev &, const int *, int, const float *,
                                                float *, hipStream_t);

template <typename T>
__global__ void k_half_from_sell(CsrDev A, const T *__restrict__ sval, T *__restrict__ hv)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)(A.half_nl + 1) * A.half_pad) return;
  const int k = (int)(i / A.half_pad), r = (int)(i % A.half_pad);
  T v = T(0);
  if (r < A.n) {  // the stencil form's slices are all ng * SELL_V slots wide
    const int64_t ng = (A.sten_n + SELL_V - 1) / SELL_V;
    v = sval[(r / 64) * 64 * ng * SELL_V + (k / SELL_V) * 64 * SELL_V + (r % 64) * SELL_V +
             k % SELL_V];
  }
  hv[i] = v;
}

template <typename T>
hipError_t half_from_sell(const CsrDev &A, const T *sval, T *hv, hipStream_t s)
{
  const int64_t tot = (int64_t)(A.half_nl + 1) * A.half_pad;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_half_from_sell<T>, dim3((unsigned)((tot + TPB - 1) / TPB)), dim3(TPB), 0,
                     s, A, sval, hv);
  return hipGetLastError();
}
template hipError_t half_from_sell<double>(const CsrDev &, const double *, double *, hipStream_t);
template hipError_t half_from_sell<float>(const CsrDev &, const float *, float *, hipStream_t);

// ev0/ev1 (both or neither): the launch's own start and end timestamps
// (hipExtLaunchKernelGGL), i.e. the kernel's duration without the dispatch gap
template <typename T, int KIND>
static hipError_t spmv_k(const SpmvTune &t, const CsrDev &A0, const T *val, const T *x, T *y,
                         const T *b, T *partial, const int *done, long long *iter,
                         hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
  dim3 grid(t.blocks), block(TPB);
  CsrDev A = A0;
  A.slice_map = t.map;
#define SSF_LAUNCH(KERN, ...)                                                      \
  do {                                                                             \
    if (ev0) hipExtLaunchKernelGGL(KERN, grid, block, 0, s, ev0, ev1, 0, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERN, grid, block, 0, s, __VA_ARGS__);                 \
  } while (0)
#define SSF_VEC(G) SSF_LAUNCH((k_spmv_vec<T, G, KIND>), A, val, x, y, b, partial, done, iter)
#define SSF_STR(I, G) SSF_LAUNCH((k_spmv_stream<T, I, G, KIND>), A, val, x, y, b, partial, done, iter)
  const bool have_half = sizeof(T) == 8 ? A.half_v64 != nullptr : A.half_v32 != nullptr;
  if (t.format == 2 && A.nslice > 0 && t.idx16 == 4 && A.half_nl > 0 && have_half &&
      A.half_b > 1) {
    const dim3 hblock(t.lanes == 2 ? 512 : 256);
#define SSF_HALFB(B, NLN)                                                                        \
  if (ev0) hipExtLaunchKernelGGL((k_spmv_halfb<T, KIND, B, NLN>), grid, hblock, 0, s, ev0, ev1, 0, \
                                 A, x, y, b, partial, done, iter);                              \
  else hipLaunchKernelGGL((k_spmv_halfb<T, KIND, B, NLN>), grid, hblock, 0, s, A, x, y, b,       \
                          partial, done, iter);
    if (A.half_b == 3 && A.half_nl == 13) { SSF_HALFB(3, 13) }
    else if (A.half_b == 2 && A.half_nl == 4) { SSF_HALFB(2, 4) }
    else return hipErrorInvalidValue;
#undef SSF_HALFB
  } else if (t.format == 2 && A.nslice > 0 && t.idx16 == 4 && A.half_nl > 0 && have_half) {
    const dim3 hblock(t.lanes == 2 ? 512 : 256);
#define SSF_HALF(NL)                                                                        \
  case NL:                                                                                  \
    if (ev0) hipExtLaunchKernelGGL((k_spmv_half<T, KIND, NL>), grid, hblock, 0, s, ev0, ev1, 0, \
                                   A, x, y, b, partial, done, iter);                        \
    else hipLaunchKernelGGL((k_spmv_half<T, KIND, NL>), grid, hblock, 0, s, A, x, y, b,     \
                            partial, done, iter);                                           \
    break;
    switch (A.half_nl) {
      SSF_HALF(1) SSF_HALF(2) SSF_HALF(3) SSF_HALF(4) SSF_HALF(5) SSF_HALF(6) SSF_HALF(7)
      SSF_HALF(8) SSF_HALF(9) SSF_HALF(10) SSF_HALF(11) SSF_HALF(12) SSF_HALF(13)
      SSF_HALF(14) SSF_HALF(15)
      default: return hipErrorInvalidValue;
    }
#undef SSF_HALF
  } else if (t.format == 2 && A.nslice > 0 && t.idx16 >= 3 && A.sten_n > 0 &&
      (A.sten_b != 1 || A.sten_n > SELL_STEN_MAX || t.lanes == 3 || t.lanes == 4)) {
    // groups per round: t.lanes 3 or 4 takes that many, else 7 (5 for blocks of 2)
#define SSF_BSTEN(D, B) SSF_LAUNCH((k_spmv_sell_bsten<T, KIND, D, B>), A, x, y, b, partial, done, iter)
    if (A.sten_b == 3) {
      if (t.lanes == 3) SSF_BSTEN(3, 3);
      else if (t.lanes == 4) SSF_BSTEN(4, 3);
      else SSF_BSTEN(7, 3);
    } else if (A.sten_b == 2) {
      if (t.lanes == 3) SSF_BSTEN(3, 2);
      else if (t.lanes == 4) SSF_BSTEN(4, 2);
      else SSF_BSTEN(5, 2);
    } else {
      if (t.lanes == 3) SSF_BSTEN(3, 1);
      else if (t.lanes == 4) SSF_BSTEN(4, 1);
      else SSF_BSTEN(7, 1);
    }
#undef SSF_BSTEN
  } else if (t.format == 2 && A.nslice > 0 && t.idx16 >= 3 && A.sten_n > 0) {
    switch ((A.sten_n + SELL_V - 1) / SELL_V) {
#define SSF_STEN(G) case G: SSF_LAUNCH((k_spmv_sell_sten<T, KIND, G>), A, x, y, b, partial, done, iter); break;
      SSF_STEN(1) SSF_STEN(2) SSF_STEN(3) SSF_STEN(4) SSF_STEN(5) SSF_STEN(6) SSF_STEN(7)
      default: SSF_LAUNCH((k_spmv_sell_sten<T, KIND, 8>), A, x, y, b, partial, done, iter); break;
#undef SSF_STEN
    }
  } else if (t.format == 2 && A.nslice > 0) {
    if (t.idx16 == 2 && A.sell_pid) {
      if (t.lanes >= 8) SSF_LAUNCH((k_spmv_sell_pat<T, KIND, 8>), A, x, y, b, partial, done, iter);
      else if (t.lanes >= 4) SSF_LAUNCH((k_spmv_sell_pat<T, KIND, 4>), A, x, y, b, partial, done, iter);
      else SSF_LAUNCH((k_spmv_sell_pat<T, KIND, 2>), A, x, y, b, partial, done, iter);
    } else if (t.idx16 && A.sell_col16)
      SSF_LAUNCH((k_spmv_sell<T, 1, KIND>), A, x, y, b, partial, done, iter);
    else
      SSF_LAUNCH((k_spmv_sell<T, 0, KIND>), A, x, y, b, partial, done, iter);
  } else if (t.format == 1 && A.ntiles > 0) {
    const bool i16 = t.idx16 && A.col16;
    switch (t.lanes) {
      case 4: if (i16) SSF_STR(true, 4); else SSF_STR(false, 4); break;
      case 8: if (i16) SSF_STR(true, 8); else SSF_STR(false, 8); break;
      default: if (i16) SSF_STR(true, 16); else SSF_STR(false, 16); break;
    }
  } else {
    switch (t.lanes) {
      case 4: SSF_VEC(4); break;
      case 8: SSF_VEC(8); break;
      case 16: SSF_VEC(16); break;
      case 32: SSF_VEC(32); break;
      default: SSF_VEC(64); break;
    }
  }
#undef SSF_VEC
#undef SSF_STR
#undef SSF_LAUNCH
  return hipGetLastError();
}

template <typename T>
hipError_t spmv(const SpmvTune &t, int kind, const CsrDev &A, const T *val, const T *x, T *y,
                const T *b, T *partial, const int *done, long long *iter, hipStream_t s,
                hipEvent_t ev0, hipEvent_t ev1)
{
  if (kind == 0) return spmv_k<T, 0>(t, A, val, x, y, b, partial, done, iter, s, ev0, ev1);
  if (kind == 1) return spmv_k<T, 1>(t, A, val, x, y, b, partial, done, iter, s, ev0, ev1);
  return spmv_k<T, 2>(t, A, val, x, y, b, partial, done, iter, s, ev0, ev1);
}
template hipError_t spmv<double>(const SpmvTune &, int, const CsrDev &, const double *,
                                 const double *, double *, const double *, double *,
                                 const int *, long long *, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t spmv<float>(const SpmvTune &, int, const CsrDev &, const float *,
                                const float *, float *, const float *, float *, const int *,
                                long long *, hipStream_t, hipEvent_t, hipEvent_t);

// p = A d in the half stencil form with the four dots of the fused CG step
template <typename T>
static hipError_t spmv_half_dots(const SpmvTune &t, const CsrDev &A0, const T *d, T *p,
                                 const T *r, T *partial, const int *done, long long *iter,
                                 hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
  CsrDev A = A0;
  A.slice_map = t.map;
  const dim3 grid(t.blocks), block(t.lanes == 2 ? 512 : 256);
  if (A.half_b > 1) {
#define SSF_HALFB3(B, NLN)                                                                      \
  if (ev0) hipExtLaunchKernelGGL((k_spmv_halfb<T, 3, B, NLN>), grid, block, 0, s, ev0, ev1, 0,   \
                                 A, d, p, r, partial, done, iter);                              \
  else hipLaunchKernelGGL((k_spmv_halfb<T, 3, B, NLN>), grid, block, 0, s, A, d, p, r, partial,  \
                          done, iter);
    if (A.half_b == 3 && A.half_nl == 13) { SSF_HALFB3(3, 13) }
    else if (A.half_b == 2 && A.half_nl == 4) { SSF_HALFB3(2, 4) }
    else return hipErrorInvalidValue;
#undef SSF_HALFB3
    return hipGetLastError();
  }
#define SSF_HALF3(NL)                                                                        \
  case NL:                                                                                   \
    if (ev0) hipExtLaunchKernelGGL((k_spmv_half<T, 3, NL>), grid, block, 0, s, ev0, ev1, 0,   \
                                   A, d, p, r, partial, done, iter);                         \
    else hipLaunchKernelGGL((k_spmv_half<T, 3, NL>), grid, block, 0, s, A, d, p, r, partial, \
                            done, iter);                                                     \
    break;
  switch (A.half_nl) {
    SSF_HALF3(1) SSF_HALF3(2) SSF_HALF3(3) SSF_HALF3(4) SSF_HALF3(5) SSF_HALF3(6) SSF_HALF3(7)
    SSF_HALF3(8) SSF_HALF3(9) SSF_HALF3(10) SSF_HALF3(11) SSF_HALF3(12) SSF_HALF3(13)
    SSF_HALF3(14) SSF_HALF3(15)
    default: return hipErrorInvalidValue;
  }
#undef SSF_HALF3
  return hipGetLastError();
}

// ----------------------------------------------------------------------- PCG

// r = y - q (q = A x, or 0), h = M^-1 r, d = h, partials y.y and r.h
template <typename T, bool JAC, bool HAVEX>
__global__ __launch_bounds__(TPB) void k_pcg_init(int n, const T *__restrict__ diag,
                                                  const T *__restrict__ y, T *__restrict__ r,
                                                  T *__restrict__ h, T *__restrict__ d,
                                                  T *__restrict__ part)
{
  int r0, r1;
  chunk(n, r0, r1);
  T ayy = T(0), arh = T(0);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const T yi = y[i];
    const T ri = HAVEX ? yi - r[i] : yi;
    const T hi = JAC ? ri / diag[i] : ri;
    r[i] = ri;
    if (JAC) h[i] = hi;
    d[i] = hi;
    ayy += yi * yi;
    arh += ri * hi;
  }
  const T a = block_sum(ayy);
  const T c = block_sum(arh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[2 * gridDim.x + blockIdx.x] = c;
  }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_pcg_init2(const T *__restrict__ part, int nb,
                                                   PcgState<T> *st, T zero)
{
  const T nory = reduce_partials(part, nb);
  const T nom = reduce_partials(part + 2 * nb, nb);
  if (threadIdx.x == 0) {
    st->nory = nory;
    st->nom[0] = nom;
    st->nom[1] = nom;
    st->iter = -1;
    st->status = 0;
    st->done = 0;
    st->scale = T(1);
    if (nory < zero) {
      st->status = 3;
      st->done = 1;
      st->ani = 0;
      st->ares = nory;
    }
  }
}

// 16-byte groups of a vector: four FP32 or two FP64 rows per access. Every vector
// is its own allocation, so row chunks aligned to a group are 16-byte aligned.
template <typename T>
struct Grp {
  static constexpr int L = 16 / (int)sizeof(T);
  typedef T v __attribute__((ext_vector_type(16 / sizeof(T))));
};
// groups per thread loaded before a kernel knows the scalars of the previous
// one (the re-reduction of its partials), so the two latencies overlap; a
// block's chunk of a 2M-row vector (512 to 1024 blocks) takes 1 to 4 of them
constexpr int PCG_PF = 4;

template <typename T, bool JAC>
__global__ __launch_bounds__(TPB) void k_pcg_update(int n, PcgState<T> *st,
                                                    const T *__restrict__ pden, int npden,
                                                    T zero, T *__restrict__ x,
                                                    T *__restrict__ r, const T *__restrict__ d,
                                                    const T *__restrict__ p,
                                                    const T *__restrict__ diag,
                                                    T *__restrict__ h, T *__restrict__ prr)
{
  if (st->done) return;
  typedef typename Grp<T>::v V;
  constexpr int L = Grp<T>::L;
  const int nb = gridDim.x;
  int r0, r1;
  chunk_al<L>(n, r0, r1);
  const int rL = r0 + (r1 - r0) / L * L;
  V xv[PCG_PF], dv[PCG_PF], pv[PCG_PF], rv[PCG_PF], gv[PCG_PF];
#pragma unroll
  for (int j = 0; j < PCG_PF; j++) {
    const int i = r0 + L * (threadIdx.x + j * TPB);
    if (i < rL) {
      xv[j] = *(const V *)(x + i);
      dv[j] = *(const V *)(d + i);
      pv[j] = *(const V *)(p + i);
      rv[j] = *(const V *)(r + i);
      if (JAC) gv[j] = *(const V *)(diag + i);
    }
  }
  const T den = reduce_partials(pden, npden);
  const long long it = st->iter;
  const T nom = st->nom[it & 1];
  if (fabs(den) < zero) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = 2;
      st->ani = it;
      st->ares = nom;
      st->done = 1;
    }
    return;
  }
  const T alpha = nom / den;
  T rr = T(0), rh = T(0);
  // x += a d, r -= a p, h = r / diag, partials r.r and r.h, rows in order per thread
  auto group = [&](V &xg, const V &dg, const V &pg, V &rg, const V &gg, int i) {
#pragma unroll
    for (int c = 0; c < L; c++) {
      xg[c] += alpha * dg[c];
      rg[c] -= alpha * pg[c];
      rr += rg[c] * rg[c];
    }
    if (JAC) {
      V hg;
#pragma unroll
      for (int c = 0; c < L; c++) {
        hg[c] = rg[c] / gg[c];
        rh += rg[c] * hg[c];
      }
      *(V *)(h + i) = hg;
    }
    *(V *)(x + i) = xg;
    *(V *)(r + i) = rg;
  };
#pragma unroll
  for (int j = 0; j < PCG_PF; j++) {
    const int i = r0 + L * (threadIdx.x + j * TPB);
    if (i < rL) group(xv[j], dv[j], pv[j], rv[j], gv[j], i);
  }
  for (int i = r0 + L * (threadIdx.x + PCG_PF * TPB); i < rL; i += L * TPB) {
    V xg = *(const V *)(x + i), rg = *(const V *)(r + i);
    const V dg = *(const V *)(d + i), pg = *(const V *)(p + i);
    V gg;
    if (JAC) gg = *(const V *)(diag + i);
    group(xg, dg, pg, rg, gg, i);
  }
  for (int i = rL + threadIdx.x; i < r1; i += TPB) {  // the chunk's last rows
    x[i] += alpha * d[i];
    const T ri = r[i] - alpha * p[i];
    r[i] = ri;
    rr += ri * ri;
    if (JAC) {
      const T hi = ri / diag[i];
      h[i] = hi;
      rh += ri * hi;
    }
  }
  const T a = block_sum(rr);
  if (threadIdx.x == 0) prr[blockIdx.x] = a;
  if (JAC) {
    const T c = block_sum(rh);
    if (threadIdx.x == 0) prr[nb + blockIdx.x] = c;
  }
}

// d = beta d + h over the block's chunk, the first groups of each thread given
template <typename T>
__device__ __forceinline__ void pcg_dir_rows(int r0, int rL, int r1, T beta, T *__restrict__ d,
                                             const T *__restrict__ h,
                                             const typename Grp<T>::v *dv,
                                             const typename Grp<T>::v *hv)
{
  typedef typename Grp<T>::v V;
  constexpr int L = Grp<T>::L;
#pragma unroll
  for (int j = 0; j < PCG_PF; j++) {
    const int i = r0 + L * (threadIdx.x + j * TPB);
    if (i < rL) {
      V dg;
#pragma unroll
      for (int c = 0; c < L; c++) dg[c] = beta * dv[j][c] + hv[j][c];
      *(V *)(d + i) = dg;
    }
  }
  for (int i = r0 + L * (threadIdx.x + PCG_PF * TPB); i < rL; i += L * TPB) {
    V dg = *(const V *)(d + i);
    const V hg = *(const V *)(h + i);
#pragma unroll
    for (int c = 0; c < L; c++) dg[c] = beta * dg[c] + hg[c];
    *(V *)(d + i) = dg;
  }
  for (int i = rL + threadIdx.x; i < r1; i += TPB) d[i] = beta * d[i] + h[i];
}

template <typename T, bool JAC, int MODE>
__global__ __launch_bounds__(TPB) void k_pcg_dir(int n, PcgState<T> *st,
                                                 const T *__restrict__ prr, T res, T zero,
                                                 T *__restrict__ d, const T *__restrict__ h)
{
  if (st->done) return;
  typedef typename Grp<T>::v V;
  constexpr int L = Grp<T>::L;
  int r0, r1;
  chunk_al<L>(n, r0, r1);
  const int rL = r0 + (r1 - r0) / L * L;
  V dv[PCG_PF], hv[PCG_PF];
#pragma unroll
  for (int j = 0; j < PCG_PF; j++) {
    const int i = r0 + L * (threadIdx.x + j * TPB);
    if (i < rL) {
      dv[j] = *(const V *)(d + i);
      hv[j] = *(const V *)(h + i);
    }
  }
  const int nb = gridDim.x;
  const T rr = reduce_partials(prr, nb);
  const T nomn = JAC ? reduce_partials(prr + nb, nb) : rr;
  const long long it = st->iter;
  const T nom = st->nom[it & 1];
  const T nory = st->nory;
  if (MODE == MODE_MIXED_INNER) {
    // the direction is always updated, then rr < nory (the refresh threshold)
    // pauses the recurrence until the FP64 refresh
    pcg_dir_rows<T>(r0, rL, r1, nomn / nom, d, h, dv, hv);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->nom[(it + 1) & 1] = nomn;
      if (rr < nory) {
        st->ani = it;
        st->status = 1;
        st->done = 1;
      }
    }
    return;
  }
  bool stop;
  if (MODE == MODE_SCR_CG_PREC) stop = fabs(rr / nory) < res;                         // scr.cpp:1188
  else if (MODE == MODE_SCR_CG) stop = (sqrt(nomn / nory) < res) || (fabs(nomn) < zero);  // :862-870
  else stop = sqrt(nomn / nory) < res;                                                // cr.cpp:1217
  if (stop) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->ani = it;
      st->ares = (MODE == MODE_SCR_CG_PREC) ? nomn / nory : nomn;
      st->status = 1;
      st->nom[(it + 1) & 1] = nomn;
      st->done = 1;
    }
    return;
  }
  pcg_dir_rows<T>(r0, rL, r1, nomn / nom, d, h, dv, hv);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->nom[(it + 1) & 1] = nomn;
}

// One step of the FP32 recurrence of the mixed solve after p = A d
// (k_spmv_half<T, 3>): alpha = r.r / d.p, x += alpha d, r -= alpha p, and with
// ||r - alpha p||^2 = r.r - 2 alpha r.p + alpha^2 p.p known from the same dots,
// beta and d = r + beta d in the same pass. Two kernels per iteration instead of
// three; r.r is the true dot of every step, so the recurrence's estimate carries
// one step's rounding only. As in MODE_MIXED_INNER, the direction is updated and
// then ||r||^2 < nory pauses the recurrence until the FP64 refresh.
template <typename T>
__global__ __launch_bounds__(TPB) void k_cg_fused_step(int n, PcgState<T> *st,
                                                       const T *__restrict__ part, int np,
                                                       T zero, T *__restrict__ x,
                                                       T *__restrict__ r, T *__restrict__ d,
                                                       const T *__restrict__ p)
{
  if (st->done) return;
  typedef typename Grp<T>::v V;
  constexpr int L = Grp<T>::L;
  int r0, r1;
  chunk_al<L>(n, r0, r1);
  const int rL = r0 + (r1 - r0) / L * L;
  V xv[PCG_PF], dv[PCG_PF], pv[PCG_PF], rv[PCG_PF];
#pragma unroll
  for (int j = 0; j < PCG_PF; j++) {
    const int i = r0 + L * (threadIdx.x + j * TPB);
    if (i < rL) {
      xv[j] = *(const V *)(x + i);
      dv[j] = *(const V *)(d + i);
      pv[j] = *(const V *)(p + i);
      rv[j] = *(const V *)(r + i);
    }
  }
  // the four dots, re-reduced in a fixed order by every block
  T q[4] = {T(0), T(0), T(0), T(0)};
  for (int i = threadIdx.x; i < np; i += TPB)
#pragma unroll
    for (int j = 0; j < 4; j++) q[j] += part[j * np + i];
  __shared__ T lds[4][TPB / 64];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    q[j] = wave_sum(q[j]);
    if ((threadIdx.x & 63) == 0) lds[j][threadIdx.x >> 6] = q[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; j++) {
    q[j] = lds[j][0];
#pragma unroll
    for (int w = 1; w < TPB / 64; w++) q[j] += lds[j][w];
  }
  const T dp = q[0], rr = q[1], rp = q[2], pp = q[3];
  const long long it = st->iter;
  if (fabs(dp) < zero) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = 2;
      st->ani = it;
      st->ares = rr;
      st->done = 1;
    }
    return;
  }
  const T alpha = rr / dp;
  const T rrn = rr - T(2) * alpha * rp + alpha * alpha * pp;
  const T beta = rrn / rr;
  auto group = [&](V &xg, V &dg, const V &pg, V &rg, int i) {
#pragma unroll
    for (int c = 0; c < L; c++) {
      xg[c] += alpha * dg[c];
      rg[c] -= alpha * pg[c];
      dg[c] = rg[c] + beta * dg[c];
    }
    *(V *)(x + i) = xg;
    *(V *)(r + i) = rg;
    *(V *)(d + i) = dg;
  };
#pragma unroll
  for (int j = 0; j < PCG_PF; j++) {
    const int i = r0 + L * (threadIdx.x + j * TPB);
    if (i < rL) group(xv[j], dv[j], pv[j], rv[j], i);
  }
  for (int i = r0 + L * (threadIdx.x + PCG_PF * TPB); i < rL; i += L * TPB) {
    V xg = *(const V *)(x + i), dg = *(const V *)(d + i), rg = *(const V *)(r + i);
    const V pg = *(const V *)(p + i);
    group(xg, dg, pg, rg, i);
  }
  for (int i = rL + threadIdx.x; i < r1; i += TPB) {  // the chunk's last rows
    x[i] += alpha * d[i];
    const T ri = r[i] - alpha * p[i];
    r[i] = ri;
    d[i] = ri + beta * d[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->nom[(it + 1) & 1] = rrn;
    if (rrn < st->nory) {
      st->ani = it;
      st->status = 1;
      st->done = 1;
    }
  }
}

// the element-wise PCG kernels use at most this many blocks
static int vec_blocks(const SpmvTune &t, int n)
{
  int nb = std::min(t.blocks, 1024);
  const int need = (n + TPB * 4 - 1) / (TPB * 4);
  nb = std::max(8, std::min(nb, need));
  return (nb + 7) / 8 * 8;
}

template <typename T>
hipError_t pcg_init(const SpmvTune &t, bool jac, bool havex, const CsrDev &A, const T *val,
                    const T *diag, const T *x, const T *y, T *r, T *h, T *d, T *part,
                    PcgState<T> *st, T zero, hipStream_t s)
{
  if (havex) {
    hipError_t e = spmv<T>(t, 0, A, val, x, r, nullptr, nullptr, nullptr, nullptr, s);
    if (e != hipSuccess) return e;
  }
  const int nb = vec_blocks(t, A.n);
  dim3 grid(nb), block(TPB);
  if (jac && havex) hipLaunchKernelGGL((k_pcg_init<T, true, true>), grid, block, 0, s, A.n, diag, y, r, h, d, part);
  else if (jac) hipLaunchKernelGGL((k_pcg_init<T, true, false>), grid, block, 0, s, A.n, diag, y, r, h, d, part);
  else if (havex) hipLaunchKernelGGL((k_pcg_init<T, false, true>), grid, block, 0, s, A.n, diag, y, r, h, d, part);
  else hipLaunchKernelGGL((k_pcg_init<T, false, false>), grid, block, 0, s, A.n, diag, y, r, h, d, part);
  hipLaunchKernelGGL((k_pcg_init2<T>), dim3(1), dim3(TPB), 0, s, part, nb, st, zero);
  return hipGetLastError();
}
template hipError_t pcg_init<double>(const SpmvTune &, bool, bool, const CsrDev &, const double *,
                                     const double *, const double *, const double *, double *,
                                     double *, double *, double *, PcgState<double> *, double,
                                     hipStream_t);
template hipError_t pcg_init<float>(const SpmvTune &, bool, bool, const CsrDev &, const float *,
                                    const float *, const float *, const float *, float *, float *,
                                    float *, float *, PcgState<float> *, float, hipStream_t);

// Partial layout in part[]: [0, t.blocks) d.p of K1; then, for the element-wise
// kernels with nv blocks, [P1, P1 + nv) r.r and [P1 + nv, P1 + 2 nv) r.h, with
// P1 = t.blocks.
template <typename T>
hipError_t pcg_iteration(const SpmvTune &t, int mode, bool jac, const CsrDev &A, const T *val,
                         const T *diag, T *x, T *r, T *d, T *p, T *h, T *part,
                         PcgState<T> *st, T res, T zero, hipStream_t s, hipEvent_t ev0,
                         hipEvent_t ev1)
{
  const bool have_half = sizeof(T) == 8 ? A.half_v64 != nullptr : A.half_v32 != nullptr;
  if (mode == MODE_MIXED_INNER && !jac && t.format == 2 && t.idx16 == 4 && A.half_nl > 0 &&
      have_half) {
    // the FP32 recurrence of the mixed solve in two kernels (k_cg_fused_step)
    hipError_t e = spmv_half_dots<T>(t, A, d, p, r, part, &st->done, &st->iter, s, ev0, ev1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_cg_fused_step<T>), dim3(vec_blocks(t, A.n)), dim3(TPB), 0, s, A.n, st,
                       part, t.blocks, zero, x, r, d, p);
    return hipGetLastError();
  }
  hipError_t e = spmv<T>(t, 1, A, val, d, p, nullptr, part, &st->done, &st->iter, s, ev0, ev1);
  if (e != hipSuccess) return e;
  const int nv = vec_blocks(t, A.n);
  dim3 grid(nv), block(TPB);
  T *prr = part + t.blocks;
  const int n = A.n;
  if (jac) hipLaunchKernelGGL((k_pcg_update<T, true>), grid, block, 0, s, n, st, part, t.blocks, zero, x, r, d, p, diag, h, prr);
  else hipLaunchKernelGGL((k_pcg_update<T, false>), grid, block, 0, s, n, st, part, t.blocks, zero, x, r, d, p, diag, h, prr);
  const T *hh = jac ? h : r;
#define SSF_DIR(J, M) hipLaunchKernelGGL((k_pcg_dir<T, J, M>), grid, block, 0, s, n, st, prr, res, zero, d, hh)
  if (jac) {
    if (mode == MODE_SCR_CG_PREC) SSF_DIR(true, MODE_SCR_CG_PREC);
    else if (mode == MODE_SCR_CG) SSF_DIR(true, MODE_SCR_CG);
    else if (mode == MODE_MIXED_INNER) SSF_DIR(true, MODE_MIXED_INNER);
    else SSF_DIR(true, MODE_CR_CG_PREC);
  } else {
    if (mode == MODE_SCR_CG_PREC) SSF_DIR(false, MODE_SCR_CG_PREC);
    else if (mode == MODE_SCR_CG) SSF_DIR(false, MODE_SCR_CG);
    else if (mode == MODE_MIXED_INNER) SSF_DIR(false, MODE_MIXED_INNER);
    else SSF_DIR(false, MODE_CR_CG_PREC);
  }
#undef SSF_DIR
  return hipGetLastError();
}
template hipError_t pcg_iteration<double>(const SpmvTune &, int, bool, const CsrDev &,
                                          const double *, const double *, double *, double *,
                                          double *, double *, double *, double *,
                                          PcgState<double> *, double, double, hipStream_t,
                                          hipEvent_t, hipEvent_t);
template hipError_t pcg_iteration<float>(const SpmvTune &, int, bool, const CsrDev &,
                                         const float *, const float *, float *, float *, float *,
                                         float *, float *, float *, PcgState<float> *, float,
                                         float, hipStream_t, hipEvent_t, hipEvent_t);

// ------------------------------------------- PCG with an external preconditioner

// partials of a.b in the element-wise kernels' row chunks; skipped once stopped
__global__ __launch_bounds__(TPB) void k_prec_dot(int n, const PcgState<double> *st,
                                                  const double *__restrict__ a,
                                                  const double *__restrict__ b,
                                                  double *__restrict__ part)
{
  if (st && st->done) return;
  int r0, r1;
  chunk(n, r0, r1);
  double acc = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) acc += a[i] * b[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// d = h and the partials of r.h
__global__ __launch_bounds__(TPB) void k_prec_start(int n, const double *__restrict__ r,
                                                    const double *__restrict__ h,
                                                    double *__restrict__ d,
                                                    double *__restrict__ part)
{
  int r0, r1;
  chunk(n, r0, r1);
  double acc = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double hi = h[i];
    d[i] = hi;
    acc += r[i] * hi;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

hipError_t pcg_init_prec(const SpmvTune &t, bool havex, const CsrDev &A, const double *val,
                         const double *x, const double *y, double *r, double *h, double *d,
                         double *part, PcgState<double> *st, double zero, hipStream_t s,
                         const std::function<hipError_t()> &prec)
{
  if (havex) {
    hipError_t e = spmv<double>(t, 0, A, val, x, r, nullptr, nullptr, nullptr, nullptr, s);
    if (e != hipSuccess) return e;
  }
  const int nb = vec_blocks(t, A.n);
  dim3 grid(nb), block(TPB);
  // r = y - A x (or y), partials y.y into part[0, nb)
  if (havex) hipLaunchKernelGGL((k_pcg_init<double, false, true>), grid, block, 0, s, A.n, nullptr, y, r, h, d, part);
  else hipLaunchKernelGGL((k_pcg_init<double, false, false>), grid, block, 0, s, A.n, nullptr, y, r, h, d, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = prec();  // h = M^-1 r
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_prec_start, grid, block, 0, s, A.n, r, h, d, part + 2 * nb);
  hipLaunchKernelGGL((k_pcg_init2<double>), dim3(1), dim3(TPB), 0, s, part, nb, st, zero);
  return hipGetLastError();
}

hipError_t pcg_iteration_prec(const SpmvTune &t, const CsrDev &A, const double *val, double *x,
                              double *r, double *d, double *p, double *h, double *part,
                              PcgState<double> *st, double res, double zero, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1,
                              const std::function<hipError_t()> &prec)
{
  hipError_t e = spmv<double>(t, 1, A, val, d, p, nullptr, part, &st->done, &st->iter, s, ev0, ev1);
  if (e != hipSuccess) return e;
  const int nv = vec_blocks(t, A.n);
  dim3 grid(nv), block(TPB);
  double *prr = part + t.blocks;
  const int n = A.n;
  // x += a d, r -= a p, partials r.r into prr[0, nv)
  hipLaunchKernelGGL((k_pcg_update<double, false>), grid, block, 0, s, n, st, part, t.blocks, zero,
                     x, r, d, p, nullptr, nullptr, prr);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = prec();  // h = M^-1 r
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_prec_dot, grid, block, 0, s, n, st, r, h, prr + nv);
  // the cg_prec test on r.r, nom = r.h, d = b d + h
  hipLaunchKernelGGL((k_pcg_dir<double, true, MODE_SCR_CG_PREC>), grid, block, 0, s, n, st, prr,
                     res, zero, d, h);
  return hipGetLastError();
}

}  // namespace ssf

// ------------------------------------------- mixed precision: reliable updates
//
// The FP32 inner solve is CG on the symmetrically scaled operator
//   As = float(S A S),  S = diag(1/sqrt|a_ii|),
// which is Jacobi-PCG on A in exact arithmetic (the iterates map by
// x = S xs, r = S^-1 rs): the diagonal of As is 1, so the recurrence carries no
// preconditioned residual and each iteration moves x, r, d, p only.
// It runs on the residual scaled by s = 1/||r|| (the FP64 residual norm at the
// last refresh, so FP32 values stay O(1)). Every K iterations (and whenever the
// FP32 recurrence meets its test) the correction is folded into the FP64
// solution and the residual is recomputed in FP64:
//   x64 += S xs / s, xs = 0;  r64 = b - A x64 (FP64);  s' = 1/||r64||;
//   rs = s' S r64, d *= s'/s;  FP64 test ||r64||^2 < res ||b||^2.
// The FP32 search direction is carried across the refresh (rescaled with the
// residual, which leaves alpha*d consistent), so the Krylov recurrence continues
// on the corrected residual and the iteration count stays close to the FP64
// recurrence while each iteration moves the FP32 traffic.

namespace ssf {

// 1/sqrt|a_ii| (1 where a_ii = 0)
__global__ void k_inv_sqrt_diag(int n, const double *__restrict__ diag, double *__restrict__ isd)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = fabs(diag[i]);
  isd[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// the scaled FP32 copy of the full CSR: v32[k] = float(v[k] (isd[i] isd[col[k]])), the
// scale product taken first so that mirrored entries round alike (a symmetric
// operator stays bitwise symmetric)
__global__ void k_csr_scale32(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                              const double *__restrict__ v, const double *__restrict__ isd,
                              float *__restrict__ v32)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double si = isd[i];
  for (int k = rowptr[i]; k < rowptr[i + 1]; k++) v32[k] = (float)(v[k] * (si * isd[col[k]]));
}

// the same for the SELL-64 values; one thread per layout row. A slot's column
// comes from whichever column form the layout holds (all agree on the entries
// that are not padding, and padding holds 0)
__global__ void k_sell_scale32(CsrDev A, const double *__restrict__ v,
                               const double *__restrict__ isd, float *__restrict__ v32)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.nslice * 64) return;
  const int sl = r / 64, lane = r % 64;
  const int off = A.sell_off[sl];
  const int w = (A.sell_off[sl + 1] - off) / 64;
  const int rc = min(r, A.n - 1);
  const double si = isd[rc];
  for (int q = 0; q < w; q++) {
    const int k = off + (q / SELL_V) * 64 * SELL_V + lane * SELL_V + q % SELL_V;
    int c;
    if (A.sell_col) c = A.sell_col[k];
    else if (A.sell_col16) c = sl * 64 + A.sell_col16[k];
    else if (A.sell_pid) c = rc + A.sell_dict[A.sell_dict_off[A.sell_pid[r]] + q];
    else c = r - r % A.sten_b + A.sten_d[q];
    c = min(max(c, 0), A.n - 1);
    v32[k] = (float)(v[k] * (si * isd[c]));
  }
}

// b = float(a * scale * isd)
__global__ void k_to_f32_scaled(int n, const double *__restrict__ a,
                                const double *__restrict__ isd, float *__restrict__ b,
                                double scale)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] * scale * isd[i]);
}

hipError_t inv_sqrt_diag(int n, const double *diag, double *isd, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_inv_sqrt_diag, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, s, n, diag, isd);
  return hipGetLastError();
}

hipError_t csr_scale32(int n, const int *rowptr, const int *col, const double *val,
                       const double *isd, float *val32, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_scale32, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, s, n, rowptr, col,
                     val, isd, val32);
  return hipGetLastError();
}

hipError_t mixed_operator(const CsrDev &A, const double *diag, const double *val, float *val32,
                          const double *sval, float *sval32, double *isd, hipStream_t s)
{
  const int n = A.n;
  if (n <= 0) return hipSuccess;
  hipError_t e = inv_sqrt_diag(n, diag, isd, s);
  if (e == hipSuccess && val && val32 && A.rowptr)
    e = csr_scale32(n, A.rowptr, A.col, val, isd, val32, s);
  if (e != hipSuccess) return e;
  if (sval && sval32 && A.nslice > 0)
    hipLaunchKernelGGL(k_sell_scale32, dim3((A.nslice * 64 + TPB - 1) / TPB), dim3(TPB), 0, s, A,
                       sval, isd, sval32);
  return hipGetLastError();
}

hipError_t mixed_rhs32(int n, const double *r64, const double *isd, float *r32, double scale,
                       hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32_scaled, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, s, n, r64, isd,
                     r32, scale);
  return hipGetLastError();
}

// x64 += S x32 / s, x32 = 0; block partials of ||x64||^2 into px
__global__ __launch_bounds__(TPB) void k_mx_xupdate(int n, const MixState *ms,
                                                    const double *__restrict__ isd,
                                                    double *__restrict__ x64,
                                                    float *__restrict__ x32,
                                                    double *__restrict__ px)
{
  if (ms->final) return;
  const double inv = 1.0 / ms->scale;
  int r0, r1;
  chunk(n, r0, r1);
  double xx = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double xi = x64[i] + (double)x32[i] * inv * isd[i];
    x64[i] = xi;
    x32[i] = 0.0f;
    xx += xi * xi;
  }
  const double t = block_sum(xx);
  if (threadIdx.x == 0) px[blockIdx.x] = t;
}

// The computed residual fl(b - A x) carries a rounding error of at most
// m eps (|b| + |A||x|) per row for m entries per row, typically sqrt(m) eps; in
// 2-norm about MIX_FLOOR (||b|| + ||A||_inf ||x||) with MIX_FLOOR = 64 eps. A
// solve is only declared stuck at the FP64 floor below that level.
constexpr double MIX_FLOOR = 64.0 * 2.220446049250313e-16;

__device__ __forceinline__ bool below_floor(double rr64, double nory64, double norm_a, double xx)
{
  const double f = MIX_FLOOR * (sqrt(nory64) + norm_a * sqrt(xx));
  return rr64 < f * f;
}

__device__ __forceinline__ double refresh_scale(double rr64, double old)
{
  return rr64 > 0.0 ? 1.0 / sqrt(rr64) : old;
}

// rs = float(s' S r64), d *= s'/s (or d = rs on a restart); block partials of
// rs.rs into prr[0, nb) and, as the recurrence's r.h, prr[nb, 2 nb)
__global__ __launch_bounds__(TPB) void k_mx_refresh(int n, const MixState *ms,
                                                    const double *__restrict__ part64, int nb64,
                                                    const double *__restrict__ r64,
                                                    const double *__restrict__ isd,
                                                    float *__restrict__ r32,
                                                    float *__restrict__ d32,
                                                    float *__restrict__ prr)
{
  if (ms->final) return;
  const double rr64 = reduce_partials(part64, nb64);
  const double s = refresh_scale(rr64, ms->scale);
  const float c = (float)(s / ms->scale);
  // a true residual 1e4 x above its best means the carried direction is spoiled:
  // restart it from the residual
  const bool restart = rr64 > 1e4 * ms->best;
  int r0, r1;
  chunk(n, r0, r1);
  float rr = 0.f;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const float ri = (float)(r64[i] * s * isd[i]);
    r32[i] = ri;
    d32[i] = restart ? ri : d32[i] * c;
    rr += ri * ri;
  }
  const float a = block_sum(rr);
  if (threadIdx.x == 0) {
    prr[blockIdx.x] = a;
    prr[gridDim.x + blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(TPB) void k_mx_check(MixState *ms, PcgState<float> *st,
                                                  const double *__restrict__ part64, int nb64,
                                                  const float *__restrict__ prr, int nv,
                                                  const double *__restrict__ px, double res)
{
  if (ms->final) return;
  const double rr64 = reduce_partials(part64, nb64);
  const float nomn = reduce_partials(prr + nv, nv);
  const double xx = reduce_partials(px, nv);
  if (threadIdx.x != 0) return;
  const double prev = ms->rr64;
  ms->rr64 = rr64;
  ms->updates += 1;
  if (rr64 < res * ms->nory64) {
    ms->final = 1;
    ms->status = 1;
    st->done = 1;
    return;
  }
  // FP64 accuracy floor above res: the residual is within the rounding floor
  // of its own computation, and either the FP32 recurrence met its reduction
  // three refreshes in a row while the FP64 residual did not even halve, or the
  // FP64 residual has not improved by 10 % in ten refreshes. (Far above that
  // floor the residual norm of CG may stall or grow for hundreds of iterations
  // on ill-conditioned systems while the error keeps falling.)
  const bool met = st->done && st->status == 1;
  ms->bad = (met && rr64 > 0.25 * prev) ? ms->bad + 1 : 0;
  if (rr64 < 0.81 * ms->best) ms->since_best = 0;
  else ms->since_best += 1;
  if (rr64 < ms->best) ms->best = rr64;
  if (below_floor(rr64, ms->nory64, ms->norm_a, xx) && (ms->bad >= 3 || ms->since_best >= 10)) {
    ms->final = 1;
    ms->status = 4;
    return;
  }
  ms->scale = refresh_scale(rr64, ms->scale);
  // next refresh once the FP32 residual falls by delta from its value now, or
  // to the final target if that is closer
  st->nory = (float)((double)nomn * fmax(ms->delta2, res * ms->nory64 / rr64));
  st->nom[(st->iter + 1) & 1] = nomn;
  st->done = 0;
  st->status = 0;
}

// after pcg_init of the FP32 recurrence (st->nory = rs.rs of its start)
__global__ void k_mx_init(MixState *ms, PcgState<float> *st, const double *__restrict__ slots,
                          const double *__restrict__ norm_a, double res, double delta)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ms->norm_a = *norm_a;
  ms->nory64 = slots[1];
  ms->rr64 = slots[0];
  ms->best = slots[0];
  ms->scale = refresh_scale(slots[0], 1.0);
  ms->delta2 = delta * delta;
  ms->bad = 0;
  ms->since_best = 0;
  ms->updates = 0;
  ms->final = 0;
  ms->status = 0;
  if (slots[0] < res * slots[1]) {
    ms->final = 1;
    ms->status = 1;
    st->done = 1;
    return;
  }
  st->nory = (float)((double)st->nory * fmax(ms->delta2, res * slots[1] / slots[0]));
}

hipError_t mixed_init(MixState *ms, PcgState<float> *st, const double *slots,
                      const double *norm_a, double res, double delta, hipStream_t s)
{
  hipLaunchKernelGGL(k_mx_init, dim3(1), dim3(64), 0, s, ms, st, slots, norm_a, res, delta);
  return hipGetLastError();
}

hipError_t mixed_update(const SpmvTune &t64, const SpmvTune &t32, const CsrDev &A,
                        const double *val64, const double *isd, const double *b, double *x64,
                        double *r64, double *part64, double *px, float *x32, float *r32,
                        float *d32, float *part32, PcgState<float> *st, MixState *ms,
                        double res, hipStream_t s)
{
  const int n = A.n;
  const int nv = vec_blocks(t32, n);
  hipLaunchKernelGGL(k_mx_xupdate, dim3(nv), dim3(TPB), 0, s, n, ms, isd, x64, x32, px);
  hipError_t e = spmv<double>(t64, 2, A, val64, x64, r64, b, part64, &ms->final, nullptr, s);
  if (e != hipSuccess) return e;
  float *prr = part32 + t32.blocks;
  hipLaunchKernelGGL(k_mx_refresh, dim3(nv), dim3(TPB), 0, s, n, ms, part64, t64.blocks, r64,
                     isd, r32, d32, prr);
  hipLaunchKernelGGL(k_mx_check, dim3(1), dim3(TPB), 0, s, ms, st, part64, t64.blocks, prr, nv,
                     px, res);
  return hipGetLastError();
}

// ||A||_inf of a CSR matrix: per-block maxima of the absolute row sums into
// partial[0, gridDim), then their maximum into *out (order-independent)
__global__ __launch_bounds__(TPB) void k_row_absmax(int n, const int *__restrict__ rowptr,
                                                    const double *__restrict__ val,
                                                    double *__restrict__ partial)
{
  int r0, r1;
  chunk(n, r0, r1);
  double m = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    double s = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) s += fabs(val[k]);
    m = fmax(m, s);
  }
  __shared__ double lds[TPB];
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ __launch_bounds__(TPB) void k_max_partials(const double *__restrict__ partial, int nb,
                                                      double *__restrict__ out)
{
  __shared__ double lds[TPB];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += TPB) m = fmax(m, partial[i]);
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

hipError_t row_absmax(int n, const int *rowptr, const double *val, double *partial, int nb,
                      double *out, hipStream_t s)
{
  if (n <= 0) return hipMemsetAsync(out, 0, sizeof(double), s);
  hipLaunchKernelGGL(k_row_absmax, dim3(nb), dim3(TPB), 0, s, n, rowptr, val, partial);
  hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(TPB), 0, s, partial, nb, out);
  return hipGetLastError();
}

// the same norm from the SELL-64 layout: a row's entries in column order (its
// padding adds zeros), so the value equals the CSR one bit for bit
__global__ __launch_bounds__(TPB) void k_sell_row_absmax(int n, const int *__restrict__ sell_off,
                                                         const double *__restrict__ v,
                                                         double *__restrict__ partial)
{
  int r0, r1;
  chunk(n, r0, r1);
  double m = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const int off = sell_off[i / 64], lane = i % 64;
    const int w = (sell_off[i / 64 + 1] - off) / 64;
    double s = 0.0;
    for (int q = 0; q < w; q++) s += fabs(v[off + (q / SELL_V) * 64 * SELL_V + lane * SELL_V + q % SELL_V]);
    m = fmax(m, s);
  }
  __shared__ double lds[TPB];
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

hipError_t sell_row_absmax(int n, const int *sell_off, const double *v, double *partial, int nb,
                           double *out, hipStream_t s)
{
  if (n <= 0) return hipMemsetAsync(out, 0, sizeof(double), s);
  hipLaunchKernelGGL(k_sell_row_absmax, dim3(nb), dim3(TPB), 0, s, n, sell_off, v, partial);
  hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(TPB), 0, s, partial, nb, out);
  return hipGetLastError();
}

}  // namespace ssf
