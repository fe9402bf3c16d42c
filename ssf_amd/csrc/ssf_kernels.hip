// gfx950 kernels of the SIFEL linear-system path: CSR SpMV with fused dots, the
// preconditioned CG recurrence of GEFEL/scr.cpp:1093 (and its cg / comprow
// variants), reference-layout <-> full-CSR value maps, and the batched element
// kernels with atomic-free coloured assembly.
//
// Reductions are deterministic: every block writes one partial per quantity, and
// every consumer block re-reduces the whole partial array in the same fixed order,
// so each block holds bit-identical scalars without inter-block synchronisation
// and a system gives the same result whatever else runs on the device.
#include "ssf_kernels.hpp"

namespace ssf {

// ----------------------------------------------------------------- helpers

// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD works
// on one contiguous eighth of the rows and the x-vector window of neighbouring
// rows stays in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
  return ((nb & 7) == 0) ? (b & 7) * (nb >> 3) + (b >> 3) : b;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the block; the result is valid in every thread. Fixed order.
template <typename T>
__device__ T block_sum(T v)
{
  __shared__ T lds[TPB / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = lds[0];
#pragma unroll
  for (int i = 1; i < TPB / 64; i++) s += lds[i];
  __syncthreads();
  return s;
}

template <typename T>
__device__ T reduce_partials(const T *__restrict__ p, int nb)
{
  T s = T(0);
  for (int i = threadIdx.x; i < nb; i += TPB) s += p[i];
  return block_sum(s);
}

template <typename T>
__device__ __forceinline__ T ldnt(const T *p)
{
  return __builtin_nontemporal_load(p);
}

// dot of one CSR row with x by G lanes; the sum is valid in the group's lane 0
template <typename T, int G>
__device__ __forceinline__ T row_dot(const int *__restrict__ rowptr, const int *__restrict__ col,
                                     const T *__restrict__ val, const T *__restrict__ x,
                                     int row, int lane)
{
  const int beg = rowptr[row], end = rowptr[row + 1];
  T s = T(0);
  for (int k = beg + lane; k < end; k += G) s += ldnt(val + k) * x[ldnt(col + k)];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
  return s;
}

__device__ __forceinline__ void chunk(int n, int &r0, int &r1)
{
  const int nb = gridDim.x;
  const int rpb = (n + nb - 1) / nb;
  const int lb = xcd_block(blockIdx.x, nb);
  r0 = lb * rpb;
  r1 = min(n, r0 + rpb);
}

// ------------------------------------------------------------------- SpMV

template <typename T, int G, int KIND>
__global__ __launch_bounds__(TPB) void k_spmv(int n, const int *__restrict__ rowptr,
                                              const int *__restrict__ col,
                                              const T *__restrict__ val,
                                              const T *__restrict__ x, T *__restrict__ y,
                                              const T *__restrict__ b,
                                              T *__restrict__ partial, const int *done,
                                              long long *iter)
{
  if (done && *done) return;
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1;
  int r0, r1;
  chunk(n, r0, r1);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  constexpr int NG = TPB / G;
  T acc = T(0);
  for (int row = r0 + grp; row < r1; row += NG) {
    T s = row_dot<T, G>(rowptr, col, val, x, row, lane);
    if (lane == 0) {
      if (KIND == 2) {
        s = b[row] - s;
        acc += s * s;
      } else if (KIND == 1) {
        acc += x[row] * s;
      }
      y[row] = s;
    }
  }
  if (KIND != 0) {
    const T t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

template <typename T, int KIND>
static hipError_t spmv_g(const SpmvTune &t, int n, const int *rowptr, const int *col,
                         const T *val, const T *x, T *y, const T *b, T *partial,
                         const int *done, long long *iter, hipStream_t s)
{
  dim3 grid(t.blocks), block(TPB);
  switch (t.lanes) {
    case 4: hipLaunchKernelGGL((k_spmv<T, 4, KIND>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, done, iter); break;
    case 8: hipLaunchKernelGGL((k_spmv<T, 8, KIND>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, done, iter); break;
    case 16: hipLaunchKernelGGL((k_spmv<T, 16, KIND>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, done, iter); break;
    case 32: hipLaunchKernelGGL((k_spmv<T, 32, KIND>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, done, iter); break;
    default: hipLaunchKernelGGL((k_spmv<T, 64, KIND>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, done, iter); break;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t spmv(const SpmvTune &t, int kind, int n, const int *rowptr, const int *col,
                const T *val, const T *x, T *y, const T *b, T *partial, const int *done,
                long long *iter, hipStream_t s)
{
  if (kind == 0) return spmv_g<T, 0>(t, n, rowptr, col, val, x, y, b, partial, done, iter, s);
  if (kind == 1) return spmv_g<T, 1>(t, n, rowptr, col, val, x, y, b, partial, done, iter, s);
  return spmv_g<T, 2>(t, n, rowptr, col, val, x, y, b, partial, done, iter, s);
}
template hipError_t spmv<double>(const SpmvTune &, int, int, const int *, const int *,
                                 const double *, const double *, double *, const double *,
                                 double *, const int *, long long *, hipStream_t);
template hipError_t spmv<float>(const SpmvTune &, int, int, const int *, const int *,
                                const float *, const float *, float *, const float *,
                                float *, const int *, long long *, hipStream_t);

// --------------------------------------------------------------------- PCG
//
// Per iteration i (GEFEL/scr.cpp:1142-1197):
//   K1  p = A d, partial d.p                      (k_spmv KIND 1, increments iter)
//   K2  den = d.p; |den| < zero -> stop (status 2)  x += a d, r -= a p, h = M^-1 r,
//       partials r.r and r.h                       (k_pcg_update)
//   K3  rr = r.r, nom' = r.h, stopping test of the mode, else d = b d + h
//                                                  (k_pcg_dir)
// The stop of K2/K3 is decided identically by every block from the same
// partials, so no block continues once another would have stopped.

template <typename T>
__device__ __forceinline__ T jdiv(T r, T dg);
template <>
__device__ __forceinline__ double jdiv<double>(double r, double dg) { return r / dg; }
template <>
__device__ __forceinline__ float jdiv<float>(float r, float dg) { return r / dg; }

template <typename T, int G, bool JAC, bool HAVEX>
__global__ __launch_bounds__(TPB) void k_pcg_init(int n, const int *__restrict__ rowptr,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val,
                                                  const T *__restrict__ diag,
                                                  const T *__restrict__ x,
                                                  const T *__restrict__ y, T *__restrict__ r,
                                                  T *__restrict__ h, T *__restrict__ d,
                                                  T *__restrict__ part)
{
  int r0, r1;
  chunk(n, r0, r1);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  constexpr int NG = TPB / G;
  T ayy = T(0), arh = T(0);
  for (int row = r0 + grp; row < r1; row += NG) {
    T s = HAVEX ? row_dot<T, G>(rowptr, col, val, x, row, lane) : T(0);
    if (lane == 0) {
      const T yi = y[row];
      const T ri = yi - s;
      const T hi = JAC ? jdiv<T>(ri, diag[row]) : ri;
      r[row] = ri;
      if (JAC) h[row] = hi;
      d[row] = hi;
      ayy += yi * yi;
      arh += ri * hi;
    }
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

template <typename T, bool JAC>
__global__ __launch_bounds__(TPB) void k_pcg_update(int n, PcgState<T> *st,
                                                    const T *__restrict__ pden, T zero,
                                                    T *__restrict__ x, T *__restrict__ r,
                                                    const T *__restrict__ d,
                                                    const T *__restrict__ p,
                                                    const T *__restrict__ diag,
                                                    T *__restrict__ h, T *__restrict__ prr)
{
  if (st->done) return;
  const int nb = gridDim.x;
  const T den = reduce_partials(pden, nb);
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
  int r0, r1;
  chunk(n, r0, r1);
  T rr = T(0), rh = T(0);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    x[i] += alpha * d[i];
    const T ri = r[i] - alpha * p[i];
    r[i] = ri;
    rr += ri * ri;
    if (JAC) {
      const T hi = jdiv<T>(ri, diag[i]);
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

template <typename T, bool JAC, int MODE>
__global__ __launch_bounds__(TPB) void k_pcg_dir(int n, PcgState<T> *st,
                                                 const T *__restrict__ prr, T res, T zero,
                                                 T *__restrict__ d, const T *__restrict__ h)
{
  if (st->done) return;
  const int nb = gridDim.x;
  const T rr = reduce_partials(prr, nb);
  const T nomn = JAC ? reduce_partials(prr + nb, nb) : rr;
  const long long it = st->iter;
  const T nom = st->nom[it & 1];
  const T nory = st->nory;
  bool stop;
  if (MODE == MODE_SCR_CG_PREC) stop = fabs(rr / nory) < res;                 // scr.cpp:1188
  else if (MODE == MODE_SCR_CG) stop = (sqrt(nomn / nory) < res) || (fabs(nomn) < zero); // :862-870
  else stop = sqrt(nomn / nory) < res;                                        // cr.cpp:1217
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
  const T beta = nomn / nom;
  int r0, r1;
  chunk(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) d[i] = beta * d[i] + h[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->nom[(it + 1) & 1] = nomn;
}

template <typename T, bool JAC, bool HAVEX>
static void init_g(const SpmvTune &t, int n, const int *rowptr, const int *col, const T *val,
                   const T *diag, const T *x, const T *y, T *r, T *h, T *d, T *part,
                   hipStream_t s)
{
  dim3 grid(t.blocks), block(TPB);
  switch (t.lanes) {
    case 4: hipLaunchKernelGGL((k_pcg_init<T, 4, JAC, HAVEX>), grid, block, 0, s, n, rowptr, col, val, diag, x, y, r, h, d, part); break;
    case 8: hipLaunchKernelGGL((k_pcg_init<T, 8, JAC, HAVEX>), grid, block, 0, s, n, rowptr, col, val, diag, x, y, r, h, d, part); break;
    case 16: hipLaunchKernelGGL((k_pcg_init<T, 16, JAC, HAVEX>), grid, block, 0, s, n, rowptr, col, val, diag, x, y, r, h, d, part); break;
    case 32: hipLaunchKernelGGL((k_pcg_init<T, 32, JAC, HAVEX>), grid, block, 0, s, n, rowptr, col, val, diag, x, y, r, h, d, part); break;
    default: hipLaunchKernelGGL((k_pcg_init<T, 64, JAC, HAVEX>), grid, block, 0, s, n, rowptr, col, val, diag, x, y, r, h, d, part); break;
  }
}

template <typename T>
hipError_t pcg_init(const SpmvTune &t, bool jac, bool havex, int n, const int *rowptr,
                    const int *col, const T *val, const T *diag, const T *x, const T *y,
                    T *r, T *h, T *d, T *part, PcgState<T> *st, T zero, hipStream_t s)
{
  if (jac && havex) init_g<T, true, true>(t, n, rowptr, col, val, diag, x, y, r, h, d, part, s);
  else if (jac) init_g<T, true, false>(t, n, rowptr, col, val, diag, x, y, r, h, d, part, s);
  else if (havex) init_g<T, false, true>(t, n, rowptr, col, val, diag, x, y, r, h, d, part, s);
  else init_g<T, false, false>(t, n, rowptr, col, val, diag, x, y, r, h, d, part, s);
  hipLaunchKernelGGL((k_pcg_init2<T>), dim3(1), dim3(TPB), 0, s, part, t.blocks, st, zero);
  return hipGetLastError();
}
template hipError_t pcg_init<double>(const SpmvTune &, bool, bool, int, const int *, const int *,
                                     const double *, const double *, const double *,
                                     const double *, double *, double *, double *, double *,
                                     PcgState<double> *, double, hipStream_t);
template hipError_t pcg_init<float>(const SpmvTune &, bool, bool, int, const int *, const int *,
                                    const float *, const float *, const float *, const float *,
                                    float *, float *, float *, float *, PcgState<float> *, float,
                                    hipStream_t);

template <typename T>
hipError_t pcg_iteration(const SpmvTune &t, int mode, bool jac, int n, const int *rowptr,
                         const int *col, const T *val, const T *diag, T *x, T *r, T *d,
                         T *p, T *h, T *part, PcgState<T> *st, T res, T zero, hipStream_t s)
{
  hipError_t e = spmv<T>(t, 1, n, rowptr, col, val, d, p, nullptr, part, &st->done, &st->iter, s);
  if (e != hipSuccess) return e;
  dim3 grid(t.blocks), block(TPB);
  T *prr = part + t.blocks;
  if (jac) hipLaunchKernelGGL((k_pcg_update<T, true>), grid, block, 0, s, n, st, part, zero, x, r, d, p, diag, h, prr);
  else hipLaunchKernelGGL((k_pcg_update<T, false>), grid, block, 0, s, n, st, part, zero, x, r, d, p, diag, h, prr);
  const T *hh = jac ? h : r;
  if (jac) {
    if (mode == MODE_SCR_CG_PREC) hipLaunchKernelGGL((k_pcg_dir<T, true, MODE_SCR_CG_PREC>), grid, block, 0, s, n, st, prr, res, zero, d, hh);
    else if (mode == MODE_SCR_CG) hipLaunchKernelGGL((k_pcg_dir<T, true, MODE_SCR_CG>), grid, block, 0, s, n, st, prr, res, zero, d, hh);
    else hipLaunchKernelGGL((k_pcg_dir<T, true, MODE_CR_CG_PREC>), grid, block, 0, s, n, st, prr, res, zero, d, hh);
  } else {
    if (mode == MODE_SCR_CG_PREC) hipLaunchKernelGGL((k_pcg_dir<T, false, MODE_SCR_CG_PREC>), grid, block, 0, s, n, st, prr, res, zero, d, hh);
    else if (mode == MODE_SCR_CG) hipLaunchKernelGGL((k_pcg_dir<T, false, MODE_SCR_CG>), grid, block, 0, s, n, st, prr, res, zero, d, hh);
    else hipLaunchKernelGGL((k_pcg_dir<T, false, MODE_CR_CG_PREC>), grid, block, 0, s, n, st, prr, res, zero, d, hh);
  }
  return hipGetLastError();
}
template hipError_t pcg_iteration<double>(const SpmvTune &, int, bool, int, const int *,
                                          const int *, const double *, const double *,
                                          double *, double *, double *, double *, double *,
                                          double *, PcgState<double> *, double, double,
                                          hipStream_t);
template hipError_t pcg_iteration<float>(const SpmvTune &, int, bool, int, const int *,
                                         const int *, const float *, const float *, float *,
                                         float *, float *, float *, float *, float *,
                                         PcgState<float> *, float, float, hipStream_t);

// ------------------------------------------------------- element-wise helpers

__global__ void k_expand_ref(int64_t negm, const int *__restrict__ lo, const int *__restrict__ up,
                             const double *__restrict__ a, double *__restrict__ val)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= negm) return;
  const double v = a[k];
  val[lo[k]] = v;
  const int u = up[k];
  if (u >= 0) val[u] = v;
}

hipError_t expand_ref(int64_t negm, const int *map_lo, const int *map_up, const double *a,
                      double *val, hipStream_t s)
{
  if (negm == 0) return hipSuccess;
  const int64_t nb = (negm + TPB - 1) / TPB;
  hipLaunchKernelGGL(k_expand_ref, dim3((unsigned)nb), dim3(TPB), 0, s, negm, map_lo, map_up, a, val);
  return hipGetLastError();
}

__global__ void k_extract_diag(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                               const double *__restrict__ val, double *__restrict__ diag,
                               float *__restrict__ diag32)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // columns ascending: binary search for the diagonal
  int lo = rowptr[i], hi = rowptr[i + 1] - 1;
  double v = 0.0;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = col[mid];
    if (c == i) {
      v = val[mid];
      break;
    }
    if (c < i) lo = mid + 1;
    else hi = mid - 1;
  }
  diag[i] = v;
  if (diag32) diag32[i] = (float)v;
}

hipError_t extract_diag(int n, const int *rowptr, const int *col, const double *val,
                        double *diag, float *diag32, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_extract_diag, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, s, n, rowptr, col, val, diag, diag32);
  return hipGetLastError();
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b, double scale)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] * scale);
}

hipError_t to_f32(int64_t n, const double *a, float *b, double scale, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, a, b, scale);
  return hipGetLastError();
}

__global__ void k_axpy_f32(int64_t n, double alpha, const float *__restrict__ e, double *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += alpha * (double)e[i];
}

hipError_t axpy_f32(int64_t n, double alpha, const float *e, double *x, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy_f32, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, alpha, e, x);
  return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_dot_partial(int64_t n, const double *__restrict__ a,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ part)
{
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    acc += a[i] * b[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

hipError_t dot_partial(int64_t n, const double *a, const double *b, double *part, int nb,
                       hipStream_t s)
{
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(TPB), 0, s, n, a, b, part);
  return hipGetLastError();
}

__global__ void k_fill(int64_t n, double v, double *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

hipError_t fill(int64_t n, double v, double *x, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, v, x);
  return hipGetLastError();
}

__global__ void k_reduce_slot(const double *__restrict__ part, int nb, double *out)
{
  const double v = reduce_partials(part, nb);
  if (threadIdx.x == 0) *out = v;
}

hipError_t reduce_to_host_slot(const double *part, int nb, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce_slot, dim3(1), dim3(TPB), 0, s, part, nb, out);
  return hipGetLastError();
}

// ----------------------------------------------------------------- assembly

// slot[(i*nd + j) * ne_total + p] (aos: slot[p * nd*nd + i*nd + j]) = reference-layout
// entry of element order[p]'s
// local pair (i, j), or -1 where localize skips it (code number <= 0, or the
// upper triangle of symcomprow: GEFEL/scr.cpp:427-433).
__global__ void k_build_slots(int storage, int aos, int64_t ne_total, int nd,
                              const int *__restrict__ order, const int *__restrict__ cn,
                              const int *__restrict__ adr, const int *__restrict__ ci,
                              int *__restrict__ slot)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nq = (int64_t)nd * nd;
  const int64_t total = ne_total * nq;
  if (t >= total) return;
  const int64_t p = aos ? t / nq : t % ne_total;
  const int ij = aos ? (int)(t % nq) : (int)(t / ne_total);
  const int i = ij / nd, j = ij % nd;
  const int e = order[p];
  const int ii = cn[(int64_t)e * nd + i] - 1;
  const int jj = cn[(int64_t)e * nd + j] - 1;
  int sl = -1;
  if (ii >= 0 && jj >= 0 && !(storage == SSF_STORAGE_SYMM_COMP_ROWS && ii < jj)) {
    int lo = adr[ii], hi = adr[ii + 1] - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int c = ci[mid];
      if (c == jj) {
        sl = mid;
        break;
      }
      if (c < jj) lo = mid + 1;
      else hi = mid - 1;
    }
  }
  slot[t] = sl;
}

hipError_t build_slots(int storage, int aos, int64_t ne_total, int nd, const int *order,
                       const int *cn, const int *ref_adr, const int *ref_ci, int *slot,
                       hipStream_t s)
{
  const int64_t total = ne_total * nd * nd;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_build_slots, dim3((unsigned)((total + TPB - 1) / TPB)), dim3(TPB), 0, s,
                     storage, aos, ne_total, nd, order, cn, ref_adr, ref_ci, slot);
  return hipGetLastError();
}

// natural derivatives of the trilinear hexahedron (GEFEL/basefun.cpp:2317-2375)
__device__ __forceinline__ void hex_nat(double xi, double eta, double zeta, double *nx,
                                        double *ny, double *nz)
{
  const double xp = 1.0 + xi, xm = 1.0 - xi, yp = 1.0 + eta, ym = 1.0 - eta;
  const double zp = 1.0 + zeta, zm = 1.0 - zeta;
  nx[0] = yp * zp * 0.125; nx[1] = -(yp * zp * 0.125); nx[2] = -(ym * zp * 0.125); nx[3] = ym * zp * 0.125;
  nx[4] = yp * zm * 0.125; nx[5] = -(yp * zm * 0.125); nx[6] = -(ym * zm * 0.125); nx[7] = ym * zm * 0.125;
  ny[0] = xp * zp * 0.125; ny[1] = xm * zp * 0.125; ny[2] = -(xm * zp * 0.125); ny[3] = -(xp * zp * 0.125);
  ny[4] = xp * zm * 0.125; ny[5] = xm * zm * 0.125; ny[6] = -(xm * zm * 0.125); ny[7] = -(xp * zm * 0.125);
  nz[0] = xp * yp * 0.125; nz[1] = xm * yp * 0.125; nz[2] = xm * ym * 0.125; nz[3] = xp * ym * 0.125;
  nz[4] = -(xp * yp * 0.125); nz[5] = -(xm * yp * 0.125); nz[6] = -(xm * ym * 0.125); nz[7] = -(xp * ym * 0.125);
}

// global derivatives and |J| (GEFEL/difcalc.cpp:487-571); |J| sign fixed as
// MEFEL/SRC/linhex.cpp:619-623 / TRFEL/SRC/linhext.cpp:809-812
__device__ __forceinline__ double hex_global(const double *x, const double *y, const double *z,
                                             double xi, double eta, double zeta, double *dx,
                                             double *dy, double *dz)
{
  double nx[8], ny[8], nz[8];
  hex_nat(xi, eta, zeta, nx, ny, nz);
  double d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0, d7 = 0, d8 = 0, d9 = 0;
#pragma unroll
  for (int a = 0; a < 8; a++) {
    d1 += nx[a] * x[a]; d2 += ny[a] * x[a]; d3 += nz[a] * x[a];
    d4 += nx[a] * y[a]; d5 += ny[a] * y[a]; d6 += nz[a] * y[a];
    d7 += nx[a] * z[a]; d8 += ny[a] * z[a]; d9 += nz[a] * z[a];
  }
  const double jac = d1 * d5 * d9 + d4 * d8 * d3 + d7 * d2 * d6 - d7 * d5 * d3 - d8 * d6 * d1 -
                     d9 * d4 * d2;
  const double ij = 1.0 / jac;
  const double c00 = d5 * d9 - d8 * d6, c01 = d7 * d6 - d9 * d4, c02 = d4 * d8 - d7 * d5;
  const double c10 = d8 * d3 - d9 * d2, c11 = d1 * d9 - d7 * d3, c12 = d7 * d2 - d8 * d1;
  const double c20 = d2 * d6 - d5 * d3, c21 = d4 * d3 - d6 * d1, c22 = d1 * d5 - d4 * d2;
#pragma unroll
  for (int a = 0; a < 8; a++) {
    dx[a] = (nx[a] * c00 + ny[a] * c01 + nz[a] * c02) * ij;
    dy[a] = (nx[a] * c10 + ny[a] * c11 + nz[a] * c12) * ij;
    dz[a] = (nx[a] * c20 + ny[a] * c21 + nz[a] * c22) * ij;
  }
  return fabs(jac);
}

#define GPT 0.577350269189626  /* GEFEL/intp.cpp:26-27, weights 1 */

// hex8 heat conduction (TRFEL/SRC/linhext.cpp:774-840, isotrmat::matcond3d):
// K_e = sum_gp k |J| B^T B with B = grad N (3 x 8).
__device__ void ke_hex8_heat(const double *x, const double *y, const double *z, double kc,
                             double *K)
{
#pragma unroll
  for (int q = 0; q < 64; q++) K[q] = 0.0;
  for (int g = 0; g < 8; g++) {
    const double xi = (g & 4) ? GPT : -GPT, eta = (g & 2) ? GPT : -GPT, zeta = (g & 1) ? GPT : -GPT;
    double dx[8], dy[8], dz[8];
    const double jac = hex_global(x, y, z, xi, eta, zeta, dx, dy, dz);
    const double w = kc * jac;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = i; j < 8; j++) {
        const double s = (dx[i] * dx[j] + dy[i] * dy[j] + dz[i] * dz[j]) * w;
        K[i * 8 + j] += s;
        if (j != i) K[j * 8 + i] += s;
      }
  }
}

// quad4 plane stress (MEFEL/SRC/plelemlq.cpp:611-690, elastisomat::matstiff_plstress)
__device__ void ke_quad4_pstress(const double *x, const double *y, double e, double nu,
                                 double t, double *K)
{
  const double c = e / (1.0 - nu * nu);
  const double d00 = c, d01 = c * nu, d22 = e / 2.0 / (1.0 + nu);
#pragma unroll
  for (int q = 0; q < 64; q++) K[q] = 0.0;
  for (int g = 0; g < 4; g++) {
    const double xi = (g & 2) ? GPT : -GPT, eta = (g & 1) ? GPT : -GPT;
    double nx[4], ny[4];
    nx[0] = 0.25 * (1.0 + eta); nx[1] = 0.25 * (-1.0 - eta); nx[2] = 0.25 * (-1.0 + eta); nx[3] = 0.25 * (1.0 - eta);
    ny[0] = 0.25 * (1.0 + xi); ny[1] = 0.25 * (1.0 - xi); ny[2] = 0.25 * (-1.0 + xi); ny[3] = 0.25 * (-1.0 - xi);
    double d1 = 0, d2 = 0, d3 = 0, d4 = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      d1 += nx[a] * x[a]; d2 += ny[a] * x[a]; d3 += nx[a] * y[a]; d4 += ny[a] * y[a];
    }
    const double jac = d1 * d4 - d2 * d3;
    const double ij = 1.0 / jac;
    double bx[4], by[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      bx[a] = (nx[a] * d4 - ny[a] * d3) * ij;
      by[a] = (ny[a] * d1 - nx[a] * d2) * ij;
    }
    const double w = fabs(jac * t);  // thickness approx: bilinear of equal nodal values
    // B = [[bx0 0 bx1 0 ..], [0 by0 0 by1 ..], [by0 bx0 by1 bx1 ..]]
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        // 2x2 block (a,b) of B^T D B
        const double kxx = bx[a] * d00 * bx[b] + by[a] * d22 * by[b];
        const double kxy = bx[a] * d01 * by[b] + by[a] * d22 * bx[b];
        const double kyx = by[a] * d01 * bx[b] + bx[a] * d22 * by[b];
        const double kyy = by[a] * d00 * by[b] + bx[a] * d22 * bx[b];
        K[(2 * a) * 8 + 2 * b] += kxx * w;
        K[(2 * a) * 8 + 2 * b + 1] += kxy * w;
        K[(2 * a + 1) * 8 + 2 * b] += kyx * w;
        K[(2 * a + 1) * 8 + 2 * b + 1] += kyy * w;
      }
  }
}

// one thread per element of a colour; adds K_e into the reference-layout values
// through the slot map, and -K_e u_p into the Dirichlet lift (colour-exclusive).
template <int KIND>
__global__ __launch_bounds__(TPB) void k_assemble8(int64_t p0, int64_t p1, int64_t ne_total,
                                                   const int *__restrict__ order,
                                                   const int *__restrict__ conn,
                                                   const int *__restrict__ cn,
                                                   const double *__restrict__ xyz,
                                                   const double *__restrict__ ma,
                                                   const double *__restrict__ mb,
                                                   const double *__restrict__ th,
                                                   const int *__restrict__ slot,
                                                   double *__restrict__ a_ref,
                                                   const double *__restrict__ pv,
                                                   double *__restrict__ lift)
{
  const int64_t p = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int e = order[p];
  double K[64];
  if (KIND == SSF_ELEM_HEX8_HEAT) {
    double x[8], y[8], z[8];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int nd = conn[(int64_t)e * 8 + a];
      x[a] = xyz[3 * (int64_t)nd]; y[a] = xyz[3 * (int64_t)nd + 1]; z[a] = xyz[3 * (int64_t)nd + 2];
    }
    ke_hex8_heat(x, y, z, ma[e], K);
  } else {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int nd = conn[(int64_t)e * 4 + a];
      x[a] = xyz[3 * (int64_t)nd]; y[a] = xyz[3 * (int64_t)nd + 1];
    }
    ke_quad4_pstress(x, y, ma[e], mb[e], th ? th[e] : 1.0, K);
  }
#pragma unroll
  for (int q = 0; q < 64; q++) {
    const int s = slot[(int64_t)q * ne_total + p];
    if (s >= 0) a_ref[s] += K[q];
  }
  if (lift) {
    int c[8];
    bool any = false;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      c[i] = cn[(int64_t)e * 8 + i];
      any |= c[i] < 0;
    }
    if (any) {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (c[i] < 1) continue;
        double f = 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (c[j] < 0) f += K[i * 8 + j] * pv[-c[j] - 1];
        lift[c[i] - 1] -= f;
      }
    }
  }
}

hipError_t assemble_hex8_elast(int64_t p0, int64_t p1, int64_t ne_total, const int *order,
                               const int *conn, const int *cn, const double *xyz,
                               const double *ma, const double *mb, const int *slot,
                               double *a_ref, const double *pv, double *lift, hipStream_t s);

hipError_t assemble_elements(int kind, int64_t p0, int64_t p1, int64_t ne_total,
                             const int *order, const int *conn, const int *cn,
                             const double *xyz, const double *mat_a, const double *mat_b,
                             const double *thick, const int *slot, double *a_ref,
                             const double *pv, double *lift, hipStream_t s)
{
  if (p1 <= p0) return hipSuccess;
  if (kind == SSF_ELEM_HEX8_ELAST)
    return assemble_hex8_elast(p0, p1, ne_total, order, conn, cn, xyz, mat_a, mat_b, slot,
                               a_ref, pv, lift, s);
  const unsigned nb = (unsigned)((p1 - p0 + TPB - 1) / TPB);
  if (kind == SSF_ELEM_HEX8_HEAT)
    hipLaunchKernelGGL((k_assemble8<SSF_ELEM_HEX8_HEAT>), dim3(nb), dim3(TPB), 0, s, p0, p1, ne_total, order, conn, cn, xyz, mat_a, mat_b, thick, slot, a_ref, pv, lift);
  else
    hipLaunchKernelGGL((k_assemble8<SSF_ELEM_QUAD4_PSTRESS>), dim3(nb), dim3(TPB), 0, s, p0, p1, ne_total, order, conn, cn, xyz, mat_a, mat_b, thick, slot, a_ref, pv, lift);
  return hipGetLastError();
}

__global__ void k_localize(int64_t p0, int64_t p1, int64_t ne_total, int nd,
                           const int *__restrict__ order, const int *__restrict__ slot,
                           const double *__restrict__ lm, double *__restrict__ a_ref)
{
  const int64_t p = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  const int e = order[p];
  const int nq = nd * nd;
  for (int q = 0; q < nq; q++) {
    const int s = slot[(int64_t)q * ne_total + p];
    if (s >= 0) a_ref[s] += lm[(int64_t)e * nq + q];
  }
}

hipError_t localize_elements(int64_t p0, int64_t p1, int64_t ne_total, int nd, const int *order,
                             const int *slot, const double *lm, double *a_ref, hipStream_t s)
{
  if (p1 <= p0) return hipSuccess;
  hipLaunchKernelGGL(k_localize, dim3((unsigned)((p1 - p0 + TPB - 1) / TPB)), dim3(TPB), 0, s,
                     p0, p1, ne_total, nd, order, slot, lm, a_ref);
  return hipGetLastError();
}

}  // namespace ssf
