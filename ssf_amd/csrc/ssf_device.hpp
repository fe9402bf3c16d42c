// Device helpers shared by the gfx950 kernels of libssf_gpu.
#pragma once

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

// Orders LDS accesses between the lanes of one wavefront (data that lanes of
// this wavefront wrote and other lanes of it read): the wavefront's LDS
// operations execute in order, so it waits for its own outstanding LDS operations
// only (s_waitcnt lgkmcnt(0)); global loads and stores stay in flight. The asm
// statements keep the compiler from moving memory accesses across the point.
__device__ __forceinline__ void wave_lds_sync()
{
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // gfx9 encoding: vmcnt 63, expcnt 7, lgkmcnt 0
  asm volatile("" ::: "memory");
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

// an element at a 32-bit byte offset from a base shared by the wavefront (the
// base stays in scalar registers, the offset takes one vector register)
template <typename U>
__device__ __forceinline__ const U *at_byte(const void *base, unsigned off)
{
  return (const U *)((const char *)base + off);
}

// NG groups of SELL_V slots of one row in the block stencil form: all value
// loads, then all gathers (slot k at column base + so[k], clamped into range),
// then the products in slot (= column) order. vb: byte offset of the first
// group's values in sval; so: the groups' offsets (wave-uniform: scalar loads).
template <typename T, int NG>
__device__ __forceinline__ T sten_groups(const T *__restrict__ sval, unsigned vb,
                                         const int *__restrict__ so, const T *__restrict__ x,
                                         int base, int nm1, T sum)
{
  typedef T tv __attribute__((ext_vector_type(SELL_V)));
  tv v[NG];
  T g[NG][SELL_V];
#pragma unroll
  for (int k = 0; k < NG; k++) v[k] = ldnt(at_byte<tv>(sval, vb + k * 64 * SELL_V * sizeof(T)));
#pragma unroll
  for (int k = 0; k < NG; k++)
#pragma unroll
    for (int u = 0; u < SELL_V; u++) {
      const int c = min(max(base + so[k * SELL_V + u], 0), nm1);
      g[k][u] = *at_byte<T>(x, (unsigned)c * sizeof(T));
    }
#pragma unroll
  for (int k = 0; k < NG; k++)
#pragma unroll
    for (int u = 0; u < SELL_V; u++) sum += v[k][u] * g[k][u];
  return sum;
}


// Row r of the symmetric half of the stencil form (CsrDev::half_*): NL lower
// slots at offsets off(k) (ascending, negative; off is an accessor, so offsets
// held in kernel arguments stay in scalar registers), stored diagonal-major in
// hv[k pad + r], then the diagonal at k = NL; the upper entry a(r, r - off(k)) is
// slot k of row r - off(k). The sum runs in column order (lower slots ascending,
// the diagonal, the upper ones ascending), all loads first. xr = x[min(r, n-1)].
template <typename T, int NL, typename Off>
__device__ __forceinline__ T half_row(const T *__restrict__ hv, int64_t pad, Off off,
                                      const T *__restrict__ x, int r, int nm1, T &xr)
{
  T vo[NL + 1], vm[NL], gl[NL], gu[NL];
#pragma unroll
  for (int k = 0; k <= NL; k++) vo[k] = hv[k * pad + r];
#pragma unroll
  for (int k = 0; k < NL; k++) vm[k] = hv[k * pad + (r - off(k))];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    gl[k] = x[min(max(r + off(k), 0), nm1)];
    gu[k] = x[min(r - off(k), nm1)];
  }
  xr = x[min(r, nm1)];
  T sum = T(0);
#pragma unroll
  for (int k = 0; k < NL; k++) sum += vo[k] * gl[k];
  sum += vo[NL] * xr;
#pragma unroll
  for (int k = NL - 1; k >= 0; k--) sum += vm[k] * gu[k];
  return sum;
}

// Node q of the block variant (B unknowns per node, NLN lower neighbour nodes q +
// m(j), m an accessor as above): the B row sums of the node in column order
// (lower nodes ascending, the diagonal block, the upper nodes ascending; within
// a node its unknowns in order) into sum, x of the node's own unknowns into xd.
// Slice-major storage: entry i of node 64 s + l at hv[(s NV + i) 64 + l], NV =
// NLN B^2 + B (B + 1) / 2, block j's entry (t, e) at i = (j B + t) B + e, the
// diagonal block's (t, e <= t) at NLN B^2 + t (t + 1) / 2 + e; the upper block
// towards q - m(j) is the transpose of block j of that node.
template <typename T, int B, int NLN, typename Off>
__device__ __forceinline__ void half_node(const T *__restrict__ hv, Off m,
                                          const T *__restrict__ x, int q, int nn, T (&sum)[B],
                                          T (&xd)[B])
{
  constexpr int ND = NLN * B * B, NV = ND + B * (B + 1) / 2;
  const T *__restrict__ own = hv + (int64_t)(q >> 6) * NV * 64 + (q & 63);
#pragma unroll
  for (int t = 0; t < B; t++) sum[t] = T(0);
#pragma unroll
  for (int j = 0; j < NLN; j++) {  // lower neighbour nodes
    const int c = min(max(q + m(j), 0), nn - 1);
    T xc[B];
#pragma unroll
    for (int e = 0; e < B; e++) xc[e] = x[B * c + e];
#pragma unroll
    for (int t = 0; t < B; t++)
#pragma unroll
      for (int e = 0; e < B; e++) sum[t] += own[((j * B + t) * B + e) * 64] * xc[e];
  }
  const int cd = min(q, nn - 1);
#pragma unroll
  for (int e = 0; e < B; e++) xd[e] = x[B * cd + e];
#pragma unroll
  for (int t = 0; t < B; t++)  // the diagonal block, its upper entries mirrored
#pragma unroll
    for (int e = 0; e < B; e++)
      sum[t] += own[(ND + (e <= t ? t * (t + 1) / 2 + e : e * (e + 1) / 2 + t)) * 64] * xd[e];
#pragma unroll
  for (int j = NLN - 1; j >= 0; j--) {  // upper neighbour nodes
    const int qm = q - m(j);
    const T *__restrict__ mir = hv + (int64_t)(qm >> 6) * NV * 64 + (qm & 63);
    const int c = min(qm, nn - 1);
    T xc[B];
#pragma unroll
    for (int e = 0; e < B; e++) xc[e] = x[B * c + e];
#pragma unroll
    for (int t = 0; t < B; t++)
#pragma unroll
      for (int e = 0; e < B; e++) sum[t] += mir[((j * B + e) * B + t) * 64] * xc[e];
  }
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

// the block's rows [r0, r1) with r0 a multiple of A (for A-wide vector accesses)
template <int A>
__device__ __forceinline__ void chunk_al(int n, int &r0, int &r1)
{
  const int nb = gridDim.x;
  const int rpb = ((n + nb - 1) / nb + A - 1) / A * A;
  const int lb = xcd_block(blockIdx.x, nb);
  r0 = min(n, lb * rpb);
  r1 = min(n, r0 + rpb);
}

__device__ __forceinline__ void chunk(int n, int &r0, int &r1)
{
  const int nb = gridDim.x;
  const int rpb = (n + nb - 1) / nb;
  const int lb = xcd_block(blockIdx.x, nb);
  r0 = lb * rpb;
  r1 = min(n, r0 + rpb);
}

}  // namespace ssf
