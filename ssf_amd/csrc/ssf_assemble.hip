// Element assembly by entry (gmatrix::localize, GEFEL/scr.cpp:422-458,
// GEFEL/cr.cpp:327) on gfx950.
//
// The reference adds every element matrix into a[] element after element, row i
// then column j of each, skipping code numbers <= 0 and, for symcomprow, the
// upper triangle. Here the element kernels only write their matrices to a
// buffer, and a plan built once per pattern and code-number table lists, for
// every entry of the reference layout, its contributions (element e, local row
// i, local column j) in exactly that order: a stable radix sort of the pairs
// (entry, e nd^2 + i nd + j) emitted in (e, i, j) order. One thread per entry then
// adds its contributions to the entry's value in list order. Every value is
// written once, with no atomics, no colouring and the reference's summation
// order; coupled code numbers (two DOFs of one element, or of several elements,
// sharing an unknown) need nothing special.
//
// The Dirichlet lift of assemble_init (TRFEL/SRC/globmatt.cpp:1635) follows the
// same scheme per unknown: the element kernels write -K_e u_p of the elements
// with prescribed values, and a host-built list per unknown adds them in element
// order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "ssf_device.hpp"

namespace ssf {

// contributions of element e's local row i: code numbers ii = cn[e][i] - 1 and
// jj = cn[e][j] - 1 both >= 0, and jj <= ii for symcomprow (GEFEL/scr.cpp:427-433)
__device__ __forceinline__ bool asm_contrib(bool sym, int ii, int jj)
{
  return ii >= 0 && jj >= 0 && !(sym && jj > ii);
}

__global__ __launch_bounds__(TPB) void k_asm_count(int64_t nrow, int nd, int sym,
                                                   const int *__restrict__ cn,
                                                   int *__restrict__ cnt)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (e, i)
  if (t >= nrow) return;
  const int64_t e = t / nd;
  const int ii = cn[t] - 1;
  int c = 0;
  for (int j = 0; j < nd; j++) c += asm_contrib(sym, ii, cn[e * nd + j] - 1) ? 1 : 0;
  cnt[t] = c;
}

// the (entry, rank) pairs of element e's row i at pos[(e, i)], columns ascending
__global__ __launch_bounds__(TPB) void k_asm_emit(int64_t nrow, int nd, int sym,
                                                  const int *__restrict__ cn,
                                                  const int *__restrict__ adr,
                                                  const int *__restrict__ ci,
                                                  const int *__restrict__ pos,
                                                  int *__restrict__ key, int *__restrict__ val,
                                                  int *__restrict__ bad)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrow) return;
  const int64_t e = t / nd;
  const int i = (int)(t % nd);
  const int ii = cn[t] - 1;
  int p = pos[t];
  for (int j = 0; j < nd; j++) {
    const int jj = cn[e * nd + j] - 1;
    if (!asm_contrib(sym, ii, jj)) continue;
    int lo = adr[ii], hi = adr[ii + 1] - 1, sl = -1;
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
    if (sl < 0) *bad = 1;  // the pattern lacks the entry: not built from these codes
    key[p] = sl < 0 ? 0 : sl;
    val[p] = (int)(e * nd * nd + i * nd + j);
    p++;
  }
}

// first contribution of every entry from the sorted entries: cstart[s] for
// s in (key[k-1], key[k]] is k (k = nc closes the list); each s is set once
__global__ __launch_bounds__(TPB) void k_asm_starts(int64_t nc, int64_t negm,
                                                    const int *__restrict__ key,
                                                    int *__restrict__ cstart)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nc) return;
  const int64_t prev = k == 0 ? -1 : key[k - 1];
  const int64_t cur = k == nc ? negm : key[k];
  for (int64_t s = prev + 1; s <= cur; s++) cstart[s] = (int)k;
}

int asm_plan_build(int storage, int64_t n, int64_t negm, int64_t ne, int nd, const int *cn,
                   const int *adr, const int *ci, hipStream_t s, int **cstart_out,
                   int **cidx_out, int64_t *nc_out)
{
  const bool sym = storage == SSF_STORAGE_SYMM_COMP_ROWS;
  const int64_t nrow = ne * nd;
  SSF_CHECK(ne * nd * nd < (1LL << 31) && negm < (1LL << 31) && nrow < (1LL << 31), SSF_EUNSUP,
            "assembly plan: more than 2^31 element-matrix entries");
  (void)n;
  int *cnt = nullptr, *pos = nullptr, *bad = nullptr;
  SSF_HIP(dalloc(&cnt, nrow + 1, s));
  SSF_HIP(dalloc(&pos, nrow + 1, s));
  SSF_HIP(dalloc(&bad, 1, s));
  SSF_HIP(hipMemsetAsync(cnt + nrow, 0, sizeof(int), s));
  SSF_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const unsigned gr = (unsigned)((nrow + TPB - 1) / TPB);
  if (nrow) hipLaunchKernelGGL(k_asm_count, dim3(gr), dim3(TPB), 0, s, nrow, nd, sym ? 1 : 0, cn, cnt);
  SSF_HIP(hipGetLastError());
  size_t tb = 0;
  SSF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, pos, (int)(nrow + 1), s));
  void *tmp = nullptr;
  SSF_HIP(hipMallocAsync(&tmp, std::max<size_t>(tb, 1), s));
  SSF_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, pos, (int)(nrow + 1), s));
  (void)hipFreeAsync(tmp, s);
  int nc32 = 0;
  SSF_HIP(hipMemcpyAsync(&nc32, pos + nrow, sizeof(int), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  const int64_t nc = nc32;
  dfree(cnt, s);
  int *key = nullptr, *val = nullptr, *skey = nullptr, *sval = nullptr, *cstart = nullptr;
  SSF_HIP(dalloc(&key, nc, s));
  SSF_HIP(dalloc(&val, nc, s));
  SSF_HIP(dalloc(&skey, nc, s));
  SSF_HIP(dalloc(&sval, nc, s));
  if (nrow)
    hipLaunchKernelGGL(k_asm_emit, dim3(gr), dim3(TPB), 0, s, nrow, nd, sym ? 1 : 0, cn, adr, ci,
                       pos, key, val, bad);
  SSF_HIP(hipGetLastError());
  dfree(pos, s);
  int bits = 1;
  while (bits < 31 && (int64_t(1) << bits) <= negm) bits++;
  // stable: the contributions of an entry keep their (e, i, j) order
  tb = 0;
  SSF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, val, sval, (int)nc, 0, bits, s));
  SSF_HIP(hipMallocAsync(&tmp, std::max<size_t>(tb, 1), s));
  SSF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, skey, val, sval, (int)nc, 0, bits, s));
  (void)hipFreeAsync(tmp, s);
  dfree(key, s);
  dfree(val, s);
  SSF_HIP(dalloc(&cstart, negm + 1, s));
  hipLaunchKernelGGL(k_asm_starts, dim3((unsigned)((nc + 1 + TPB - 1) / TPB)), dim3(TPB), 0, s, nc,
                     negm, skey, cstart);
  SSF_HIP(hipGetLastError());
  int hbad = 0;
  SSF_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  dfree(skey, s);
  dfree(bad, s);
  if (hbad) {
    dfree(sval, s);
    dfree(cstart, s);
    set_error("assembly: an element couples unknowns the matrix pattern has no entry for "
              "(the pattern was built from other code numbers)");
    return SSF_EINVAL;
  }
  *cstart_out = cstart;
  *cidx_out = sval;
  *nc_out = nc;
  return SSF_OK;
}

// a[s] += the contributions of entry s in list order; a contribution
// e nd^2 + i nd + j reads value t of element e in the element-matrix buffer: the
// full matrix (P = nd^2, t = i nd + j) or its lower triangle (P = nd (nd + 1) / 2,
// t of (max(i, j), min(i, j)), symmetric K_e), at (e / W) W P + t W + e % W
template <bool SYM>
__global__ __launch_bounds__(TPB) void k_asm_gather(int64_t negm, int nd, int wshift,
                                                    const int *__restrict__ cstart,
                                                    const int *__restrict__ cidx,
                                                    const double *__restrict__ kbuf,
                                                    double *__restrict__ a)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= negm) return;
  const int k0 = cstart[s], k1 = cstart[s + 1];
  if (k0 == k1) return;
  const int nd2 = nd * nd;
  const int64_t P = SYM ? (int64_t)nd * (nd + 1) / 2 : nd2;
  const int wmask = (1 << wshift) - 1;
  double acc = a[s];
  for (int k = k0; k < k1; k++) {
    const int v = cidx[k];
    const int e = v / nd2, ij = v - e * nd2;
    int t = ij;
    if (SYM) {
      const int i = ij / nd, j = ij - i * nd;
      t = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
    }
    acc += kbuf[(int64_t)(e & ~wmask) * P + ((int64_t)t << wshift) + (e & wmask)];
  }
  a[s] = acc;
}

hipError_t asm_gather(int64_t negm, int nd, bool sym, int wshift, const int *cstart,
                      const int *cidx, const double *kbuf, double *a, hipStream_t s)
{
  if (negm == 0) return hipSuccess;
  // (a mapping that gives each XCD a contiguous eighth of the entries measured
  // 12 % slower on the 64x64x256 elasticity column: 2.93 against 2.62 ms)
  const unsigned g = (unsigned)((negm + TPB - 1) / TPB);
  if (sym) hipLaunchKernelGGL(k_asm_gather<true>, dim3(g), dim3(TPB), 0, s, negm, nd, wshift, cstart, cidx, kbuf, a);
  else hipLaunchKernelGGL(k_asm_gather<false>, dim3(g), dim3(TPB), 0, s, negm, nd, wshift, cstart, cidx, kbuf, a);
  return hipGetLastError();
}

// lift[d] += the per-element lift values of unknown d in element order
__global__ __launch_bounds__(TPB) void k_lift_gather(int64_t n, const int *__restrict__ lstart,
                                                     const int *__restrict__ lidx,
                                                     const double *__restrict__ fbuf,
                                                     double *__restrict__ lift)
{
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  const int k0 = lstart[d], k1 = lstart[d + 1];
  if (k0 == k1) return;
  double acc = lift[d];
  for (int k = k0; k < k1; k++) acc += fbuf[lidx[k]];
  lift[d] = acc;
}

hipError_t lift_gather(int64_t n, const int *lstart, const int *lidx, const double *fbuf,
                       double *lift, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lift_gather, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n,
                     lstart, lidx, fbuf, lift);
  return hipGetLastError();
}

// The lift plan on the host: the elements with a prescribed code number get the
// consecutive indices pidx[e] >= 0 (else -1); unknown d's contributions are
// pidx[e] nd + i for the (e, i) with cn[e][i] = d + 1, in element order.
void lift_plan_host(int64_t n, int64_t ne, int nd, const int64_t *cn, std::vector<int> &pidx,
                    std::vector<int> &lstart, std::vector<int> &lidx, int64_t *npe)
{
  pidx.assign(ne, -1);
  int64_t np = 0;
  for (int64_t e = 0; e < ne; e++)
    for (int i = 0; i < nd; i++)
      if (cn[e * nd + i] < 0) {
        pidx[e] = (int)np++;
        break;
      }
  lstart.assign(n + 1, 0);
  for (int64_t e = 0; e < ne; e++)
    if (pidx[e] >= 0)
      for (int i = 0; i < nd; i++)
        if (cn[e * nd + i] > 0) lstart[cn[e * nd + i]]++;
  for (int64_t d = 0; d < n; d++) lstart[d + 1] += lstart[d];
  lidx.assign(lstart[n], 0);
  std::vector<int> fillp(lstart.begin(), lstart.end() - 1);
  for (int64_t e = 0; e < ne; e++)
    if (pidx[e] >= 0)
      for (int i = 0; i < nd; i++)
        if (cn[e * nd + i] > 0) lidx[fillp[cn[e * nd + i] - 1]++] = pidx[e] * nd + i;
  *npe = np;
}

}  // namespace ssf
