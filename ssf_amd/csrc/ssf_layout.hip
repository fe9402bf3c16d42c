// SELL-64 layout of the full matrix of a symcomprow pattern, built on the device
// from the pattern itself (adr, ci: lower rows, ascending, diagonal last).
//
// Full row i = the lower part of scr row i (ascending, diagonal last), then the
// upper part: the rows j > i holding column i, in ascending j. Entry q of full
// row r sits at soff[r / 64] + (q / V) * 64 V + (r % 64) V + q % V (V = SELL_V).
// For every scr entry k this builds its SELL position in row j (lo[k]) and, off
// the diagonal, its mirrored position in row i (up[k]); padding entries hold
// the row's own column and value 0.
//
// The ascending order of the upper part comes from a stable radix sort of the
// strictly lower entries by column (entries enter in row order), so the layout,
// and every summation order over it, is deterministic.
#include <hipcub/hipcub.hpp>

#include "ssf_kernels.hpp"

namespace ssf {

__global__ void k_lay_rows(int n, const int *__restrict__ adr, int *__restrict__ erow,
                           int *__restrict__ ll)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = adr[i], b = adr[i + 1];
  ll[i] = b - a;
  for (int k = a; k < b; k++) erow[k] = i;
}

// upper counts per column; sort keys: column for strictly lower entries, n for
// the diagonal (sorted to the end)
__global__ void k_lay_keys(int64_t negm, int n, const int *__restrict__ ci,
                           const int *__restrict__ erow, int *__restrict__ ucnt,
                           int *__restrict__ key, int *__restrict__ val)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= negm) return;
  const int c = ci[k];
  const bool off = c != erow[k];
  if (off) atomicAdd(ucnt + c, 1);
  key[k] = off ? c : n;
  val[k] = (int)k;
}

// slice widths (64 * longest row, rounded up to SELL_V); full row lengths
__global__ void k_lay_width(int n, int nslice, const int *__restrict__ ll,
                            const int *__restrict__ ucnt, int *__restrict__ flen,
                            int *__restrict__ sw)
{
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl > nslice) return;
  if (sl == nslice) {
    sw[sl] = 0;
    return;
  }
  int w = 0;
  for (int r = sl * 64; r < min(n, sl * 64 + 64); r++) {
    const int len = ll[r] + ucnt[r];
    flen[r] = len;
    w = max(w, len);
  }
  sw[sl] = (w + SELL_V - 1) / SELL_V * SELL_V * 64;
}

__device__ __forceinline__ int64_t lay_pos(const int *__restrict__ soff, int r, int q)
{
  return (int64_t)soff[r >> 6] + (int64_t)(q / SELL_V) * 64 * SELL_V + (r & 63) * SELL_V + q % SELL_V;
}

// every position of slice sl: the row's own column (padding); real entries are
// written over it afterwards
__global__ void k_lay_pad(int n, int nslice, const int *__restrict__ soff,
                          int *__restrict__ scol)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = (int)(t / 64), l = (int)(t % 64);
  if (sl >= nslice) return;
  const int r = sl * 64 + l;
  const int w = (soff[sl + 1] - soff[sl]) / 64;
  const int own = min(r, n - 1);
  for (int q = 0; q < w; q++) scol[lay_pos(soff, r, q)] = own;
}

// lower positions: entry k of scr row j is entry k - adr[j] of full row j
__global__ void k_lay_lower(int64_t negm, const int *__restrict__ adr, const int *__restrict__ ci,
                            const int *__restrict__ erow, const int *__restrict__ soff,
                            int *__restrict__ scol, int *__restrict__ lo, int *__restrict__ up)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= negm) return;
  const int j = erow[k];
  const int64_t p = lay_pos(soff, j, (int)(k - adr[j]));
  scol[p] = ci[k];
  lo[k] = (int)p;
  up[k] = -1;
}

// upper positions: the p-th sorted strictly lower entry (row j, column i) is
// entry ll[i] + (p - ustart[i]) of full row i
__global__ void k_lay_upper(int64_t nlow, const int *__restrict__ skey,
                            const int *__restrict__ sval, const int *__restrict__ ustart,
                            const int *__restrict__ ll, const int *__restrict__ erow,
                            const int *__restrict__ soff, int *__restrict__ scol,
                            int *__restrict__ up)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlow) return;
  const int i = skey[p];
  const int k = sval[p];
  const int q = ll[i] + (int)(p - ustart[i]);
  const int64_t pos = lay_pos(soff, i, q);
  scol[pos] = erow[k];
  up[k] = (int)pos;
}

// 16-bit offsets from the slice's first row; *bad set if one does not fit
__global__ void k_lay_col16(int64_t tot, int nslice, const int *__restrict__ soff,
                            const int *__restrict__ scol, short *__restrict__ c16,
                            int *__restrict__ bad)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= tot) return;
  // slice of position k: binary search in soff
  int lo = 0, hi = nslice;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (soff[mid] <= k) lo = mid;
    else hi = mid;
  }
  const int o = scol[k] - lo * 64;
  if (o < -32768 || o > 32767) {
    *bad = 1;
    return;
  }
  c16[k] = (short)o;
}

// SELL values and diagonal from the reference-layout values
__global__ void k_lay_expand(int64_t negm, const int *__restrict__ lo, const int *__restrict__ up,
                             const double *__restrict__ a, double *__restrict__ sv)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= negm) return;
  const double v = a[k];
  sv[lo[k]] = v;
  if (up[k] >= 0) sv[up[k]] = v;
}

__global__ void k_lay_diag(int n, const int *__restrict__ adr, const double *__restrict__ a,
                           double *__restrict__ diag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = a[adr[i + 1] - 1];  // symcomprow: the diagonal is last
}

static unsigned nblocks(int64_t work) { return (unsigned)((work + TPB - 1) / TPB); }

template <typename T>
static hipError_t amalloc(T **p, int64_t count, hipStream_t s)
{
  *p = nullptr;
  return hipMallocAsync((void **)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T), s);
}

template <typename T>
static void afree(T *&p, hipStream_t s)
{
  if (p) (void)hipFreeAsync(p, s);
  p = nullptr;
}

hipError_t sell_layout_build(int n, int64_t negm, const int *adr, const int *ci, hipStream_t s,
                             SellLayout *out)
{
  *out = SellLayout();
  if (n <= 0) return hipSuccess;
  const int nslice = (n + 63) / 64;
  int *erow = nullptr, *ll = nullptr, *ucnt = nullptr, *ustart = nullptr, *flen = nullptr;
  int *key = nullptr, *val = nullptr, *skey = nullptr, *sval = nullptr, *sw = nullptr;
  int *soff = nullptr, *scol = nullptr, *lo = nullptr, *up = nullptr, *bad = nullptr;
  short *c16 = nullptr;
  void *tmp = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(amalloc(&erow, negm, s)); c(amalloc(&ll, n, s)); c(amalloc(&ucnt, n + 1, s));
  c(amalloc(&ustart, n + 1, s)); c(amalloc(&flen, n, s));
  c(amalloc(&key, negm, s)); c(amalloc(&val, negm, s)); c(amalloc(&skey, negm, s));
  c(amalloc(&sval, negm, s)); c(amalloc(&sw, nslice + 1, s)); c(amalloc(&soff, nslice + 1, s));
  c(amalloc(&lo, negm, s)); c(amalloc(&up, negm, s)); c(amalloc(&bad, 1, s));
  if (e == hipSuccess) c(hipMemsetAsync(ucnt, 0, (n + 1) * sizeof(int), s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lay_rows, dim3(nblocks(n)), dim3(TPB), 0, s, n, adr, erow, ll);
    hipLaunchKernelGGL(k_lay_keys, dim3(nblocks(negm)), dim3(TPB), 0, s, negm, n, ci, erow, ucnt,
                       key, val);
    hipLaunchKernelGGL(k_lay_width, dim3(nblocks(nslice + 1)), dim3(TPB), 0, s, n, nslice, ll,
                       ucnt, flen, sw);
    c(hipGetLastError());
  }
  // slice offsets and the first upper slot of every column (exclusive scans)
  size_t tb = 0, tb2 = 0, tb3 = 0;
  int bits = 1;
  while ((1 << bits) <= n) bits++;
  if (e == hipSuccess) c(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, sw, soff, nslice + 1, s));
  if (e == hipSuccess) c(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, ucnt, ustart, n + 1, s));
  if (e == hipSuccess)
    c(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, key, skey, val, sval, (int)negm, 0, bits, s));
  const size_t tbytes = std::max(tb, std::max(tb2, tb3));
  if (e == hipSuccess) c(hipMallocAsync(&tmp, std::max<size_t>(tbytes, 1), s));
  if (e == hipSuccess) c(hipcub::DeviceScan::ExclusiveSum(tmp, tb, sw, soff, nslice + 1, s));
  if (e == hipSuccess) c(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, ucnt, ustart, n + 1, s));
  if (e == hipSuccess)
    c(hipcub::DeviceRadixSort::SortPairs(tmp, tb3, key, skey, val, sval, (int)negm, 0, bits, s));
  int tot = 0, nlow = 0;
  if (e == hipSuccess) c(hipMemcpyAsync(&tot, soff + nslice, sizeof(int), hipMemcpyDeviceToHost, s));
  if (e == hipSuccess) c(hipMemcpyAsync(&nlow, ustart + n, sizeof(int), hipMemcpyDeviceToHost, s));
  if (e == hipSuccess) c(hipStreamSynchronize(s));
  if (e == hipSuccess) c(amalloc(&scol, tot, s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lay_pad, dim3(nblocks((int64_t)nslice * 64)), dim3(TPB), 0, s, n, nslice,
                       soff, scol);
    hipLaunchKernelGGL(k_lay_lower, dim3(nblocks(negm)), dim3(TPB), 0, s, negm, adr, ci, erow, soff,
                       scol, lo, up);
    hipLaunchKernelGGL(k_lay_upper, dim3(nblocks(nlow)), dim3(TPB), 0, s, (int64_t)nlow, skey,
                       sval, ustart, ll, erow, soff, scol, up);
    c(hipGetLastError());
  }
  int hbad = 1;
  if (e == hipSuccess) c(amalloc(&c16, tot, s));
  if (e == hipSuccess) c(hipMemsetAsync(bad, 0, sizeof(int), s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lay_col16, dim3(nblocks(tot)), dim3(TPB), 0, s, (int64_t)tot, nslice,
                       soff, scol, c16, bad);
    c(hipGetLastError());
  }
  if (e == hipSuccess) c(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  if (e == hipSuccess) c(hipStreamSynchronize(s));
  afree(erow, s); afree(ucnt, s); afree(ustart, s); afree(flen, s);
  afree(key, s); afree(val, s); afree(skey, s); afree(sval, s); afree(sw, s); afree(bad, s);
  afree(ll, s);
  if (tmp) (void)hipFreeAsync(tmp, s);
  if (e != hipSuccess) {
    afree(soff, s); afree(scol, s); afree(lo, s); afree(up, s); afree(c16, s);
    return e;
  }
  if (hbad) afree(c16, s);
  else afree(scol, s);
  out->nslice = nslice;
  out->tot = tot;
  out->off = soff;
  out->col = scol;
  out->col16 = c16;
  out->lo = lo;
  out->up = up;
  return hipSuccess;
}

hipError_t sell_expand(int n, int64_t negm, const int *adr, const int *lo, const int *up,
                       const double *a_ref, double *sv, double *diag, hipStream_t s)
{
  if (negm > 0)
    hipLaunchKernelGGL(k_lay_expand, dim3(nblocks(negm)), dim3(TPB), 0, s, negm, lo, up, a_ref, sv);
  if (n > 0) hipLaunchKernelGGL(k_lay_diag, dim3(nblocks(n)), dim3(TPB), 0, s, n, adr, a_ref, diag);
  return hipGetLastError();
}

}  // namespace ssf
