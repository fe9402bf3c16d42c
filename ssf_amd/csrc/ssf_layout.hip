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
//
// Stencil form: when the full rows' column offsets from the first row of their
// block of B rows (B = 1, 2, 3) form one list (at most 32 for B = 1, 96 else)
// and one slot per offset in every row costs at most 5 % more entries, every
// row gets those slots instead: slot q of row r holds its entry at column
// (r - r % B) + sten[q], or 0. No column array is kept; the list is found with
// small hash tables (per block in LDS, then merged), the same rule as the host
// layout's.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

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

// ---- stencil form
constexpr int STEN_TAB = 256;  // hash table slots per block size (> SELL_STEN_MAX_B)

__device__ __forceinline__ unsigned sten_hash(int o) { return ((unsigned)o * 2654435761u) >> 24; }

// o into the LDS table; false when the table is full
__device__ __forceinline__ bool sten_put_lds(int *tab, int o)
{
  const unsigned h = sten_hash(o);
  for (int p = 0; p < STEN_TAB; p++) {
    int *slot = tab + ((h + p) & (STEN_TAB - 1));
    const int v = *(volatile int *)slot;
    if (v == o) return true;
    if (v == INT_MIN) {
      const int old = atomicCAS(slot, INT_MIN, o);
      if (old == INT_MIN || old == o) return true;
    }
  }
  return false;
}

// o into the global table, counting the distinct offsets in *cnt
__device__ __forceinline__ void sten_put_global(int *tab, int *cnt, int o)
{
  const unsigned h = sten_hash(o);
  for (int p = 0; p < STEN_TAB; p++) {
    int *slot = tab + ((h + p) & (STEN_TAB - 1));
    const int v = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == o) return;
    if (v == INT_MIN) {
      const int old = atomicCAS(slot, INT_MIN, o);
      if (old == INT_MIN) {
        atomicAdd(cnt, 1);
        return;
      }
      if (old == o) return;
    }
  }
  atomicAdd(cnt, STEN_TAB + 1);  // more offsets than any stencil form holds
}

__global__ void k_lay_sten_init(int *__restrict__ tabs, int *__restrict__ cnts)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * STEN_TAB) tabs[t] = INT_MIN;
  if (t < 3) cnts[t] = 0;
}

// the offsets of every full-row entry from the first row of its block of B rows:
// scr entry k (row j, column i) is column i of full row j and, off the
// diagonal, column j of full row i. Each block collects its distinct offsets in
// LDS, then adds them to the global tables (tabs[B - 1], cnts[B - 1]).
__global__ __launch_bounds__(TPB) void k_lay_sten_collect(int64_t negm, const int *__restrict__ erow,
                                                          const int *__restrict__ ci,
                                                          int *__restrict__ tabs,
                                                          int *__restrict__ cnts)
{
  __shared__ int ltab[3][STEN_TAB];
  __shared__ int lfull[3];
  for (int t = threadIdx.x; t < 3 * STEN_TAB; t += blockDim.x) (&ltab[0][0])[t] = INT_MIN;
  if (threadIdx.x < 3) lfull[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < negm;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int j = erow[k], i = ci[k];
#pragma unroll
    for (int B = 1; B <= 3; B++) {
      if (*(volatile int *)&lfull[B - 1]) continue;
      bool ok = sten_put_lds(ltab[B - 1], i - (j - j % B));
      if (ok && i != j) ok = sten_put_lds(ltab[B - 1], j - (i - i % B));
      if (!ok) lfull[B - 1] = 1;
    }
  }
  __syncthreads();
#pragma unroll
  for (int B = 1; B <= 3; B++) {
    if (lfull[B - 1]) {
      if (threadIdx.x == 0) atomicAdd(cnts + B - 1, STEN_TAB + 1);
      continue;
    }
    for (int t = threadIdx.x; t < STEN_TAB; t += blockDim.x) {
      const int o = ltab[B - 1][t];
      if (o != INT_MIN) sten_put_global(tabs + (B - 1) * STEN_TAB, cnts + B - 1, o);
    }
  }
}

__global__ void k_lay_sten_off(int nslice, int w, int *__restrict__ soff)
{
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl <= nslice) soff[sl] = sl * 64 * w;
}

// slot of offset o in the sorted list (it is there: the list holds them all)
__device__ __forceinline__ int sten_slot(const int *__restrict__ sten, int ns, int o)
{
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sten[mid] < o) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// SELL positions of scr entry k in the stencil form: in full row j (lo) and,
// off the diagonal, in full row i (up)
__global__ void k_lay_sten_pos(int64_t negm, const int *__restrict__ erow,
                               const int *__restrict__ ci, const int *__restrict__ sten, int ns,
                               int B, int w, int *__restrict__ lo, int *__restrict__ up)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= negm) return;
  const int j = erow[k], i = ci[k];
  const int q = sten_slot(sten, ns, i - (j - j % B));
  lo[k] = (j >> 6) * 64 * w + (q / SELL_V) * 64 * SELL_V + (j & 63) * SELL_V + q % SELL_V;
  if (i != j) {
    const int q2 = sten_slot(sten, ns, j - (i - i % B));
    up[k] = (i >> 6) * 64 * w + (q2 / SELL_V) * 64 * SELL_V + (i & 63) * SELL_V + q2 % SELL_V;
  } else {
    up[k] = -1;
  }
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
  // the stencil form when it applies (same rule as the host layout's)
  std::vector<int> su;
  int sb = 1;
  if (e == hipSuccess && negm > 0) {
    int *tabs = nullptr, *cnts = nullptr;
    c(amalloc(&tabs, 3 * STEN_TAB, s));
    c(amalloc(&cnts, 3, s));
    std::vector<int> htab(3 * STEN_TAB);
    int hc[3] = {0, 0, 0};
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_lay_sten_init, dim3(nblocks(3 * STEN_TAB)), dim3(TPB), 0, s, tabs, cnts);
      const unsigned g = (unsigned)std::min<int64_t>(1024, nblocks(negm));
      hipLaunchKernelGGL(k_lay_sten_collect, dim3(g), dim3(TPB), 0, s, negm, erow, ci, tabs, cnts);
      c(hipGetLastError());
    }
    if (e == hipSuccess)
      c(hipMemcpyAsync(htab.data(), tabs, htab.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    if (e == hipSuccess) c(hipMemcpyAsync(hc, cnts, sizeof hc, hipMemcpyDeviceToHost, s));
    if (e == hipSuccess) c(hipStreamSynchronize(s));
    afree(tabs, s);
    afree(cnts, s);
    for (int B = 1; e == hipSuccess && B <= 3 && su.empty(); B++) {
      const int maxn = B == 1 ? SELL_STEN_MAX : SELL_STEN_MAX_B;
      if (n % B != 0 || hc[B - 1] <= 0 || hc[B - 1] > maxn) continue;
      std::vector<int> u;
      for (int t = 0; t < STEN_TAB; t++)
        if (htab[(B - 1) * STEN_TAB + t] != INT_MIN) u.push_back(htab[(B - 1) * STEN_TAB + t]);
      std::sort(u.begin(), u.end());
      const int64_t w = ((int64_t)u.size() + SELL_V - 1) / SELL_V * SELL_V;
      const int64_t stot = w * 64 * nslice;
      if (stot > (int64_t)tot + tot / 20) continue;
      // the stencil kernels address values and x by 32-bit byte offsets
      if (std::max<int64_t>(stot, n) * (int64_t)sizeof(double) >= (int64_t(1) << 32)) continue;
      su.swap(u);
      sb = B;
    }
  }
  if (e == hipSuccess && !su.empty()) {
    const int w = (int)((su.size() + SELL_V - 1) / SELL_V * SELL_V);
    std::vector<int> sp(su);
    sp.resize(w, 0);
    int *dsten = nullptr;
    c(amalloc(&dsten, w, s));
    if (e == hipSuccess) c(hipMemcpyAsync(dsten, sp.data(), w * sizeof(int), hipMemcpyHostToDevice, s));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_lay_sten_off, dim3(nblocks(nslice + 1)), dim3(TPB), 0, s, nslice, w, soff);
      hipLaunchKernelGGL(k_lay_sten_pos, dim3(nblocks(negm)), dim3(TPB), 0, s, negm, erow, ci,
                         dsten, (int)su.size(), sb, w, lo, up);
      c(hipGetLastError());
    }
    if (e == hipSuccess) c(hipStreamSynchronize(s));  // sp goes out of scope
    afree(erow, s); afree(ucnt, s); afree(ustart, s); afree(flen, s);
    afree(key, s); afree(val, s); afree(skey, s); afree(sval, s); afree(sw, s); afree(bad, s);
    afree(ll, s);
    if (tmp) (void)hipFreeAsync(tmp, s);
    if (e != hipSuccess) {
      afree(soff, s); afree(lo, s); afree(up, s); afree(dsten, s);
      return e;
    }
    out->nslice = nslice;
    out->tot = (int64_t)w * 64 * nslice;
    out->off = soff;
    out->lo = lo;
    out->up = up;
    out->sten = dsten;
    out->sten_b = sb;
    out->sten_h = su;
    return hipSuccess;
  }
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
