// symcomprow pattern of an element mesh built on the device.
//
// symcomprow::initiate (GEFEL/scr.cpp:556-570: numcontr_elem :112, addresses
// :211, fillarray_elem :163, sort_and_colindex_tko :317) collects for every row
// the columns j <= i coupled to it by some element, ascending, so the diagonal
// comes last. Here every element emits one 64-bit key (row << b | column) per
// unordered pair of its DOFs with positive code numbers (row = the larger
// code - 1); a radix sort orders the keys by row, then column; the first key of
// each run of equal keys is an entry of the pattern, and the first entry of
// each row starts it. The arrays equal the reference's bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "ssf_kernels.hpp"

namespace ssf {
namespace {

constexpr unsigned long long kNoKey = ~0ULL;

// element e's pair p = (a, b), a <= b, in the order (0,0), (0,1), ..., (1,1), ...
__global__ void k_pat_keys(int64_t ne, int nd, const int *__restrict__ cn, int bits,
                           unsigned long long *__restrict__ key)
{
  const int npe = nd * (nd + 1) / 2;
  const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (t >= ne * npe) return;
  const int64_t e = t / npe;
  int p = (int)(t - e * npe), a = 0;
  while (p >= nd - a) {
    p -= nd - a;
    a++;
  }
  const int b = a + p;
  const int ca = cn[e * nd + a], cb = cn[e * nd + b];
  if (ca < 1 || cb < 1) {
    key[t] = kNoKey;
    return;
  }
  const unsigned long long r = (unsigned long long)(max(ca, cb) - 1);
  const unsigned long long c = (unsigned long long)(min(ca, cb) - 1);
  key[t] = (r << bits) | c;
}

// flag[k] = 1 for the first key of a run of equal keys (not the sentinel)
__global__ void k_pat_flags(int64_t m, const unsigned long long *__restrict__ key,
                            int *__restrict__ flag)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= m) return;
  const unsigned long long v = key[k];
  flag[k] = (v != kNoKey && (k == 0 || key[k - 1] != v)) ? 1 : 0;
}

// ci[pos] = column; the first entry of a row sets adr[row] (rows of no entry
// keep -1 and are reported)
__global__ void k_pat_fill(int64_t m, int bits, const unsigned long long *__restrict__ key,
                           const int *__restrict__ flag, const int *__restrict__ pos,
                           int *__restrict__ adr, int *__restrict__ ci)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= m || !flag[k]) return;
  const unsigned long long v = key[k];
  const unsigned long long mask = (1ULL << bits) - 1;
  const int row = (int)(v >> bits);
  ci[pos[k]] = (int)(v & mask);
  if (k == 0 || (int)(key[k - 1] >> bits) != row) adr[row] = pos[k];
}

__global__ void k_pat_fill_int(int64_t m, int v, int *__restrict__ a)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k < m) a[k] = v;
}

// every row must hold its diagonal (the last entry): adr set, ci[adr[i+1]-1] == i
__global__ void k_pat_check(int n, const int *__restrict__ adr, const int *__restrict__ ci,
                            int *__restrict__ bad)
{
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  if (adr[i] < 0 || adr[i + 1] <= adr[i] || ci[adr[i + 1] - 1] != i) atomicOr(bad, 1);
}

__global__ void k_pat_sum(int64_t m, const int *__restrict__ v, unsigned long long salt,
                          unsigned long long *__restrict__ out)
{
  unsigned long long s = 0;
  for (int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x; k < m; k += (int64_t)gridDim.x * TPB)
    s += (unsigned long long)((long long)v[k] + 1) * mix64((unsigned long long)k ^ salt);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

// full row lengths of a symcomprow pattern: the lower row plus the mirrored entries
__global__ void k_pat_upper(int64_t negm, int n, const int *__restrict__ adr,
                            const int *__restrict__ ci, int *__restrict__ len)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= negm) return;
  atomicAdd(len + ci[k], 1);  // counts the diagonal too
}

// SELL-64 entries: per slice 64 x the longest row rounded up to SELL_V
__global__ void k_pat_slices(int n, const int *__restrict__ adr, const int *__restrict__ len,
                             unsigned long long *__restrict__ tot)
{
  const int sl = blockIdx.x * TPB + threadIdx.x;
  const int nsl = (n + 63) / 64;
  unsigned long long v = 0;
  if (sl < nsl) {
    int w = 0;
    for (int r = sl * 64; r < min(n, sl * 64 + 64); r++)
      w = max(w, (adr[r + 1] - adr[r]) + len[r] - 1);  // len counts the diagonal once more
    v = (unsigned long long)((w + SELL_V - 1) / SELL_V * SELL_V) * 64;
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(tot, v);
}

inline unsigned grid(int64_t m) { return (unsigned)std::max<int64_t>(1, (m + TPB - 1) / TPB); }

}  // namespace

hipError_t pattern_checksum_dev(const int *adr, int64_t n, const int *ci, int64_t negm,
                                hipStream_t s, uint64_t *out)
{
  unsigned long long *d = nullptr;
  hipError_t e = hipMallocAsync((void **)&d, sizeof(*d), s);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(d, 0, sizeof(*d), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_sum, dim3(std::min<unsigned>(grid(n + 1), 1024)), dim3(TPB), 0, s,
                       n + 1, adr, kSaltAdr, d);
    hipLaunchKernelGGL(k_pat_sum, dim3(std::min<unsigned>(grid(negm), 4096)), dim3(TPB), 0, s,
                       negm, ci, kSaltCi, d);
    e = hipGetLastError();
  }
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFreeAsync(d, s);
  *out = h;
  return e;
}

hipError_t sell_entries_dev(int64_t n, int64_t negm, const int *adr, const int *ci, hipStream_t s,
                            int64_t *tot_out)
{
  int *len = nullptr;
  unsigned long long *tot = nullptr;
  hipError_t e = hipMallocAsync((void **)&len, n * sizeof(int), s);
  if (e == hipSuccess) e = hipMallocAsync((void **)&tot, sizeof(*tot), s);
  if (e == hipSuccess) e = hipMemsetAsync(len, 0, n * sizeof(int), s);
  if (e == hipSuccess) e = hipMemsetAsync(tot, 0, sizeof(*tot), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_upper, dim3(grid(negm)), dim3(TPB), 0, s, negm, (int)n, adr, ci, len);
    hipLaunchKernelGGL(k_pat_slices, dim3(grid((n + 63) / 64)), dim3(TPB), 0, s, (int)n, adr, len,
                       tot);
    e = hipGetLastError();
  }
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, tot, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (len) (void)hipFreeAsync(len, s);
  if (tot) (void)hipFreeAsync(tot, s);
  *tot_out = (int64_t)h;
  return e;
}

int64_t pattern_build_bytes(int64_t ne, int nd)
{
  const int64_t m = ne * (int64_t)nd * (nd + 1) / 2;
  return m * (8 + 8 + 4 + 4) + m * 8 + (1 << 20);  // keys, sorted keys, flags, positions, sort temp
}

int pattern_build_dev(int64_t n, int64_t ne, int nd, const int *cn, hipStream_t s, int **adr_out,
                      int **ci_out, int64_t *negm_out)
{
  *adr_out = nullptr;
  *ci_out = nullptr;
  *negm_out = 0;
  const int64_t m = ne * (int64_t)nd * (nd + 1) / 2;
  SSF_CHECK(n > 0 && n < (1LL << 31) - 1 && m < (1LL << 31), SSF_EUNSUP,
            "device pattern: more than 2^31 element DOF pairs");
  int bits = 1;
  while ((1LL << bits) <= n) bits++;  // 2^bits > n: no entry key is all ones
  unsigned long long *key = nullptr, *skey = nullptr;
  int *flag = nullptr, *pos = nullptr, *adr = nullptr, *ci = nullptr, *bad = nullptr;
  void *tmp = nullptr;
  size_t tb = 0, tb2 = 0;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(hipMallocAsync((void **)&key, std::max<int64_t>(m, 1) * sizeof(*key), s));
  c(hipMallocAsync((void **)&skey, std::max<int64_t>(m, 1) * sizeof(*skey), s));
  c(hipMallocAsync((void **)&flag, std::max<int64_t>(m, 1) * sizeof(int), s));
  c(hipMallocAsync((void **)&pos, std::max<int64_t>(m, 1) * sizeof(int), s));
  c(hipMallocAsync((void **)&adr, (n + 1) * sizeof(int), s));
  c(hipMallocAsync((void **)&bad, sizeof(int), s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_keys, dim3(grid(m)), dim3(TPB), 0, s, ne, nd, cn, bits, key);
    c(hipGetLastError());
  }
  if (e == hipSuccess)
    c(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, key, skey, (int)m, 0, 2 * bits, s));
  if (e == hipSuccess) c(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, flag, pos, (int)m, s));
  if (e == hipSuccess) c(hipMallocAsync(&tmp, std::max<size_t>({tb, tb2, (size_t)1}), s));
  if (e == hipSuccess)
    c(hipcub::DeviceRadixSort::SortKeys(tmp, tb, key, skey, (int)m, 0, 2 * bits, s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_flags, dim3(grid(m)), dim3(TPB), 0, s, m, skey, flag);
    c(hipGetLastError());
  }
  if (e == hipSuccess) c(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, flag, pos, (int)m, s));
  int last_pos = 0, last_flag = 0;
  if (e == hipSuccess && m > 0) {
    c(hipMemcpyAsync(&last_pos, pos + m - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    c(hipMemcpyAsync(&last_flag, flag + m - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    c(hipStreamSynchronize(s));
  }
  const int64_t negm = (int64_t)last_pos + last_flag;
  if (e == hipSuccess) c(hipMallocAsync((void **)&ci, std::max<int64_t>(negm, 1) * sizeof(int), s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_fill_int, dim3(grid(n + 1)), dim3(TPB), 0, s, n + 1, -1, adr);
    hipLaunchKernelGGL(k_pat_fill_int, dim3(1), dim3(TPB), 0, s, (int64_t)1, (int)negm, adr + n);
    hipLaunchKernelGGL(k_pat_fill, dim3(grid(m)), dim3(TPB), 0, s, m, bits, skey, flag, pos, adr, ci);
    c(hipGetLastError());
  }
  int h_bad = 0;
  if (e == hipSuccess) c(hipMemsetAsync(bad, 0, sizeof(int), s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_check, dim3(grid(n)), dim3(TPB), 0, s, (int)n, adr, ci, bad);
    c(hipGetLastError());
  }
  if (e == hipSuccess) c(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  c(hipStreamSynchronize(s));
  if (tmp) (void)hipFreeAsync(tmp, s);
  for (void *p : {(void *)key, (void *)skey, (void *)flag, (void *)pos, (void *)bad})
    if (p) (void)hipFreeAsync(p, s);
  if (e != hipSuccess || h_bad) {
    if (adr) (void)hipFreeAsync(adr, s);
    if (ci) (void)hipFreeAsync(ci, s);
    if (e != hipSuccess) {
      set_error(std::string("device pattern: ") + hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? SSF_ENOMEM : SSF_EHIP;
    }
    set_error("device pattern: an unknown that no element refers to (every row needs its diagonal)");
    return SSF_EINVAL;
  }
  *adr_out = adr;
  *ci_out = ci;
  *negm_out = negm;
  return SSF_OK;
}

}  // namespace ssf
