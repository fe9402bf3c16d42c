// Cuthill-McKee node renumbering (gtopology::cuthill_mckee_renumb,
// GEFEL/gtopology.cpp:6242-6355, on the edge graph of adjacnodes_edge :2632) on
// gfx950.
//
// The reference numbers the nodes one at a time: it takes the nodes in number
// order and gives each unnumbered neighbour the next number, neighbours listed by
// degree (ties in node order). Run level by level this is the same order: every
// node reached from the current level is numbered by its lowest-numbered parent,
// at its position in that parent's list, so sorting the level's new nodes by the
// key (parent number, position) reproduces the sequential numbering exactly.
//   graph    unique edges of the elements: a sort of 64-bit (a, b) keys over both
//            directions; neighbours ascending; degree = row length
//   lists    stable sort of the entries by (row, degree of the neighbour)
//   start    the last node of least degree (the reference's `min >= deg`)
//   levels   per level: one thread per node of the level offers (its number,
//            position) to each unnumbered neighbour by a 64-bit atomic minimum;
//            the first offer appends the neighbour to the level's list; the list
//            is sorted by key and numbered in that order
// The level count is the eccentricity of the start node (a few hundred launches
// for a 128^3 mesh); the reference's search for the next parent scans all nodes
// per number, nn^2 steps.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "ssf_kernels.hpp"

namespace ssf {
namespace {

constexpr unsigned long long kNone = ~0ULL;

__constant__ int kHexEdges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
__constant__ int kQuadEdges[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

unsigned nblk(int64_t m) { return (unsigned)((std::max<int64_t>(m, 1) + TPB - 1) / TPB); }

// both directions of every element edge as (a << 32 | b)
__global__ void k_rn_edges(int64_t ne, int nne, int ned, const int *__restrict__ conn,
                           unsigned long long *__restrict__ key)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= ne * ned) return;
  const int64_t e = k / ned;
  const int q = (int)(k - e * ned);
  const int *ed = nne == 8 ? kHexEdges[q] : kQuadEdges[q];
  const unsigned long long a = (unsigned)conn[e * nne + ed[0]], b = (unsigned)conn[e * nne + ed[1]];
  key[2 * k] = (a << 32) | b;
  key[2 * k + 1] = (b << 32) | a;
}

__global__ void k_rn_unique(int64_t m, const unsigned long long *__restrict__ key,
                            int *__restrict__ flag)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k < m) flag[k] = (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// compacted entries: row, column; ptr[row] at the first entry of each row
__global__ void k_rn_fill(int64_t m, const unsigned long long *__restrict__ key,
                          const int *__restrict__ flag, const int *__restrict__ pos,
                          int *__restrict__ row, int *__restrict__ col)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= m || !flag[k]) return;
  const int p = pos[k];
  row[p] = (int)(key[k] >> 32);
  col[p] = (int)(key[k] & 0xffffffffULL);
}

__global__ void k_rn_ptr(int64_t nent, int nn, const int *__restrict__ row, int *__restrict__ ptr)
{
  const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= nent) return;
  const int r = row[p], rp = p > 0 ? row[p - 1] : -1;
  for (int d = rp + 1; d <= r; d++) ptr[d] = (int)p;
  if (p == nent - 1)
    for (int d = r + 1; d <= nn; d++) ptr[d] = (int)nent;
}

// key of entry p: row, then the degree of its neighbour (stable: ties keep node order)
__global__ void k_rn_degkey(int64_t nent, int dbits, const int *__restrict__ row,
                            const int *__restrict__ col, const int *__restrict__ ptr,
                            unsigned long long *__restrict__ key)
{
  const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= nent) return;
  const int c = col[p];
  key[p] = ((unsigned long long)row[p] << dbits) | (unsigned long long)(ptr[c + 1] - ptr[c]);
}

// per node: (degree, ~node): the minimum is the last node of least degree
__global__ void k_rn_startkey(int nn, const int *__restrict__ ptr,
                              unsigned long long *__restrict__ key, int *__restrict__ maxdeg)
{
  const int v = blockIdx.x * TPB + threadIdx.x;
  if (v >= nn) return;
  const int d = ptr[v + 1] - ptr[v];
  key[v] = ((unsigned long long)d << 32) | (0xffffffffULL - (unsigned)v);
  atomicMax(maxdeg, d);
}

__global__ void k_rn_init(int nn, int *__restrict__ cmk, unsigned long long *__restrict__ cand)
{
  const int v = blockIdx.x * TPB + threadIdx.x;
  if (v >= nn) return;
  cmk[v] = -1;
  cand[v] = kNone;
}

__global__ void k_rn_start(const unsigned long long *__restrict__ best, int *__restrict__ cmk,
                           int *__restrict__ inv)
{
  const int v = (int)(0xffffffffULL - (best[0] & 0xffffffffULL));
  cmk[v] = 0;
  inv[0] = v;
}

// the nodes numbered lo..hi-1 offer (number, position) to their unnumbered neighbours
__global__ void k_rn_expand(int lo, int hi, const int *__restrict__ inv, const int *__restrict__ ptr,
                            const int *__restrict__ adj, const int *__restrict__ cmk,
                            unsigned long long *__restrict__ cand, int *__restrict__ list,
                            int *__restrict__ count)
{
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= hi - lo) return;
  const int num = lo + t, u = inv[num];
  for (int e = ptr[u]; e < ptr[u + 1]; e++) {
    const int v = adj[e];
    if (cmk[v] != -1) continue;
    const unsigned long long k = ((unsigned long long)num << 32) | (unsigned)(e - ptr[u]);
    if (atomicMin(&cand[v], k) == kNone) list[atomicAdd(count, 1)] = v;
  }
}

__global__ void k_rn_gather(int cnt, const int *__restrict__ list,
                            const unsigned long long *__restrict__ cand,
                            unsigned long long *__restrict__ key)
{
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < cnt) key[i] = cand[list[i]];
}

__global__ void k_rn_number(int cnt, int m, const int *__restrict__ sorted, int *__restrict__ cmk,
                            int *__restrict__ inv, unsigned long long *__restrict__ cand)
{
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= cnt) return;
  const int v = sorted[i];
  cmk[v] = m + i;
  inv[m + i] = v;
  cand[v] = kNone;
}

struct Scratch {
  hipStream_t s;
  std::vector<void *> p;
  template <typename T>
  hipError_t get(T **out, size_t n)
  {
    *out = nullptr;
    hipError_t e = hipMallocAsync((void **)out, std::max<size_t>(n, 1) * sizeof(T), s);
    if (e == hipSuccess) p.push_back(*out);
    return e;
  }
  ~Scratch()
  {
    for (void *q : p) (void)hipFreeAsync(q, s);
  }
};

int bits_for(uint64_t v)
{
  int b = 1;
  while (b < 63 && (1ULL << b) <= v) b++;
  return b;
}

}  // namespace
}  // namespace ssf

using namespace ssf;

extern "C" int ssf_cuthill_mckee(ssf_ctx *ctx, int64_t nn, int64_t ne, int64_t nne,
                                 const int64_t *conn, int reverse, int64_t *ordering)
{
  SSF_ENTER();
  SSF_CHECK(ctx && (ne == 0 || conn) && ordering, SSF_EINVAL, "ssf_cuthill_mckee: NULL argument");
  SSF_CHECK(nne == 8 || nne == 4, SSF_EUNSUP,
            "ssf_cuthill_mckee: hex8 (nne 8) and quad4 (nne 4) element edges only");
  SSF_CHECK(nn > 0 && nn < (1LL << 31) - 1 && ne > 0 && ne * nne < (1LL << 31), SSF_EINVAL,
            "ssf_cuthill_mckee: sizes out of range (at least one element)");
  const int ned = nne == 8 ? 12 : 4;
  SSF_CHECK(2 * ne * ned < (1LL << 31), SSF_EINVAL, "ssf_cuthill_mckee: too many edges");
  std::vector<int> c32(ne * nne);
  for (int64_t k = 0; k < ne * nne; k++) {
    SSF_CHECK(conn[k] >= 0 && conn[k] < nn, SSF_EINVAL, "ssf_cuthill_mckee: node id out of range");
    c32[k] = (int)conn[k];
  }
  SSF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  Scratch S{s, {}};
  const int64_t m = 2 * ne * ned;
  const int nbits = bits_for((uint64_t)nn);  // node ids and numbers < 2^nbits <= 2^31
  int *conn_d, *flag, *pos, *row, *col, *ptr, *cmk, *inv, *list, *sorted, *count, *maxdeg;
  unsigned long long *key, *skey, *cand, *lkey, *slkey;
  int *sortedcol;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(S.get(&conn_d, ne * nne));
  c(S.get(&key, m));
  c(S.get(&skey, m));
  c(S.get(&flag, m));
  c(S.get(&pos, m));
  c(S.get(&row, m));
  c(S.get(&col, m));
  c(S.get(&sortedcol, m));
  c(S.get(&ptr, nn + 1));
  c(S.get(&cmk, nn));
  c(S.get(&inv, nn));
  c(S.get(&cand, nn));
  c(S.get(&list, nn));
  c(S.get(&sorted, nn));
  c(S.get(&lkey, nn));
  c(S.get(&slkey, nn));
  c(S.get(&count, 1));
  c(S.get(&maxdeg, 1));
  SSF_HIP(e);
  // temporary storage of the largest sort / scan / reduction
  size_t tb = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0;
  c(hipcub::DeviceRadixSort::SortKeys(nullptr, t1, key, skey, (int)m, 0, 32 + nbits, s));
  c(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, flag, pos, (int)m, s));
  c(hipcub::DeviceRadixSort::SortPairs(nullptr, t3, key, skey, col, sortedcol, (int)m, 0, 64, s));
  c(hipcub::DeviceReduce::Min(nullptr, t4, lkey, slkey, (int)nn, s));
  tb = std::max(std::max(t1, t2), std::max(t3, t4));
  void *tmp = nullptr;
  c(S.get((char **)&tmp, tb));
  SSF_HIP(e);
  SSF_HIP(hipMemcpyAsync(conn_d, c32.data(), c32.size() * sizeof(int), hipMemcpyHostToDevice, s));
  // unique edges, both directions, ascending (row, neighbour)
  hipLaunchKernelGGL(k_rn_edges, dim3(nblk(ne * ned)), dim3(TPB), 0, s, ne, (int)nne, ned, conn_d,
                     key);
  SSF_HIP(hipGetLastError());
  SSF_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, t1, key, skey, (int)m, 0, 32 + nbits, s));
  hipLaunchKernelGGL(k_rn_unique, dim3(nblk(m)), dim3(TPB), 0, s, m, skey, flag);
  SSF_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, t2, flag, pos, (int)m, s));
  hipLaunchKernelGGL(k_rn_fill, dim3(nblk(m)), dim3(TPB), 0, s, m, skey, flag, pos, row, col);
  int last_pos = 0, last_flag = 0;
  SSF_HIP(hipMemcpyAsync(&last_pos, pos + m - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipMemcpyAsync(&last_flag, flag + m - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  const int64_t nent = m > 0 ? (int64_t)last_pos + last_flag : 0;
  SSF_CHECK(nent > 0, SSF_EINVAL, "ssf_cuthill_mckee: the mesh has no edges");
  hipLaunchKernelGGL(k_rn_ptr, dim3(nblk(nent)), dim3(TPB), 0, s, nent, (int)nn, row, ptr);
  // degrees, the start node
  SSF_HIP(hipMemsetAsync(maxdeg, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_rn_startkey, dim3(nblk(nn)), dim3(TPB), 0, s, (int)nn, ptr, lkey, maxdeg);
  SSF_HIP(hipcub::DeviceReduce::Min(tmp, t4, lkey, slkey, (int)nn, s));
  int hmaxdeg = 0;
  SSF_HIP(hipMemcpyAsync(&hmaxdeg, maxdeg, sizeof(int), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  // neighbour lists by degree, stable (ties in node order)
  const int dbits = bits_for((uint64_t)hmaxdeg);
  SSF_CHECK(nbits + dbits <= 64, SSF_EUNSUP, "ssf_cuthill_mckee: degree too large");
  hipLaunchKernelGGL(k_rn_degkey, dim3(nblk(nent)), dim3(TPB), 0, s, nent, dbits, row, col, ptr,
                     key);
  SSF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, t3, key, skey, col, sortedcol, (int)nent, 0,
                                             nbits + dbits, s));
  const int *adj = sortedcol;
  // levels
  hipLaunchKernelGGL(k_rn_init, dim3(nblk(nn)), dim3(TPB), 0, s, (int)nn, cmk, cand);
  hipLaunchKernelGGL(k_rn_start, dim3(1), dim3(1), 0, s, slkey, cmk, inv);
  SSF_HIP(hipGetLastError());
  int lo = 0, hi = 1;
  while (hi < nn) {
    int cnt = 0;
    SSF_HIP(hipMemsetAsync(count, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_rn_expand, dim3(nblk(hi - lo)), dim3(TPB), 0, s, lo, hi, inv, ptr, adj,
                       cmk, cand, list, count);
    SSF_HIP(hipGetLastError());
    SSF_HIP(hipMemcpyAsync(&cnt, count, sizeof(int), hipMemcpyDeviceToHost, s));
    SSF_HIP(hipStreamSynchronize(s));
    SSF_CHECK(cnt > 0, SSF_EINVAL,
              "ssf_cuthill_mckee: the edge graph is not connected (the reference leaves such "
              "nodes unnumbered)");
    hipLaunchKernelGGL(k_rn_gather, dim3(nblk(cnt)), dim3(TPB), 0, s, cnt, list, cand, lkey);
    // keys: parent number (< hi) above the position (< maxdeg)
    SSF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, t3, lkey, slkey, list, sorted, cnt, 0,
                                               32 + nbits, s));
    hipLaunchKernelGGL(k_rn_number, dim3(nblk(cnt)), dim3(TPB), 0, s, cnt, hi, sorted, cmk, inv,
                       cand);
    SSF_HIP(hipGetLastError());
    lo = hi;
    hi += cnt;
  }
  std::vector<int> h(nn);
  SSF_HIP(hipMemcpyAsync(h.data(), inv, nn * sizeof(int), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  for (int64_t k = 0; k < nn; k++) ordering[reverse ? nn - 1 - k : k] = h[k];
  return SSF_OK;
}
