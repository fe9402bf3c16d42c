// C ABI of libssf_gpu: storage setup, values, assembly by entry, SpMV and the
// CG solves behind gmatrix::solve_system (see include/ssf_gpu.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>

#include <unordered_map>

#include "ssf_kernels.hpp"

namespace ssf {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// f(t) for t in [0, T), one host thread each
template <typename F>
static void parallel_tasks(int T, F f)
{
  if (T <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++) th.emplace_back([=]() { f(t); });
  for (auto &x : th) x.join();
}

// checksum of a reference-layout pattern, equal to pattern_checksum_dev's:
// sum of (value + 1) * mix64(index ^ salt) over adr and ci, wrapping
static uint64_t pattern_fingerprint(const std::vector<int64_t> &adr, const std::vector<int64_t> &ci)
{
  std::atomic<unsigned long long> h{0};
  auto sum = [&](const std::vector<int64_t> &v, unsigned long long salt) {
    parallel_for((int64_t)v.size(), [&](int64_t b, int64_t e) {
      unsigned long long s = 0;
      for (int64_t k = b; k < e; k++)
        s += (unsigned long long)(v[k] + 1) * mix64((unsigned long long)k ^ salt);
      h += s;
    });
  };
  sum(adr, kSaltAdr);
  sum(ci, kSaltCi);
  return h.load();
}

// ------------------------------------------------------------ patterns (host)

// symcomprow::initiate / comprow::initiate semantics (GEFEL/scr.cpp:112-330,
// GEFEL/cr.cpp:140-320): for every row ii the ascending, duplicate-free set of
// jj over all element pairs with both code numbers > 0 (and jj <= ii for the
// symmetric lower storage). Built per row from a DOF -> element incidence list.
int pattern_from_elements(int storage, int64_t n, int64_t ne, int64_t ndofe, const int64_t *cn,
                          std::vector<int64_t> &adr, std::vector<int64_t> &ci)
{
  const bool lower = storage == SSF_STORAGE_SYMM_COMP_ROWS;
  std::vector<int64_t> deg(n + 1, 0);
  for (int64_t k = 0; k < ne * ndofe; k++) {
    const int64_t c = cn[k];
    SSF_CHECK(c <= n, SSF_EINVAL, "code number larger than the number of unknowns");
    if (c > 0) deg[c - 1]++;
  }
  std::vector<int64_t> inc_ptr(n + 1, 0);
  for (int64_t i = 0; i < n; i++) inc_ptr[i + 1] = inc_ptr[i] + deg[i];
  std::vector<int64_t> inc(inc_ptr[n]);
  std::vector<int64_t> fillp(inc_ptr.begin(), inc_ptr.end() - 1);
  for (int64_t e = 0; e < ne; e++)
    for (int64_t i = 0; i < ndofe; i++) {
      const int64_t c = cn[e * ndofe + i];
      if (c > 0) inc[fillp[c - 1]++] = e;
    }
  // row lengths, then columns
  std::vector<int64_t> len(n, 0);
  std::vector<std::vector<int64_t>> rows_tmp;  // unused; rows are rebuilt below
  auto row_cols = [&](int64_t ii, std::vector<int64_t> &buf) {
    buf.clear();
    for (int64_t q = inc_ptr[ii]; q < inc_ptr[ii + 1]; q++) {
      const int64_t e = inc[q];
      for (int64_t j = 0; j < ndofe; j++) {
        const int64_t jj = cn[e * ndofe + j] - 1;
        if (jj < 0) continue;
        if (lower && jj > ii) continue;
        buf.push_back(jj);
      }
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  parallel_for(n, [&](int64_t b, int64_t e) {
    std::vector<int64_t> buf;
    for (int64_t i = b; i < e; i++) {
      row_cols(i, buf);
      len[i] = (int64_t)buf.size();
    }
  });
  adr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; i++) adr[i + 1] = adr[i] + len[i];
  ci.assign(adr[n], 0);
  parallel_for(n, [&](int64_t b, int64_t e) {
    std::vector<int64_t> buf;
    for (int64_t i = b; i < e; i++) {
      row_cols(i, buf);
      std::copy(buf.begin(), buf.end(), ci.begin() + adr[i]);
    }
  });
  return SSF_OK;
}

// ------------------------------------------------------------- SpMV tuning

// Deterministic default: the stream kernel when the matrix has a tiling (lanes
// per row from the mean row length, one to a few tiles per block), otherwise the
// row-vector kernel.
static SpmvTune default_tune(int64_t n, int64_t nnz, int ntiles, int nslice, bool idx16)
{
  SpmvTune t;
  const double avg = n ? (double)nnz / (double)n : 1.0;
  if (nslice > 0) {  // SELL-64: four slices per block pass, at most 2048 blocks
    t.format = 2;
    int64_t nb = std::max<int64_t>(8, std::min<int64_t>(2048, (nslice + 3) / 4));
    t.blocks = (int)((nb + 7) / 8 * 8);
    return t;
  }
  if (ntiles > 0) {
    t.format = 1;
    t.idx16 = idx16 ? 1 : 0;
    t.lanes = avg <= 24 ? 4 : avg <= 56 ? 8 : 16;
    int64_t nb = std::max<int64_t>(8, std::min<int64_t>(1024, ntiles));
    t.blocks = (int)((nb + 7) / 8 * 8);
    return t;
  }
  t.lanes = avg <= 6 ? 4 : avg <= 24 ? 8 : avg <= 56 ? 16 : avg <= 160 ? 32 : 64;
  const int64_t rows_per_group_min = 4;
  const int64_t groups = TPB / t.lanes;
  int64_t nb = (n + groups * rows_per_group_min - 1) / (groups * rows_per_group_min);
  nb = std::max<int64_t>(8, std::min<int64_t>(2048, nb));
  nb = (nb + 7) / 8 * 8;
  t.blocks = (int)nb;
  return t;
}

// the symmetric half of the stencil form: a symcomprow matrix whose stencil list
// (offsets from the row itself) is symmetric about its 0 slot; sets half_nl (the
// strictly lower slots, at most 15) and half_pad, or leaves half_nl = 0
// Block variant: the block stencil form of B = 2 or 3 unknowns per node whose
// offsets are every unknown of a set of neighbour nodes symmetric about the node
// itself (27 nodes of a hex8 mesh, 9 of a quad4 mesh); the stored entries get
// their SELL slot and row in the block in half_slot (device).
static void half_setup_block(ssf_matrix *m)
{
  const int B = m->sten_b, ns = (int)m->sten.size();
  if (ns % B != 0 || m->n % B != 0) return;
  std::vector<int> nodes;  // neighbour node offsets, ascending (sten is ascending)
  for (int k = 0; k < ns; k++) {
    const int d = m->sten[k], e = ((d % B) + B) % B, mo = (d - e) / B;
    if (e != k % B) return;  // every unknown of every neighbour node, in order
    if (e == 0) nodes.push_back(mo);
    else if (nodes.back() != mo) return;
  }
  const int nm = (int)nodes.size(), nl = nm / 2;
  if (nm % 2 == 0 || nodes[nl] != 0) return;
  for (int k = 0; k < nl; k++)
    if (nodes[k] != -nodes[nm - 1 - k]) return;
  if (!((B == 3 && nl == 13) || (B == 2 && nl == 4))) return;  // the kernels' instances
  const int64_t nn = m->n / B;
  const int64_t pad = ((nn + 63) / 64 * 64 - nodes[0] + 63) / 64 * 64;
  if (pad * B >= (int64_t(1) << 31)) return;
  // (slot, row in block) of every stored entry: blocks j < nl, entry (t, e), then
  // the diagonal block's lower triangle
  std::vector<int> slot;
  auto slot_of = [&](int d) {
    return (int)(std::lower_bound(m->sten.begin(), m->sten.end(), d) - m->sten.begin());
  };
  for (int j = 0; j < nl; j++)
    for (int t = 0; t < B; t++)
      for (int e = 0; e < B; e++) {
        slot.push_back(slot_of(B * nodes[j] + e));
        slot.push_back(t);
      }
  for (int t = 0; t < B; t++)
    for (int e = 0; e <= t; e++) {
      slot.push_back(slot_of(e));
      slot.push_back(t);
    }
  if (dalloc(&m->half_slot, slot.size(), m->ctx->stream) != hipSuccess ||
      dalloc(&m->half_m_d, nl, m->ctx->stream) != hipSuccess ||
      hipMemcpyAsync(m->half_slot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice,
                     m->ctx->stream) != hipSuccess ||
      hipMemcpyAsync(m->half_m_d, nodes.data(), nl * sizeof(int), hipMemcpyHostToDevice,
                     m->ctx->stream) != hipSuccess ||
      hipStreamSynchronize(m->ctx->stream) != hipSuccess) {
    dfree(m->half_slot, m->ctx->stream);
    dfree(m->half_m_d, m->ctx->stream);
    return;
  }
  m->half_b = B;
  m->half_m.assign(nodes.begin(), nodes.begin() + nl);
  m->half_nl = nl;
  m->half_pad = pad;
}

// values per row or node of the half form
static int64_t half_stored(const ssf_matrix *m)
{
  const int B = m->half_b;
  return B == 1 ? m->half_nl + 1 : (int64_t)m->half_nl * B * B + B * (B + 1) / 2;
}

static void half_setup(ssf_matrix *m)
{
  m->half_nl = 0;
  m->half_pad = 0;
  m->half_b = 1;
  const int ns = (int)m->sten.size();
  if (m->storage != SSF_STORAGE_SYMM_COMP_ROWS || m->nslice <= 0 || ns == 0) return;
  if (m->sten_b == 2 || m->sten_b == 3) {
    half_setup_block(m);
    return;
  }
  if (m->sten_b != 1 || ns % 2 == 0 || ns > 31) return;
  const int nl = ns / 2;
  if (m->sten[nl] != 0) return;
  for (int k = 0; k < nl; k++)
    if (m->sten[k] != -m->sten[ns - 1 - k]) return;
  const int64_t pad = ((int64_t)m->nslice * 64 - m->sten[0] + 63) / 64 * 64;
  // rows are addressed as 32-bit indices
  if (pad >= (int64_t(1) << 31)) return;
  m->half_nl = nl;
  m->half_pad = pad;
}

// the half stencil form: blocks of 256 threads, one wavefront per slice and at
// most 512 blocks (a few slices per wavefront)
static void half_default_tune(ssf_matrix *m)
{
  SpmvTune t;
  t.format = 2;
  t.lanes = 1;
  t.idx16 = 4;
  const int64_t ns = (m->n / m->half_b + 63) / 64;  // slices of 64 rows or nodes
  const int64_t nb = std::max<int64_t>(8, std::min<int64_t>(512, (ns + 3) / 4));
  t.blocks = (int)((nb + 7) / 8 * 8);
  m->tune64 = m->tune32 = t;
}

// Stream tiling: consecutive whole rows with at most STREAM_CAP entries per tile;
// 16-bit column offsets relative to each tile's first row when they all fit.
static bool build_tiles(int64_t n, const std::vector<int> &rowptr, const std::vector<int> &col,
                        std::vector<int> &tiles, std::vector<short> &col16)
{
  tiles.clear();
  col16.clear();
  tiles.push_back(0);
  int64_t cur = 0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t len = rowptr[i + 1] - rowptr[i];
    if (len > STREAM_CAP) {
      tiles.clear();
      return false;
    }
    if (cur + len > STREAM_CAP) {
      tiles.push_back((int)i);
      cur = 0;
    }
    cur += len;
  }
  if (tiles.back() != n || tiles.size() == 1) tiles.push_back((int)n);
  bool fits = true;
  col16.resize(col.size());
  for (size_t t = 0; t + 1 < tiles.size() && fits; t++) {
    const int r0 = tiles[t];
    for (int k = rowptr[r0]; k < rowptr[tiles[t + 1]]; k++) {
      const int off = col[k] - r0;
      if (off < -32768 || off > 32767) {
        fits = false;
        break;
      }
      col16[k] = (short)off;
    }
  }
  if (!fits) col16.clear();
  return true;
}

bool build_sell_patterns(int64_t n, const std::vector<int> &off, const std::vector<int> &scol,
                         std::vector<unsigned short> &pid, std::vector<int> &dict,
                         std::vector<int> &dict_off, int64_t max_dict)
{
  const int64_t nslice = (int64_t)off.size() - 1;
  const int64_t nrow = nslice * 64;
  // a hash of every layout row's (width, offsets), in parallel
  std::vector<uint64_t> h(nrow);
  auto offset_at = [&](int64_t sl, int64_t l, int64_t q) {
    const int64_t k = off[sl] + (q / SELL_V) * 64 * SELL_V + l * SELL_V + q % SELL_V;
    const int64_t rb = std::min<int64_t>(sl * 64 + l, n - 1);
    return (int)(scol[k] - rb);
  };
  parallel_for(nslice, [&](int64_t b, int64_t e) {
    for (int64_t sl = b; sl < e; sl++) {
      const int64_t w = (off[sl + 1] - off[sl]) / 64;
      for (int64_t l = 0; l < 64; l++) {
        uint64_t x = mix64((uint64_t)w);
        for (int64_t q = 0; q < w; q++) x = mix64(x ^ (uint64_t)(uint32_t)offset_at(sl, l, q));
        h[sl * 64 + l] = x;
      }
    }
  });
  // first occurrence of each pattern, compared entry by entry on equal hashes
  pid.assign(nrow, 0);
  dict.clear();
  dict_off.clear();
  std::vector<int> width;
  std::unordered_multimap<uint64_t, int> seen;
  for (int64_t sl = 0; sl < nslice; sl++) {
    const int w = (off[sl + 1] - off[sl]) / 64;
    for (int64_t l = 0; l < 64; l++) {
      const uint64_t x = h[sl * 64 + l];
      int found = -1;
      auto range = seen.equal_range(x);
      for (auto it = range.first; it != range.second && found < 0; ++it) {
        const int p = it->second;
        if (width[p] != w) continue;
        bool same = true;
        for (int q = 0; q < w && same; q++) same = dict[dict_off[p] + q] == offset_at(sl, l, q);
        if (same) found = p;
      }
      if (found < 0) {
        if ((int64_t)dict_off.size() >= 65535 || (int64_t)dict.size() + w > max_dict) return false;
        found = (int)dict_off.size();
        dict_off.push_back((int)dict.size());
        width.push_back(w);
        for (int q = 0; q < w; q++) dict.push_back(offset_at(sl, l, q));
        seen.emplace(x, found);
      }
      pid[sl * 64 + l] = (unsigned short)found;
    }
  }
  if (dict.empty()) dict.push_back(0);
  return true;
}

// SELL-64 layout of a full CSR: slices of 64 rows, each row's entries in
// column order, padded to the slice's longest row (rounded up to SELL_V) with
// value 0 at the row's own column; src maps a SELL entry to its CSR entry.
// The column offsets col - (r - r % B) of the rows r that occur in the matrix,
// ascending, when there are at most maxn of them and a layout with one slot per
// offset in every row holds at most 5 % more entries than the compact one
// (tot); else empty. B = 1 gives the offsets from the row itself; B = 2, 3 those
// from the first row of the row's block (the unknowns of one node numbered
// together), whose lists do not depend on the row's place in its block.
static std::vector<int> stencil_offsets(int64_t n, const std::vector<int> &rowptr,
                                        const std::vector<int> &col, int64_t tot, int B,
                                        int maxn)
{
  // each host thread collects the offsets of its rows (giving up past maxn),
  // then the lists are merged
  const int T = host_threads();
  std::vector<std::vector<int>> part(T);
  std::atomic<bool> over{false};
  parallel_tasks(T, [&](int t) {
    std::vector<int> &u = part[t];
    const int64_t b = n * t / T, e = n * (t + 1) / T;
    for (int64_t r = b; r < e && !over.load(std::memory_order_relaxed); r++) {
      const int base = (int)(r - r % B);
      for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
        const int o = col[k] - base;
        if (std::find(u.begin(), u.end(), o) == u.end()) {
          if ((int)u.size() == maxn) {
            over = true;
            break;
          }
          u.push_back(o);
        }
      }
    }
  });
  if (over) return {};
  std::vector<int> u;
  for (const std::vector<int> &pt : part) u.insert(u.end(), pt.begin(), pt.end());
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  if ((int)u.size() > maxn) return {};
  const int64_t w = ((int64_t)u.size() + SELL_V - 1) / SELL_V * SELL_V;
  if (u.empty() || w * 64 * ((n + 63) / 64) > tot + tot / 20) return {};
  return u;
}

bool build_sell(int64_t n, const std::vector<int> &rowptr, const std::vector<int> &col,
                std::vector<int> &off, std::vector<int> &scol, std::vector<int> &src,
                std::vector<short> &scol16, std::vector<int> *sten, int *sten_b)
{
  const int64_t nslice = (n + 63) / 64;
  off.assign(nslice + 1, 0);
  int64_t tot = 0;
  for (int64_t sl = 0; sl < nslice; sl++) {
    int64_t w = 0;
    for (int64_t r = sl * 64; r < std::min<int64_t>(n, sl * 64 + 64); r++)
      w = std::max<int64_t>(w, rowptr[r + 1] - rowptr[r]);
    w = (w + SELL_V - 1) / SELL_V * SELL_V;
    off[sl] = (int)tot;
    tot += w * 64;
    if (tot >= (1LL << 31)) return false;
  }
  off[nslice] = (int)tot;
  std::vector<int> u;
  int B = 1;
  if (sten) {
    // offsets from the row (at most 32, held in the kernel arguments), else from
    // the row's block of 2 or 3 unknowns (at most 96, read from device memory)
    u = stencil_offsets(n, rowptr, col, tot, 1, SELL_STEN_MAX);
    for (int bb = 2; bb <= 3 && u.empty(); bb++)
      if (n % bb == 0) {
        u = stencil_offsets(n, rowptr, col, tot, bb, SELL_STEN_MAX_B);
        if (!u.empty()) B = bb;
      }
    *sten = u;
    if (sten_b) *sten_b = u.empty() ? 1 : B;
  }
  if (!u.empty()) {
    // stencil form: every slice as wide as the offset list, slot k <-> u[k]
    const int64_t w = ((int64_t)u.size() + SELL_V - 1) / SELL_V * SELL_V;
    // the stencil kernel addresses values and x by 32-bit byte offsets: keep the
    // form only while both fit in 4 GiB at FP64
    if (std::max<int64_t>(w * 64 * nslice, n) * (int64_t)sizeof(double) >= (int64_t(1) << 32)) {
      sten->clear();
      if (sten_b) *sten_b = 1;
      u.clear();
    } else {
      for (int64_t sl = 0; sl <= nslice; sl++) off[sl] = (int)(sl * 64 * w);
      tot = off[nslice];
      scol.assign(tot, 0);
      src.assign(tot, -1);
      parallel_for(nslice, [&](int64_t b, int64_t e) {
        for (int64_t sl = b; sl < e; sl++)
          for (int64_t l = 0; l < 64; l++) {
            const int64_t r = sl * 64 + l;
            const int base = (int)(r - r % B);
            int k = r < n ? rowptr[r] : 0;
            const int kend = r < n ? rowptr[r + 1] : 0;
            for (int64_t q = 0; q < w; q++) {
              const int64_t s = off[sl] + (q / SELL_V) * 64 * SELL_V + l * SELL_V + q % SELL_V;
              if (q < (int64_t)u.size() && k < kend && col[k] - base == u[q]) {
                scol[s] = col[k];
                src[s] = k;
                k++;
              } else {
                scol[s] = (int)std::min<int64_t>(r, n - 1);  // absent: value 0
              }
            }
          }
      });
    }
  }
  if (u.empty()) {
  scol.assign(tot, 0);
  src.assign(tot, -1);
  parallel_for(nslice, [&](int64_t b, int64_t e) {
    for (int64_t sl = b; sl < e; sl++) {
      const int64_t w = (off[sl + 1] - off[sl]) / 64;
      for (int64_t l = 0; l < 64; l++) {
        const int64_t r = sl * 64 + l;
        const int64_t len = r < n ? rowptr[r + 1] - rowptr[r] : 0;
        for (int64_t q = 0; q < w; q++) {
          const int64_t k = off[sl] + (q / SELL_V) * 64 * SELL_V + l * SELL_V + q % SELL_V;
          if (q < len) {
            scol[k] = col[rowptr[r] + q];
            src[k] = rowptr[r] + (int)q;
          } else {
            scol[k] = (int)std::min<int64_t>(r, n - 1);  // in the slice
          }
        }
      }
    }
  });
  }
  // 16-bit column offsets from each slice's first row, when every offset fits
  std::vector<char> fits(nslice, 1);
  scol16.assign(tot, 0);
  parallel_for(nslice, [&](int64_t b, int64_t e) {
    for (int64_t sl = b; sl < e; sl++)
      for (int64_t k = off[sl]; k < off[sl + 1]; k++) {
        const int64_t o = (int64_t)scol[k] - sl * 64;
        if (o < -32768 || o > 32767) {
          fits[sl] = 0;
          break;
        }
        scol16[k] = (short)o;
      }
  });
  for (char f : fits)
    if (!f) {
      std::vector<short>().swap(scol16);
      break;
    }
  return true;
}

}  // namespace ssf

using namespace ssf;

// =================================================================== C ABI

extern "C" const char *ssf_last_error(void) { return g_err.c_str(); }

extern "C" int ssf_ctx_create(int device, ssf_ctx **out)
{
  SSF_ENTER();
  SSF_CHECK(out, SSF_EINVAL, "ssf_ctx_create: out is NULL");
  int count = 0;
  SSF_HIP(hipGetDeviceCount(&count));
  SSF_CHECK(device >= 0 && device < count, SSF_EINVAL, "ssf_ctx_create: no such device");
  SSF_HIP(hipSetDevice(device));
  ssf_ctx *c = new ssf_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    SSF_HIP(e);
  }
  *out = c;
  return SSF_OK;
}

extern "C" int ssf_ctx_destroy(ssf_ctx *ctx)
{
  SSF_ENTER();
  if (!ctx) return SSF_OK;
  SSF_CHECK(ctx->live_matrices == 0, SSF_ESTATE,
            "ssf_ctx_destroy: matrices created on this context are still alive");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);  // stream-ordered frees complete first
  (void)hipStreamDestroy(ctx->stream);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      if (ctx->prof_ev[a][b]) (void)hipEventDestroy(ctx->prof_ev[a][b]);
  delete ctx;
  return SSF_OK;
}

extern "C" int ssf_ctx_sync(ssf_ctx *ctx)
{
  SSF_ENTER();
  SSF_CHECK(ctx, SSF_EINVAL, "ssf_ctx_sync: NULL context");
  SSF_HIP(hipStreamSynchronize(ctx->stream));
  return SSF_OK;
}

extern "C" void *ssf_ctx_stream(ssf_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

static void free_workspace(Workspace &w, hipStream_t s)
{
  dfree(w.x, s); dfree(w.r, s); dfree(w.d, s); dfree(w.p, s); dfree(w.h, s); dfree(w.y, s);
  dfree(w.x32, s); dfree(w.r32, s); dfree(w.d32, s); dfree(w.p32, s); dfree(w.h32, s); dfree(w.y32, s);
  dfree(w.part, s); dfree(w.part32, s); dfree(w.st, s); dfree(w.probe, s);
  w.st32 = nullptr;
  w.mix = nullptr;
  if (w.probe_host) (void)hipHostFree(w.probe_host);
  w.probe_host = nullptr;
  if (w.st_host) (void)hipHostFree(w.st_host);
  w.st_host = nullptr;
  w.n = 0;
}

extern "C" int ssf_matrix_destroy(ssf_matrix *m)
{
  SSF_ENTER();
  if (!m) return SSF_OK;
  (void)hipSetDevice(m->ctx->device);
  // the frees are ordered after the work queued on the matrix's stream; work the
  // caller queued on other streams must have finished (as for any release)
  if (m->a_ref == m->val) m->a_ref = nullptr;  // comprow: the same allocation
  dfree(m->a_ref, m->ctx->stream);
  dfree(m->rowptr, m->ctx->stream); dfree(m->col, m->ctx->stream); dfree(m->val, m->ctx->stream); dfree(m->val32, m->ctx->stream);
  dfree(m->diag, m->ctx->stream); dfree(m->isd, m->ctx->stream); dfree(m->norm_a, m->ctx->stream);
  dfree(m->map_lo, m->ctx->stream); dfree(m->map_up, m->ctx->stream);
  dfree(m->asm_cstart, m->ctx->stream); dfree(m->asm_cidx, m->ctx->stream);
  dfree(m->tiles, m->ctx->stream); dfree(m->col16, m->ctx->stream);
  dfree(m->sell_off, m->ctx->stream); dfree(m->sell_col, m->ctx->stream); dfree(m->sell_src, m->ctx->stream); dfree(m->sell_v64, m->ctx->stream); dfree(m->sell_v32, m->ctx->stream);
  dfree(m->sell_col16, m->ctx->stream);
  dfree(m->sten_d, m->ctx->stream);
  dfree(m->half_v64, m->ctx->stream); dfree(m->half_v32, m->ctx->stream);
  dfree(m->half_slot, m->ctx->stream); dfree(m->half_m_d, m->ctx->stream);
  dfree(m->sell_pid, m->ctx->stream); dfree(m->sell_dict, m->ctx->stream);
  dfree(m->sell_dict_off, m->ctx->stream);
  dfree(m->sell_lo, m->ctx->stream); dfree(m->sell_up, m->ctx->stream);
  dfree(m->d_adr32, m->ctx->stream); dfree(m->d_ci32, m->ctx->stream);
  free_workspace(m->ws, m->ctx->stream);
  dual_free(m->dual, m->ctx->stream);
  ildl_free(m->ildl, m->ctx->stream);
  m->ctx->live_matrices--;
  delete m;
  return SSF_OK;
}

void ssf::matrix_drop_assembly(ssf_matrix *m, bool host_pattern)
{
  (void)hipStreamSynchronize(m->ctx->stream);
  dfree(m->asm_cstart, m->ctx->stream);
  dfree(m->asm_cidx, m->ctx->stream);
  m->asm_ne = 0;
  m->asm_ndofe = 0;
  m->asm_nc = 0;
  m->asm_cn_host.clear();
  m->asm_cn_host.shrink_to_fit();
  if (host_pattern) {  // the matrix can then only be solved and multiplied
    std::vector<int64_t>().swap(m->adr);
    std::vector<int64_t>().swap(m->ci);
    if (m->a_ref == m->val) m->a_ref = nullptr;
    dfree(m->a_ref, m->ctx->stream);
    dfree(m->map_lo, m->ctx->stream);
    dfree(m->map_up, m->ctx->stream);
  }
}

void ssf::matrix_keep_sell_only(ssf_matrix *m)
{
  (void)hipStreamSynchronize(m->ctx->stream);
  dfree(m->col, m->ctx->stream);
  dfree(m->val, m->ctx->stream);
  dfree(m->val32, m->ctx->stream);
  dfree(m->isd, m->ctx->stream);
  dfree(m->tiles, m->ctx->stream);
  dfree(m->col16, m->ctx->stream);
  dfree(m->sell_src, m->ctx->stream);
  dfree(m->sell_v32, m->ctx->stream);
  dfree(m->half_v32, m->ctx->stream);
  if (m->sell_col16) dfree(m->sell_col, m->ctx->stream);
  dfree(m->sell_lo, m->ctx->stream);
  dfree(m->sell_up, m->ctx->stream);
  dfree(m->d_adr32, m->ctx->stream);
  dfree(m->d_ci32, m->ctx->stream);
  dfree(m->a_ref, m->ctx->stream);
  m->ntiles = 0;
  m->val32_valid = false;
}

// Full CSR of a symcomprow pattern, built in parallel: row i = the lower part of
// scr row i (ascending, diagonal last), then the upper part: the rows j > i that
// hold column i, in ascending j. The upper parts are a counting-sort transpose
// over contiguous chunks of rows: chunk t's entries of column i follow those of
// chunks < t, so every column's rows stay ascending. lo[k] / up[k] give the full
// slots of scr entry k (up = -1 on the diagonal).
static bool full_from_scr(int64_t n, const int64_t *adr, const int64_t *ci,
                          std::vector<int> &rowptr, std::vector<int> &col, std::vector<int> &lo,
                          std::vector<int> &up)
{
  const int64_t negm = adr[n];
  const int T = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / 4096 + 1));
  std::vector<int> cc((size_t)T * n, 0);  // entries of column i in chunk t (strictly lower)
  auto chunk_rows = [&](int t, int64_t &b, int64_t &e) {
    b = n * t / T;
    e = n * (t + 1) / T;
  };
  parallel_tasks(T, [&](int t) {
    int64_t b, e;
    chunk_rows(t, b, e);
    int *c = cc.data() + (size_t)t * n;
    for (int64_t j = b; j < e; j++)
      for (int64_t k = adr[j]; k < adr[j + 1]; k++)
        if (ci[k] != j) c[ci[k]]++;
  });
  rowptr.assign(n + 1, 0);
  std::vector<int> ucnt(n);
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      int u = 0;
      for (int t = 0; t < T; t++) u += cc[(size_t)t * n + i];
      ucnt[i] = u;
    }
  });
  int64_t tot = 0;
  for (int64_t i = 0; i < n; i++) {
    rowptr[i] = (int)tot;
    tot += (adr[i + 1] - adr[i]) + ucnt[i];
    if (tot >= (1LL << 31)) return false;
  }
  rowptr[n] = (int)tot;
  // first upper slot of column i for chunk t
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      int run = rowptr[i] + (int)(adr[i + 1] - adr[i]);
      for (int t = 0; t < T; t++) {
        const int c = cc[(size_t)t * n + i];
        cc[(size_t)t * n + i] = run;
        run += c;
      }
    }
  });
  col.assign(tot, 0);
  lo.assign(negm, 0);
  up.assign(negm, -1);
  parallel_tasks(T, [&](int t) {
    int64_t b, e;
    chunk_rows(t, b, e);
    int *c = cc.data() + (size_t)t * n;
    for (int64_t j = b; j < e; j++)
      for (int64_t k = adr[j]; k < adr[j + 1]; k++) {
        const int64_t s = rowptr[j] + (k - adr[j]);
        col[s] = (int)ci[k];
        lo[k] = (int)s;
        if (ci[k] != j) {
          const int u = c[ci[k]]++;
          col[u] = (int)j;
          up[k] = u;
        }
      }
  });
  return true;
}

// validation of a reference-layout pattern: ascending, in range; symcomprow rows
// hold the lower triangle and end with their diagonal
static const char *check_pattern(int64_t n, const int64_t *adr, const int64_t *ci, bool sym)
{
  std::atomic<int> bad{0};
  parallel_for(n, [&](int64_t b, int64_t e) {
    int code = 0;
    for (int64_t i = b; i < e && !code; i++) {
      if (adr[i + 1] < adr[i]) { code = 1; break; }
      for (int64_t k = adr[i]; k < adr[i + 1]; k++) {
        if (ci[k] < 0 || ci[k] >= n) { code = 2; break; }
        if (k > adr[i] && ci[k] <= ci[k - 1]) { code = 3; break; }
        if (sym && ci[k] > i) { code = 4; break; }
      }
      if (!code && sym && !(adr[i + 1] > adr[i] && ci[adr[i + 1] - 1] == i)) code = 5;
    }
    if (code) {
      int z = 0;
      bad.compare_exchange_strong(z, code);
    }
  });
  switch (bad.load()) {
    case 1: return "ssf_matrix_create_csr: adr not monotone";
    case 2: return "ssf_matrix_create_csr: column index out of range";
    case 3: return "ssf_matrix_create_csr: columns not ascending";
    case 4: return "ssf_matrix_create_csr: symm_comp_rows holds the lower triangle only";
    case 5: return "ssf_matrix_create_csr: every symm_comp_rows row must end with its diagonal";
    default: return nullptr;
  }
}

// Device storage of a pattern. adr_own / ci_own (may be null) are moved into the
// matrix instead of copied; stream_tiles = false skips the tiling of the CSR
// stream kernel (matrices solved through the SELL layout only).
int ssf::matrix_create(ssf_ctx *ctx, int storage, int64_t n, const int64_t *adr,
                       const int64_t *ci, std::vector<int64_t> *adr_own,
                       std::vector<int64_t> *ci_own, bool stream_tiles, ssf_matrix **out)
{
  SSF_CHECK(ctx && out && adr && (ci || n == 0), SSF_EINVAL, "ssf_matrix_create_csr: NULL argument");
  SSF_CHECK(storage == SSF_STORAGE_SYMM_COMP_ROWS || storage == SSF_STORAGE_COMPRESSED_ROWS,
            SSF_EUNSUP, "ssf_matrix_create_csr: storage must be symm_comp_rows (11) or compressed_rows (10)");
  SSF_CHECK(n >= 0 && n < (1LL << 31), SSF_EINVAL, "ssf_matrix_create_csr: n out of range");
  SSF_CHECK(adr[0] == 0, SSF_EINVAL, "ssf_matrix_create_csr: adr[0] must be 0");
  const int64_t negm = adr[n];
  const bool sym = storage == SSF_STORAGE_SYMM_COMP_ROWS;
  if (const char *msg = check_pattern(n, adr, ci, sym)) {
    set_error(msg);
    return SSF_EINVAL;
  }
  SSF_HIP(hipSetDevice(ctx->device));
  ssf_matrix *m = new ssf_matrix();
  m->ctx = ctx;
  ctx->live_matrices++;
  m->storage = storage;
  m->n = n;
  m->negm = negm;
  std::vector<int> rowptr(n + 1), col, lo, up;
  if (!sym) {
    for (int64_t i = 0; i <= n; i++) rowptr[i] = (int)adr[i];
    col.resize(negm);
    parallel_for(negm, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; k++) col[k] = (int)ci[k];
    });
    m->nnz = negm;
  } else {
    if (!full_from_scr(n, adr, ci, rowptr, col, lo, up)) {
      ssf_matrix_destroy(m);
      set_error("ssf_matrix_create_csr: more than 2^31 entries in the full matrix");
      return SSF_EUNSUP;
    }
    m->nnz = rowptr[n];
  }
  if (adr_own && ci_own) {
    m->adr.swap(*adr_own);
    m->ci.swap(*ci_own);
  } else {
    m->adr.assign(adr, adr + n + 1);
    m->ci.assign(ci, ci + negm);
  }
  m->pattern_fp = pattern_fingerprint(m->adr, m->ci);
  std::vector<int> tiles;
  std::vector<short> col16;
  const bool have_tiles = stream_tiles && build_tiles(n, rowptr, col, tiles, col16);
  std::vector<int> soff, scol, ssrc;
  std::vector<short> scol16;
  const bool have_sell = n > 0 && build_sell(n, rowptr, col, soff, scol, ssrc, scol16, &m->sten,
                                             &m->sten_b);
  std::vector<int> sten_pad;  // (copied to the device below; lives until the final sync)
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  chk(dalloc(&m->rowptr, n + 1, ctx->stream));
  chk(dalloc(&m->col, m->nnz, ctx->stream));
  chk(dalloc(&m->val, m->nnz, ctx->stream));
  chk(dalloc(&m->diag, n, ctx->stream));
  if (have_tiles) {
    m->ntiles = (int)tiles.size() - 1;
    chk(dalloc(&m->tiles, tiles.size(), ctx->stream));
    if (e == hipSuccess)
      chk(hipMemcpyAsync(m->tiles, tiles.data(), tiles.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (!col16.empty()) {
      chk(dalloc(&m->col16, col16.size(), ctx->stream));
      if (e == hipSuccess)
        chk(hipMemcpyAsync(m->col16, col16.data(), col16.size() * sizeof(short), hipMemcpyHostToDevice, ctx->stream));
    }
  }
  if (have_sell) {
    m->nslice = (int)soff.size() - 1;
    m->sell_tot = soff.back();
    chk(dalloc(&m->sell_off, soff.size(), ctx->stream));
    chk(dalloc(&m->sell_col, m->sell_tot, ctx->stream));
    chk(dalloc(&m->sell_src, m->sell_tot, ctx->stream));
    chk(dalloc(&m->sell_v64, m->sell_tot, ctx->stream));
    // the offsets of every slot of a row: the list padded with 0 to the slot count
    sten_pad = m->sten;
    sten_pad.resize((m->sten.size() + SELL_V - 1) / SELL_V * SELL_V, 0);
    if (!sten_pad.empty()) chk(dalloc(&m->sten_d, sten_pad.size(), ctx->stream));
    if (e == hipSuccess) {
      if (!sten_pad.empty())
        chk(hipMemcpyAsync(m->sten_d, sten_pad.data(), sten_pad.size() * sizeof(int),
                           hipMemcpyHostToDevice, ctx->stream));
      chk(hipMemcpyAsync(m->sell_off, soff.data(), soff.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
      chk(hipMemcpyAsync(m->sell_col, scol.data(), scol.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
      chk(hipMemcpyAsync(m->sell_src, ssrc.data(), ssrc.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
      if (!scol16.empty()) {
        chk(dalloc(&m->sell_col16, m->sell_tot, ctx->stream));
        if (e == hipSuccess)
          chk(hipMemcpyAsync(m->sell_col16, scol16.data(), scol16.size() * sizeof(short),
                             hipMemcpyHostToDevice, ctx->stream));
      }
      chk(hipMemsetAsync(m->sell_v64, 0, std::max<int64_t>(1, m->sell_tot) * sizeof(double), ctx->stream));
    }
    // row patterns: the column stream of a mesh matrix repeats a few offset lists
    std::vector<unsigned short> pid;
    std::vector<int> dict, dict_off;
    // (the pattern kernel addresses values and x by 32-bit byte offsets)
    const bool fits32 = std::max<int64_t>(m->sell_tot, n) * (int64_t)sizeof(double) < (int64_t(1) << 32);
    if (e == hipSuccess && fits32 && build_sell_patterns(n, soff, scol, pid, dict, dict_off, 1 << 22)) {
      m->sell_npat = (int)dict_off.size();
      chk(dalloc(&m->sell_pid, pid.size(), ctx->stream));
      chk(dalloc(&m->sell_dict, dict.size(), ctx->stream));
      chk(dalloc(&m->sell_dict_off, dict_off.size(), ctx->stream));
      if (e == hipSuccess) {
        chk(hipMemcpyAsync(m->sell_pid, pid.data(), pid.size() * sizeof(unsigned short),
                           hipMemcpyHostToDevice, ctx->stream));
        chk(hipMemcpyAsync(m->sell_dict, dict.data(), dict.size() * sizeof(int),
                           hipMemcpyHostToDevice, ctx->stream));
        chk(hipMemcpyAsync(m->sell_dict_off, dict_off.data(), dict_off.size() * sizeof(int),
                           hipMemcpyHostToDevice, ctx->stream));
        chk(hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope
      }
    }
  }
  if (sym) {
    chk(dalloc(&m->a_ref, negm, ctx->stream));
    chk(dalloc(&m->map_lo, negm, ctx->stream));
    chk(dalloc(&m->map_up, negm, ctx->stream));
  } else {
    m->a_ref = m->val;
  }
  if (e == hipSuccess) {
    chk(hipMemcpyAsync(m->rowptr, rowptr.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (m->nnz) chk(hipMemcpyAsync(m->col, col.data(), m->nnz * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (sym && negm) {
      chk(hipMemcpyAsync(m->map_lo, lo.data(), negm * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
      chk(hipMemcpyAsync(m->map_up, up.data(), negm * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    }
    chk(hipMemsetAsync(m->a_ref, 0, std::max<int64_t>(1, negm) * sizeof(double), ctx->stream));
    chk(hipMemsetAsync(m->val, 0, std::max<int64_t>(1, m->nnz) * sizeof(double), ctx->stream));
    chk(hipMemsetAsync(m->diag, 0, std::max<int64_t>(1, n) * sizeof(double), ctx->stream));
    chk(hipStreamSynchronize(ctx->stream));
  }
  if (e != hipSuccess) {
    const std::string msg = std::string("ssf_matrix_create_csr: ") + hipGetErrorString(e);
    ssf_matrix_destroy(m);
    set_error(msg);
    return e == hipErrorOutOfMemory ? SSF_ENOMEM : SSF_EHIP;
  }
  // canonical int32 columns for the FP64 drop-in, 16-bit offsets where they fit
  // for the FP32 inner solve
  m->tune64 = default_tune(n, m->nnz, m->ntiles, m->nslice, false);
  m->tune32 = default_tune(n, m->nnz, m->ntiles, m->nslice, m->col16 != nullptr);
  if (m->sell_col16) m->tune64.idx16 = m->tune32.idx16 = (m->tune64.format == 2) ? 1 : 0;
  if (m->sell_pid && m->tune64.format == 2) m->tune64.idx16 = m->tune32.idx16 = 2;
  if (!m->sten.empty() && m->tune64.format == 2) m->tune64.idx16 = m->tune32.idx16 = 3;
  half_setup(m);
  if (m->half_nl > 0 && m->tune64.format == 2) half_default_tune(m);
  *out = m;
  return SSF_OK;
}

int ssf::matrix_create_sell_dev(ssf_ctx *ctx, int64_t n, int *d_adr, int *d_ci, int64_t negm,
                                ssf_matrix **out, bool half_form)
{
  hipStream_t s = ctx->stream;
  ssf_matrix *m = new ssf_matrix();
  m->ctx = ctx;
  ctx->live_matrices++;
  m->storage = SSF_STORAGE_SYMM_COMP_ROWS;
  m->n = n;
  m->negm = negm;
  m->nnz = 2 * negm - n;
  m->sell_only = true;
  m->d_adr32 = d_adr;
  m->d_ci32 = d_ci;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  chk(pattern_checksum_dev(d_adr, n, d_ci, negm, s, &m->pattern_fp));
  SellLayout L;
  if (e == hipSuccess) chk(sell_layout_build((int)n, negm, m->d_adr32, m->d_ci32, s, &L));
  if (e == hipSuccess) {
    m->nslice = L.nslice;
    m->sell_tot = L.tot;
    m->sell_off = L.off;
    m->sell_col = L.col;
    m->sell_col16 = L.col16;
    m->sell_lo = L.lo;
    m->sell_up = L.up;
    m->sten = L.sten_h;  // stencil form (no column arrays) when the layout took it
    m->sten_b = L.sten ? L.sten_b : 1;
    m->sten_d = L.sten;
    chk(dalloc(&m->a_ref, negm, s));
    chk(dalloc(&m->sell_v64, m->sell_tot, s));
    chk(dalloc(&m->diag, n, s));
  }
  if (e == hipSuccess) {
    chk(hipMemsetAsync(m->a_ref, 0, negm * sizeof(double), s));
    chk(hipMemsetAsync(m->sell_v64, 0, std::max<int64_t>(1, m->sell_tot) * sizeof(double), s));
    chk(hipMemsetAsync(m->diag, 0, n * sizeof(double), s));
    chk(hipStreamSynchronize(s));
  }
  if (e != hipSuccess) {
    const std::string msg = std::string("matrix_create_sell: ") + hipGetErrorString(e);
    ssf_matrix_destroy(m);
    set_error(msg);
    return e == hipErrorOutOfMemory ? SSF_ENOMEM : SSF_EHIP;
  }
  SpmvTune t;
  t.format = 2;
  t.blocks = (int)((std::max<int64_t>(8, std::min<int64_t>(2048, (m->nslice + 3) / 4)) + 7) / 8 * 8);
  t.idx16 = !m->sten.empty() ? 3 : m->sell_col16 ? 1 : 0;
  m->tune64 = m->tune32 = t;
  if (half_form) half_setup(m);
  if (m->half_nl > 0) half_default_tune(m);
  *out = m;
  return SSF_OK;
}

int ssf::matrix_create_sell(ssf_ctx *ctx, int64_t n, std::vector<int64_t> &adr,
                            std::vector<int64_t> &ci, ssf_matrix **out)
{
  SSF_CHECK(ctx && out && (int64_t)adr.size() == n + 1, SSF_EINVAL, "matrix_create_sell: bad argument");
  SSF_CHECK(n > 0 && n < (1LL << 31) && adr[n] < (1LL << 31) && 2 * adr[n] < (1LL << 31),
            SSF_EUNSUP, "matrix_create_sell: more than 2^31 entries");
  if (const char *msg = check_pattern(n, adr.data(), ci.data(), true)) {
    set_error(msg);
    return SSF_EINVAL;
  }
  SSF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t negm = adr[n];
  std::vector<int> adr32(n + 1), ci32(negm);
  parallel_for(n + 1, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) adr32[i] = (int)adr[i];
  });
  parallel_for(negm, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; k++) ci32[k] = (int)ci[k];
  });
  int *d_adr = nullptr, *d_ci = nullptr;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  chk(dalloc(&d_adr, n + 1, s));
  chk(dalloc(&d_ci, negm, s));
  if (e == hipSuccess) {
    chk(hipMemcpyAsync(d_adr, adr32.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_ci, ci32.data(), negm * sizeof(int), hipMemcpyHostToDevice, s));
    chk(hipStreamSynchronize(s));
  }
  if (e != hipSuccess) {
    dfree(d_adr, s);
    dfree(d_ci, s);
    SSF_HIP(e);
  }
  int rc = matrix_create_sell_dev(ctx, n, d_adr, d_ci, negm, out);
  if (rc) return rc;
  (*out)->adr.swap(adr);
  (*out)->ci.swap(ci);
  return SSF_OK;
}

extern "C" int ssf_matrix_create_csr(ssf_ctx *ctx, int storage, int64_t n, const int64_t *adr,
                                     const int64_t *ci, ssf_matrix **out)
{
  SSF_ENTER();
  return matrix_create(ctx, storage, n, adr, ci, nullptr, nullptr, true, out);
}

extern "C" int ssf_matrix_create_elem(ssf_ctx *ctx, int storage, int64_t n, int64_t ne,
                                      int64_t ndofe, const int64_t *cn, ssf_matrix **out)
{
  SSF_ENTER();
  SSF_CHECK(ctx && out && (cn || ne == 0), SSF_EINVAL, "ssf_matrix_create_elem: NULL argument");
  SSF_CHECK(storage == SSF_STORAGE_SYMM_COMP_ROWS || storage == SSF_STORAGE_COMPRESSED_ROWS,
            SSF_EUNSUP, "ssf_matrix_create_elem: storage must be symm_comp_rows (11) or compressed_rows (10)");
  SSF_CHECK(n >= 0 && ne >= 0 && ndofe > 0, SSF_EINVAL, "ssf_matrix_create_elem: bad sizes");
  std::vector<int64_t> adr, ci;
  int rc = pattern_from_elements(storage, n, ne, ndofe, cn, adr, ci);
  if (rc != SSF_OK) return rc;
  if (storage == SSF_STORAGE_SYMM_COMP_ROWS)
    for (int64_t i = 0; i < n; i++)
      SSF_CHECK(adr[i + 1] > adr[i] && ci[adr[i + 1] - 1] == i, SSF_EINVAL,
                "ssf_matrix_create_elem: unknown " + std::to_string(i + 1) + " belongs to no element");
  SSF_ENTER();
  return matrix_create(ctx, storage, n, adr.data(), ci.data(), &adr, &ci, true, out);
}

extern "C" int ssf_matrix_create_elem_device(ssf_ctx *ctx, int64_t n, int64_t ne, int64_t ndofe,
                                             const int64_t *cn, ssf_matrix **out)
{
  SSF_ENTER();
  SSF_CHECK(ctx && out && cn, SSF_EINVAL, "ssf_matrix_create_elem_device: NULL argument");
  *out = nullptr;
  SSF_CHECK(n > 0 && n < (1LL << 31) - 1 && ne > 0 && ndofe > 0 && ndofe <= 64 &&
                ne * ndofe < (1LL << 31),
            SSF_EINVAL, "ssf_matrix_create_elem_device: bad sizes");
  std::vector<int> cn32(ne * ndofe);
  std::atomic<int> bad{0};
  parallel_for(ne * ndofe, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; k++) {
      if (cn[k] > n || cn[k] < -(1LL << 31) + 1) bad = 1;
      cn32[k] = (int)cn[k];
    }
  });
  SSF_CHECK(!bad.load(), SSF_EINVAL, "ssf_matrix_create_elem_device: code number out of range");
  SSF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  int *d_cn = nullptr;
  SSF_HIP(dalloc(&d_cn, ne * ndofe, s));
  hipError_t e = hipMemcpyAsync(d_cn, cn32.data(), ne * ndofe * sizeof(int), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    dfree(d_cn, s);
    SSF_HIP(e);
  }
  int *d_adr = nullptr, *d_ci = nullptr;
  int64_t negm = 0;
  int rc = pattern_build_dev(n, ne, (int)ndofe, d_cn, s, &d_adr, &d_ci, &negm);
  dfree(d_cn, s);
  if (rc) return rc;
  if (negm >= (1LL << 30)) {
    dfree(d_adr, s);
    dfree(d_ci, s);
    set_error("ssf_matrix_create_elem_device: more than 2^31 entries in the full matrix");
    return SSF_EUNSUP;
  }
  return matrix_create_sell_dev(ctx, n, d_adr, d_ci, negm, out);
}

extern "C" int ssf_matrix_sizes(const ssf_matrix *m, int64_t *n, int64_t *negm, int64_t *nnz_full)
{
  SSF_ENTER();
  SSF_CHECK(m, SSF_EINVAL, "ssf_matrix_sizes: NULL matrix");
  if (n) *n = m->n;
  if (negm) *negm = m->negm;
  if (nnz_full) *nnz_full = m->nnz;
  return SSF_OK;
}

extern "C" int ssf_matrix_pattern(const ssf_matrix *m, int64_t *adr, int64_t *ci)
{
  SSF_ENTER();
  SSF_CHECK(m, SSF_EINVAL, "ssf_matrix_pattern: NULL matrix");
  if ((int64_t)m->adr.size() == m->n + 1) {
    if (adr) std::memcpy(adr, m->adr.data(), (m->n + 1) * sizeof(int64_t));
    if (ci && m->negm) std::memcpy(ci, m->ci.data(), m->negm * sizeof(int64_t));
    return SSF_OK;
  }
  // pattern kept on the device only
  SSF_CHECK(m->d_adr32 && m->d_ci32, SSF_ESTATE, "ssf_matrix_pattern: the pattern was released");
  SSF_HIP(hipSetDevice(m->ctx->device));
  std::vector<int> a32(m->n + 1), c32(m->negm);
  SSF_HIP(hipMemcpyAsync(a32.data(), m->d_adr32, (m->n + 1) * sizeof(int), hipMemcpyDeviceToHost,
                         m->ctx->stream));
  if (m->negm)
    SSF_HIP(hipMemcpyAsync(c32.data(), m->d_ci32, m->negm * sizeof(int), hipMemcpyDeviceToHost,
                           m->ctx->stream));
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  if (adr) std::copy(a32.begin(), a32.end(), adr);
  if (ci) std::copy(c32.begin(), c32.end(), ci);
  return SSF_OK;
}

// (re)builds the half form of the FP64 (precision 0) or FP32 SELL values
static int half_build(ssf_matrix *m, int precision)
{
  if (m->half_nl == 0) return SSF_OK;
  hipStream_t s = m->ctx->stream;
  const int nv = (int)half_stored(m);
  const int64_t tot = nv * m->half_pad;
  if (precision == 0) {
    if (!m->sell_v64) return SSF_OK;
    if (!m->half_v64) SSF_HIP(dalloc(&m->half_v64, tot, s));
    if (m->half_b > 1)
      SSF_HIP(half_from_sell_block<double>(m->csr(), m->half_slot, nv, m->sell_v64, m->half_v64, s));
    else
      SSF_HIP(half_from_sell<double>(m->csr(), m->sell_v64, m->half_v64, s));
  } else {
    if (!m->sell_v32) return SSF_OK;
    if (!m->half_v32) SSF_HIP(dalloc(&m->half_v32, tot, s));
    if (m->half_b > 1)
      SSF_HIP(half_from_sell_block<float>(m->csr(), m->half_slot, nv, m->sell_v32, m->half_v32, s));
    else
      SSF_HIP(half_from_sell<float>(m->csr(), m->sell_v32, m->half_v32, s));
  }
  return SSF_OK;
}

int ssf::matrix_half_build(ssf_matrix *m, int precision) { return half_build(m, precision); }

int64_t ssf::matrix_half_entries(const ssf_matrix *m)
{
  return m->half_nl > 0 ? half_stored(m) * m->half_pad : 0;
}

// after the reference-layout values changed: full CSR values, diagonal
static int values_changed(ssf_matrix *m)
{
  hipStream_t s = m->ctx->stream;
  m->values_version++;
  if (m->sell_only) {
    SSF_HIP(sell_expand((int)m->n, m->negm, m->d_adr32, m->sell_lo, m->sell_up, m->a_ref,
                        m->sell_v64, m->diag, s));
    m->values_set = true;
    m->val32_valid = false;
    return half_build(m, 0);
  }
  if (m->storage == SSF_STORAGE_SYMM_COMP_ROWS) SSF_HIP(expand_ref(m->negm, m->map_lo, m->map_up, m->a_ref, m->val, s));
  SSF_HIP(extract_diag((int)m->n, m->rowptr, m->col, m->val, m->diag, nullptr, s));
  if (m->sell_src) SSF_HIP(sell_gather(m->sell_tot, m->sell_src, m->val, m->sell_v64, s));
  m->values_set = true;
  m->val32_valid = false;
  return half_build(m, 0);
}

extern "C" int ssf_matrix_zero(ssf_matrix *m)
{
  SSF_ENTER();
  SSF_CHECK(m, SSF_EINVAL, "ssf_matrix_zero: NULL matrix");
  SSF_HIP(hipSetDevice(m->ctx->device));
  SSF_HIP(hipMemsetAsync(m->a_ref, 0, std::max<int64_t>(1, m->negm) * sizeof(double), m->ctx->stream));
  int rc = values_changed(m);
  if (rc) return rc;
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  return SSF_OK;
}

extern "C" int ssf_matrix_set_values(ssf_matrix *m, const double *a)
{
  SSF_ENTER();
  SSF_CHECK(m && (a || m->negm == 0), SSF_EINVAL, "ssf_matrix_set_values: NULL argument");
  SSF_HIP(hipSetDevice(m->ctx->device));
  if (m->negm) SSF_HIP(hipMemcpyAsync(m->a_ref, a, m->negm * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
  int rc = values_changed(m);
  if (rc) return rc;
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  return SSF_OK;
}

extern "C" int ssf_matrix_get_values(const ssf_matrix *m, double *a)
{
  SSF_ENTER();
  SSF_CHECK(m && (a || m->negm == 0), SSF_EINVAL, "ssf_matrix_get_values: NULL argument");
  SSF_HIP(hipSetDevice(m->ctx->device));
  if (m->negm) SSF_HIP(hipMemcpyAsync(a, m->a_ref, m->negm * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  return SSF_OK;
}

// ----------------------------------------------------------------- assembly

// The assembly plan of a code-number table (rebuilt when the code numbers
// change): the contributions of every element-matrix entry to the reference
// layout, listed per entry in the reference's localize order (ssf_assemble.hip).
// d_cn_out receives the table as int32 on the device (the caller frees it).
static int prepare_assembly(ssf_matrix *m, int64_t ne, int nd, const int64_t *cn, int **d_cn_out)
{
  hipStream_t s = m->ctx->stream;
  const int64_t ncn = ne * nd;
  bool same = m->asm_cstart && m->asm_ne == ne && m->asm_ndofe == nd &&
              (int64_t)m->asm_cn_host.size() == ncn &&
              std::equal(m->asm_cn_host.begin(), m->asm_cn_host.end(), cn);
  std::vector<int> cn32(ncn);
  std::atomic<int> too_large{0};
  parallel_for(ncn, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; k++) {
      if (cn[k] > m->n) too_large = 1;
      cn32[k] = (int)cn[k];
    }
  });
  SSF_CHECK(!too_large.load(), SSF_EINVAL, "code number larger than the number of unknowns");
  int *d_cn = nullptr;
  SSF_HIP(dalloc(&d_cn, ncn, s));
  SSF_HIP(hipMemcpyAsync(d_cn, cn32.data(), ncn * sizeof(int), hipMemcpyHostToDevice, s));
  *d_cn_out = d_cn;
  if (same) {
    SSF_HIP(hipStreamSynchronize(s));  // cn32 is released on return
    return SSF_OK;
  }
  dfree(m->asm_cstart, s);
  dfree(m->asm_cidx, s);
  m->asm_ne = 0;
  const int *d_adr = m->d_adr32, *d_ci = m->d_ci32;  // the device pattern, if kept
  int *t_adr = nullptr, *t_ci = nullptr;
  if (!d_adr) {  // reference-layout pattern for the entry search
    std::vector<int> adr32(m->n + 1), ci32(m->negm);
    parallel_for(m->n + 1, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) adr32[i] = (int)m->adr[i];
    });
    parallel_for(m->negm, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; k++) ci32[k] = (int)m->ci[k];
    });
    SSF_HIP(dalloc(&t_adr, m->n + 1, s));
    SSF_HIP(dalloc(&t_ci, m->negm, s));
    SSF_HIP(hipMemcpyAsync(t_adr, adr32.data(), (m->n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    if (m->negm)
      SSF_HIP(hipMemcpyAsync(t_ci, ci32.data(), m->negm * sizeof(int), hipMemcpyHostToDevice, s));
    SSF_HIP(hipStreamSynchronize(s));  // the host copies go out of scope
    d_adr = t_adr;
    d_ci = t_ci;
  }
  int rc = asm_plan_build(m->storage, m->n, m->negm, ne, nd, d_cn, d_adr, d_ci, s,
                          &m->asm_cstart, &m->asm_cidx, &m->asm_nc);
  dfree(t_adr, s);
  dfree(t_ci, s);
  if (rc) {
    dfree(d_cn, s);
    *d_cn_out = nullptr;
    return rc;
  }
  m->asm_ne = ne;
  m->asm_ndofe = nd;
  m->asm_cn_host.assign(cn, cn + ncn);
  return SSF_OK;
}

extern "C" int ssf_matrix_assemble(ssf_matrix *m, int kind, int64_t nn, const double *xyz,
                                   int64_t ne, const int64_t *conn, const int64_t *cn,
                                   const double *mat_a, const double *mat_b,
                                   const double *thick, const double *pv, int64_t npv,
                                   double *rhs_lift)
{
  SSF_ENTER();
  SSF_CHECK(m && xyz && conn && cn && mat_a, SSF_EINVAL, "ssf_matrix_assemble: NULL argument");
  SSF_CHECK(kind >= 0 && kind <= 3, SSF_EINVAL, "ssf_matrix_assemble: unknown element kind");
  SSF_CHECK(kind == SSF_ELEM_HEX8_HEAT || kind == SSF_ELEM_HEX8_HEAT_CAP || mat_b, SSF_EINVAL,
            "ssf_matrix_assemble: Poisson ratios missing");
  SSF_CHECK(ne >= 0 && nn >= 0 && ne < (1LL << 31) && nn < (1LL << 31), SSF_EINVAL, "ssf_matrix_assemble: bad sizes");
  const int nne = kind == SSF_ELEM_QUAD4_PSTRESS ? 4 : 8;
  const int nd = kind == SSF_ELEM_HEX8_ELAST ? 24 : 8;
  SSF_CHECK((int64_t)ne * nd * nd < (1LL << 31), SSF_EUNSUP,
            "ssf_matrix_assemble: more than 2^31 element-matrix entries");
  {
    std::atomic<int> bad{0};
    parallel_for(ne * nne, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; k++)
        if (conn[k] < 0 || conn[k] >= nn) bad = 1;
    });
    SSF_CHECK(!bad.load(), SSF_EINVAL, "ssf_matrix_assemble: element node index out of range");
  }
  if (rhs_lift)  // prescribed values are read by the lift only
    for (int64_t k = 0; k < ne * nd; k++)
      if (cn[k] < 0) SSF_CHECK(pv && -cn[k] <= npv, SSF_EINVAL, "ssf_matrix_assemble: prescribed value index out of range");
  SSF_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  int *d_cn = nullptr;
  int rc = prepare_assembly(m, ne, nd, cn, &d_cn);
  if (rc) return rc;
  // the lift: elements with prescribed values and their rows, per unknown
  std::vector<int> pidx, lstart, lidx;
  int64_t npe = 0;
  if (rhs_lift) lift_plan_host(m->n, ne, nd, cn, pidx, lstart, lidx, &npe);
  std::vector<int> conn32(ne * nne);
  parallel_for(ne * nne, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; k++) conn32[k] = (int)conn[k];
  });
  int *d_conn = nullptr, *d_pidx = nullptr, *d_lstart = nullptr, *d_lidx = nullptr;
  double *d_xyz = nullptr, *d_a = nullptr, *d_b = nullptr, *d_t = nullptr, *d_pv = nullptr, *d_lift = nullptr;
  double *d_kbuf = nullptr, *d_fbuf = nullptr;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  const int wsh = asm_wshift(kind);
  chk(dalloc(&d_kbuf, asm_kbuf_count(ne, nd, wsh), s));
  if (rhs_lift) {
    chk(dalloc(&d_pidx, ne, s));
    chk(dalloc(&d_lstart, m->n + 1, s));
    chk(dalloc(&d_lidx, lidx.size(), s));
    chk(dalloc(&d_fbuf, npe * nd, s));
  }
  chk(dalloc(&d_conn, ne * nne, s));
  chk(dalloc(&d_xyz, nn * 3, s));
  chk(dalloc(&d_a, ne, s));
  if (mat_b) chk(dalloc(&d_b, ne, s));
  if (thick) chk(dalloc(&d_t, ne, s));
  if (pv && npv) chk(dalloc(&d_pv, npv, s));
  if (rhs_lift) chk(dalloc(&d_lift, m->n, s));
  if (e == hipSuccess) {
    chk(hipMemcpyAsync(d_conn, conn32.data(), ne * nne * sizeof(int), hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_xyz, xyz, nn * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_a, mat_a, ne * sizeof(double), hipMemcpyHostToDevice, s));
    if (d_b) chk(hipMemcpyAsync(d_b, mat_b, ne * sizeof(double), hipMemcpyHostToDevice, s));
    if (d_t) chk(hipMemcpyAsync(d_t, thick, ne * sizeof(double), hipMemcpyHostToDevice, s));
    if (d_pv) chk(hipMemcpyAsync(d_pv, pv, npv * sizeof(double), hipMemcpyHostToDevice, s));
    if (d_lift) chk(hipMemsetAsync(d_lift, 0, m->n * sizeof(double), s));
    if (rhs_lift) {
      chk(hipMemcpyAsync(d_pidx, pidx.data(), ne * sizeof(int), hipMemcpyHostToDevice, s));
      chk(hipMemcpyAsync(d_lstart, lstart.data(), (m->n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
      if (!lidx.empty())
        chk(hipMemcpyAsync(d_lidx, lidx.data(), lidx.size() * sizeof(int), hipMemcpyHostToDevice, s));
    }
  }
  // element matrices (and lift values) of every element, then the sums per entry
  // (and per unknown) in the reference's order
  if (e == hipSuccess)
    chk(assemble_elements(kind, ne, d_conn, d_cn, d_xyz, d_a, d_b, d_t, d_kbuf, d_pidx, d_pv,
                          d_fbuf, s));
  if (e == hipSuccess)
    chk(asm_gather(m->negm, nd, true, wsh, m->asm_cstart, m->asm_cidx, d_kbuf, m->a_ref, s));
  if (e == hipSuccess && rhs_lift) chk(lift_gather(m->n, d_lstart, d_lidx, d_fbuf, d_lift, s));
  if (e == hipSuccess) {
    int r2 = values_changed(m);
    if (r2) e = hipErrorUnknown;
  }
  if (e == hipSuccess && rhs_lift)
    chk(hipMemcpyAsync(rhs_lift, d_lift, m->n * sizeof(double), hipMemcpyDeviceToHost, s));
  chk(hipStreamSynchronize(s));
  dfree(d_conn, s); dfree(d_xyz, s); dfree(d_a, s); dfree(d_b, s); dfree(d_t, s); dfree(d_pv, s); dfree(d_lift, s);
  dfree(d_cn, s); dfree(d_kbuf, s); dfree(d_fbuf, s); dfree(d_pidx, s); dfree(d_lstart, s);
  dfree(d_lidx, s);
  SSF_HIP(e);
  return SSF_OK;
}

extern "C" int ssf_matrix_localize_batch(ssf_matrix *m, int64_t ne, int64_t nc, const int64_t *cn,
                                         const double *lm)
{
  SSF_ENTER();
  SSF_CHECK(m && cn && lm, SSF_EINVAL, "ssf_matrix_localize_batch: NULL argument");
  SSF_CHECK(ne >= 0 && nc > 0 && nc <= 64, SSF_EINVAL, "ssf_matrix_localize_batch: bad sizes");
  if (ne == 0) return SSF_OK;
  SSF_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  int *d_cn = nullptr;
  int rc = prepare_assembly(m, ne, (int)nc, cn, &d_cn);
  if (rc) return rc;
  double *d_lm = nullptr;
  SSF_HIP(dalloc(&d_lm, ne * nc * nc, s));
  SSF_HIP(hipMemcpyAsync(d_lm, lm, ne * nc * nc * sizeof(double), hipMemcpyHostToDevice, s));
  // every entry adds its contributions of the element matrices in element order
  SSF_HIP(asm_gather(m->negm, (int)nc, false, 0, m->asm_cstart, m->asm_cidx, d_lm, m->a_ref, s));
  rc = values_changed(m);
  SSF_HIP(hipStreamSynchronize(s));
  dfree(d_lm, s);
  dfree(d_cn, s);
  return rc;
}

// --------------------------------------------------------------- workspace

static int ensure_ws(ssf_matrix *m, bool mixed)
{
  Workspace &w = m->ws;
  const int64_t n = std::max<int64_t>(1, m->n);
  const int nb = std::max(m->tune64.blocks, m->tune32.blocks);
  if (w.n != n || w.nb < nb) {
    free_workspace(w, m->ctx->stream);
    w.n = n;
    w.nb = std::max(nb, 8192);
    SSF_HIP(dalloc(&w.x, n, m->ctx->stream)); SSF_HIP(dalloc(&w.r, n, m->ctx->stream)); SSF_HIP(dalloc(&w.d, n, m->ctx->stream));
    SSF_HIP(dalloc(&w.p, n, m->ctx->stream)); SSF_HIP(dalloc(&w.h, n, m->ctx->stream)); SSF_HIP(dalloc(&w.y, n, m->ctx->stream));
    SSF_HIP(dalloc(&w.part, 4 * (size_t)w.nb, m->ctx->stream));
    SSF_HIP(dalloc(&w.st, 1, m->ctx->stream));
    SSF_HIP(hipHostMalloc((void **)&w.st_host, 2 * sizeof(PcgState<double>), hipHostMallocDefault));
  }
  if (mixed && !w.x32) {
    SSF_HIP(dalloc(&w.x32, n, m->ctx->stream)); SSF_HIP(dalloc(&w.r32, n, m->ctx->stream)); SSF_HIP(dalloc(&w.d32, n, m->ctx->stream));
    SSF_HIP(dalloc(&w.p32, n, m->ctx->stream)); SSF_HIP(dalloc(&w.h32, n, m->ctx->stream)); SSF_HIP(dalloc(&w.y32, n, m->ctx->stream));
    SSF_HIP(dalloc(&w.part32, 4 * (size_t)w.nb, m->ctx->stream));
    SSF_HIP(dalloc(&w.probe, 1, m->ctx->stream));
    w.mix = &w.probe->mix;
    w.st32 = &w.probe->st;
    SSF_HIP(hipHostMalloc((void **)&w.probe_host, 2 * sizeof(MixProbe), hipHostMallocDefault));
  }
  if (mixed && !m->val32_valid) {
    // a SELL-only matrix (device-built pattern) has no full CSR: its FP32 copy
    // and its norm come from the SELL values
    if (!m->sell_only && !m->val32) SSF_HIP(dalloc(&m->val32, m->nnz, m->ctx->stream));
    if (!m->isd) SSF_HIP(dalloc(&m->isd, m->n, m->ctx->stream));
    if (m->sell_v64 && !m->sell_v32) SSF_HIP(dalloc(&m->sell_v32, m->sell_tot, m->ctx->stream));
    SSF_HIP(mixed_operator(m->csr(), m->diag, m->sell_only ? nullptr : m->val, m->val32,
                           m->sell_v64, m->sell_v32, m->isd, m->ctx->stream));
    if (int rc = half_build(m, 1)) return rc;
    if (!m->norm_a) SSF_HIP(dalloc(&m->norm_a, 1, m->ctx->stream));
    if (m->sell_only)
      SSF_HIP(sell_row_absmax((int)m->n, m->sell_off, m->sell_v64, w.part, 1024, m->norm_a,
                              m->ctx->stream));
    else
      SSF_HIP(row_absmax((int)m->n, m->rowptr, m->val, w.part, 1024, m->norm_a, m->ctx->stream));
    m->val32_valid = true;
  }
  return SSF_OK;
}

// ----------------------------------------------------------------- SpMV

extern "C" int ssf_gmxv_dev(ssf_matrix *m, const double *a_dev, double *b_dev)
{
  SSF_ENTER();
  SSF_CHECK(m && a_dev && b_dev, SSF_EINVAL, "ssf_gmxv_dev: NULL argument");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_gmxv_dev: matrix values were never set");
  SSF_HIP(hipSetDevice(m->ctx->device));
  if (m->n == 0) return SSF_OK;
  SSF_HIP(spmv<double>(m->tune64, 0, m->csr(), m->val, a_dev, b_dev, nullptr,
                       nullptr, nullptr, nullptr, m->ctx->stream));
  return SSF_OK;
}

extern "C" int ssf_gmxv(ssf_matrix *m, const double *a, double *b)
{
  SSF_ENTER();
  SSF_CHECK(m && a && b, SSF_EINVAL, "ssf_gmxv: NULL argument");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_gmxv: matrix values were never set");
  SSF_HIP(hipSetDevice(m->ctx->device));
  int rc = ensure_ws(m, false);
  if (rc) return rc;
  hipStream_t s = m->ctx->stream;
  SSF_HIP(hipMemcpyAsync(m->ws.d, a, m->n * sizeof(double), hipMemcpyHostToDevice, s));
  rc = ssf_gmxv_dev(m, m->ws.d, m->ws.p);
  if (rc) return rc;
  SSF_HIP(hipMemcpyAsync(b, m->ws.p, m->n * sizeof(double), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  return SSF_OK;
}

// ------------------------------------------------------------------ solves

static void prof_add(ssf_ctx *c, int which, int parity)
{
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->prof_ev[parity][0], c->prof_ev[parity][1]) == hipSuccess) {
    c->prof_us[which] += 1000.0 * ms;
    c->prof_n[which] += 1;
  }
}

extern "C" int ssf_ctx_set_profiling(ssf_ctx *ctx, int on)
{
  SSF_ENTER();
  SSF_CHECK(ctx, SSF_EINVAL, "ssf_ctx_set_profiling: NULL context");
  SSF_HIP(hipSetDevice(ctx->device));
  if (on && !ctx->prof_ev[0][0])
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) SSF_HIP(hipEventCreate(&ctx->prof_ev[a][b]));
  ctx->prof = on != 0;
  ctx->prof_us[0] = ctx->prof_us[1] = 0.0;
  ctx->prof_n[0] = ctx->prof_n[1] = 0;
  return SSF_OK;
}

extern "C" int ssf_ctx_kernel_time(ssf_ctx *ctx, int which, double *avg_us, int64_t *count)
{
  SSF_ENTER();
  SSF_CHECK(ctx && avg_us && (which == 0 || which == 1), SSF_EINVAL, "ssf_ctx_kernel_time: bad argument");
  *avg_us = ctx->prof_n[which] ? ctx->prof_us[which] / ctx->prof_n[which] : 0.0;
  if (count) *count = ctx->prof_n[which];
  return SSF_OK;
}

namespace {

constexpr int CHUNK = 16;  // iterations enqueued between two host checks

template <typename T>
int run_pcg(ssf_matrix *m, const SpmvTune &t, int mode, bool jac, const T *val, const T *diag,
            T *x, const T *y, T *r, T *d, T *p, T *h, T *part, PcgState<T> *st,
            PcgState<T> *st_host, bool havex, int64_t ni, T res, T zero, int64_t *launched)
{
  hipStream_t s = m->ctx->stream;
  const int n = (int)m->n;
  SSF_HIP(pcg_init<T>(t, jac, havex, m->csr(), val, diag, x, y, r, h, d, part, st,
                      zero, s));
  hipEvent_t ev[2];
  SSF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SSF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int64_t it = 0;
  int k = 0;
  bool pending = false;
  int rc = SSF_OK;
  bool sample_valid[2] = {true, false};
  const int which = sizeof(T) == 8 ? 0 : 1;
  // chunk k is enqueued before the host waits for chunk k-1's state copy
  while (true) {
    bool launched_now = false;
    if (it < ni) {
      const int64_t c = std::min<int64_t>(CHUNK, ni - it);
      // the sampled SpMV is the chunk's second: the first follows the state copy
      // and starts on cold caches
      const int64_t qs = c > 1 ? 1 : 0;
      for (int64_t q = 0; q < c; q++) {
        const bool smp = m->ctx->prof && q == qs;
        hipError_t e = pcg_iteration<T>(t, mode, jac, m->csr(), val, diag, x, r, d, p,
                                        h, part, st, res, zero, s,
                                        smp ? m->ctx->prof_ev[k & 1][0] : nullptr,
                                        smp ? m->ctx->prof_ev[k & 1][1] : nullptr);
        if (e != hipSuccess) {
          set_error(std::string("pcg_iteration: ") + hipGetErrorString(e));
          rc = SSF_EHIP;
          break;
        }
      }
      if (rc) break;
      it += c;
      (void)hipMemcpyAsync(&st_host[k & 1], st, sizeof(PcgState<T>), hipMemcpyDeviceToHost, s);
      (void)hipEventRecord(ev[k & 1], s);
      launched_now = true;
    }
    if (pending) {
      (void)hipEventSynchronize(ev[(k - 1) & 1]);
      if (m->ctx->prof && sample_valid[(k - 1) & 1]) prof_add(m->ctx, which, (k - 1) & 1);
      sample_valid[k & 1] = !st_host[(k - 1) & 1].done;
      if (st_host[(k - 1) & 1].done) break;
    }
    if (!launched_now) break;
    pending = true;
    k++;
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  SSF_HIP(hipMemcpy(&st_host[0], st, sizeof(PcgState<T>), hipMemcpyDeviceToHost));
  *launched = it;
  return SSF_OK;
}

int true_residual(ssf_matrix *m, const double *x, const double *y, double *rel)
{
  hipStream_t s = m->ctx->stream;
  Workspace &w = m->ws;
  SSF_HIP(spmv<double>(m->tune64, 2, m->csr(), m->val, x, w.p, y, w.part,
                       nullptr, nullptr, s));
  SSF_HIP(reduce_to_host_slot(w.part, m->tune64.blocks, w.part + 2 * w.nb, s));
  SSF_HIP(dot_partial(m->n, y, y, w.part, 256, s));
  SSF_HIP(reduce_to_host_slot(w.part, 256, w.part + 2 * w.nb + 1, s));
  double v[2];
  SSF_HIP(hipMemcpyAsync(v, w.part + 2 * w.nb, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  *rel = v[1] > 0 ? std::sqrt(v[0] / v[1]) : 0.0;
  return SSF_OK;
}

}  // namespace

static int solve_fp64(ssf_matrix *m, int mode, bool jac, const ssf_solve_params *p,
                      const double *y, double *x, ssf_solve_info *info)
{
  hipStream_t s = m->ctx->stream;
  Workspace &w = m->ws;
  if (!p->iv) SSF_HIP(hipMemsetAsync(x, 0, m->n * sizeof(double), s));
  int64_t launched = 0;
  int rc = run_pcg<double>(m, m->tune64, mode, jac, m->val, m->diag, x, y,
                           w.r, w.d, w.p, w.h, w.part, w.st, w.st_host, p->iv != 0, p->ni, p->res,
                           p->zero, &launched);
  if (rc) return rc;
  const PcgState<double> &st = w.st_host[0];
  info->status = st.status;
  info->outer = 0;
  if (st.status == 3) {
    // zero right-hand side: x = 0 (symcomprow::cg, GEFEL/scr.cpp:821-830); the
    // reference's cg_prec would continue into 0/0 (GEFEL/scr.cpp:1117-1120)
    if (mode != MODE_CR_CG_PREC) SSF_HIP(hipMemsetAsync(x, 0, m->n * sizeof(double), s));
    info->ani = 0;
    info->ares = st.ares;
    info->iterations = 0;
  } else if (st.status == 1 || st.status == 2) {
    info->ani = st.ani;
    info->ares = st.ares;
    info->iterations = st.ani + 1;
  } else {
    // ni spent: cg_prec leaves ani/ares as they were; cg variants set ani = ni,
    // ares = last r.h (GEFEL/scr.cpp:876, GEFEL/cr.cpp:1227)
    if (mode != MODE_SCR_CG_PREC) {
      info->ani = p->ni;
      info->ares = st.nom[p->ni & 1];
    }
    info->iterations = launched;
  }
  return true_residual(m, x, y, &info->rel_residual);
}

// Mixed-precision solve: FP32 Jacobi-PCG on float(A) with FP32-accumulated dots,
// refined in FP64 by reliable updates every `interval` iterations (and whenever
// the FP32 recurrence meets the test): x64 += x32, r = b - A x64 in FP64, the
// FP32 residual is replaced by the FP64 one while the search direction is kept.
// Stops when ||b - A x||^2 < res ||b||^2 in FP64 (the squared test of
// GEFEL/scr.cpp:1188 applied to the true residual).
static int solve_mixed(ssf_matrix *m, const ssf_solve_params *p, const double *y, double *x,
                       ssf_solve_info *info)
{
  hipStream_t s = m->ctx->stream;
  Workspace &w = m->ws;
  const CsrDev A = m->csr();
  if (!p->iv) SSF_HIP(hipMemsetAsync(x, 0, m->n * sizeof(double), s));
  // (128 measured best on heat 128^3 and the 64 x 64 x 256 elasticity column, from
  // 16 to 512; at 512 the elasticity FP32 recurrence needs five times the iterations)
  const int64_t interval = p->inner_ni > 0 ? p->inner_ni : 128;
  // ||b||^2 and the initial FP64 residual
  double *dslot = w.part + 3 * w.nb;
  SSF_HIP(dot_partial(m->n, y, y, w.part, 256, s));
  SSF_HIP(reduce_to_host_slot(w.part, 256, dslot + 1, s));
  SSF_HIP(spmv<double>(m->tune64, 2, A, m->val, x, w.r, y, w.part, nullptr, nullptr, s));
  SSF_HIP(reduce_to_host_slot(w.part, m->tune64.blocks, dslot, s));
  double slots[2];
  SSF_HIP(hipMemcpyAsync(slots, dslot, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipStreamSynchronize(s));
  info->outer = 0;
  info->iterations = 0;
  if (slots[1] < p->zero) {  // zero right-hand side: x = 0 (GEFEL/scr.cpp:821-830)
    SSF_HIP(hipMemsetAsync(x, 0, m->n * sizeof(double), s));
    SSF_HIP(hipStreamSynchronize(s));
    info->status = 3;
    info->ani = 0;
    info->ares = slots[1];
    info->rel_residual = 0.0;
    return SSF_OK;
  }
  // FP32 recurrence on the residual scaled by 1/||r0|| (the device keeps the
  // scale from here on), x32 = 0
  const double scale = slots[0] > 0.0 ? 1.0 / std::sqrt(slots[0]) : 1.0;
  SSF_HIP(mixed_rhs32((int)m->n, w.r, m->isd, w.y32, scale, s));
  SSF_HIP(hipMemsetAsync(w.x32, 0, m->n * sizeof(float), s));
  SSF_HIP(pcg_init<float>(m->tune32, false, false, A, m->val32, nullptr, w.x32, w.y32, w.r32,
                          nullptr, w.d32, w.part32, w.st32, 1e-30f, s));
  const double delta = (p->inner_res > 0.0 && p->inner_res < 1.0) ? p->inner_res : 1e-4;
  SSF_HIP(mixed_init(w.mix, w.st32, dslot, m->norm_a, p->res, delta, s));
  // The FP32 iterations go out in units of at most MIX_UNIT iterations, the
  // host one unit ahead of what it has read back: after each unit both states
  // are copied to the host and an event is recorded, and unit u + 1 is enqueued
  // once unit u - 1 is known. A unit may end with the FP64 refresh, which closes
  // the chunk of iterations since the previous refresh. A chunk closes
  //  - after `interval` iterations, or when ni is spent;
  //  - as soon as a unit shows the FP32 recurrence paused (its residual met the
  //    refresh threshold): the iterations enqueued after the pause exit early
  //    (a few us each) and change nothing, so the refresh sees the iterate of
  //    the pause.
  // Units are shortened to end one iteration (and 5 %) after where the FP32
  // residual's mean rate of decrease since the last refresh puts the pause, and
  // the unit after a refresh that may end the solve (its chunk aimed at the
  // final residual) is short, so that few iterations are enqueued past a pause.
  // The refresh itself waits for the pause to be seen: placed on the prediction
  // it came early on heat 128^3 (the first iterations of CG fall fastest) and
  // cost another refresh and 21 iterations. Every decision reads device states
  // only, so a solve repeats bit for bit.
  constexpr int64_t MIX_UNIT = 16;
  struct Unit {
    int64_t len;
    bool refresh, sample;
  } unit[2];
  hipEvent_t ev[2];
  SSF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SSF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  MixProbe *ph = w.probe_host;
  // unit 0: the state after mixed_init, taken as a refresh
  SSF_HIP(hipMemcpyAsync(&ph[0], w.probe, sizeof(MixProbe), hipMemcpyDeviceToHost, s));
  SSF_HIP(hipEventRecord(ev[0], s));
  unit[0] = {0, true, false};
  int64_t u = 1, it = 0, in_chunk = 0, next_len = std::min<int64_t>(MIX_UNIT, interval);
  bool next_close = false;
  bool sample_valid[2] = {false, false};
  double rate = 0.0, base_lrr = 0.0;
  int64_t base_iter = 0;
  int rc = SSF_OK;
  while (true) {
    bool launched_now = false;
    if (it < p->ni) {
      const int64_t c = std::min<int64_t>(std::min<int64_t>(next_len, interval - in_chunk), p->ni - it);
      const bool close = next_close || in_chunk + c >= interval || it + c >= p->ni;
      // sampled: the unit's second FP32 SpMV (the first after a refresh starts on
      // cold caches)
      const int64_t qs = c > 1 ? 1 : 0;
      const int slot = (int)(u & 1);
      for (int64_t q = 0; q < c && rc == SSF_OK; q++)
        if (pcg_iteration<float>(m->tune32, MODE_MIXED_INNER, false, A, m->val32, nullptr, w.x32,
                                 w.r32, w.d32, w.p32, nullptr, w.part32, w.st32, 1.0f, 1e-30f, s,
                                 (m->ctx->prof && q == qs) ? m->ctx->prof_ev[slot][0] : nullptr,
                                 (m->ctx->prof && q == qs) ? m->ctx->prof_ev[slot][1] : nullptr) !=
            hipSuccess)
          rc = SSF_EHIP;
      if (rc == SSF_OK && close &&
          mixed_update(m->tune64, m->tune32, A, m->val, m->isd, y, x, w.r, w.part,
                       w.part + w.nb, w.x32, w.r32, w.d32, w.part32, w.st32, w.mix, p->res,
                       s) != hipSuccess)
        rc = SSF_EHIP;
      if (rc) {
        set_error("solve_mixed: kernel launch failed");
        break;
      }
      it += c;
      in_chunk = close ? 0 : in_chunk + c;
      (void)hipMemcpyAsync(&ph[slot], w.probe, sizeof(MixProbe), hipMemcpyDeviceToHost, s);
      (void)hipEventRecord(ev[slot], s);
      unit[slot] = {c, close, m->ctx->prof && c > 0};
      launched_now = true;
    }
    // unit u - 1 is known
    const int ps = (int)((u - 1) & 1);
    (void)hipEventSynchronize(ev[ps]);
    if (sample_valid[ps]) prof_add(m->ctx, 1, ps);
    sample_valid[ps] = false;
    const MixProbe &pv = ph[ps];
    if (pv.mix.final || !launched_now) break;
    const Unit &cur = unit[u & 1];
    // unit u's sample is a real iteration unless the recurrence is paused
    sample_valid[u & 1] = cur.sample && !pv.st.done;
    // FP32 residual and its rate of decrease since the last refresh
    const int64_t iter_v = pv.st.iter + 1;
    const double rr = (double)pv.st.nom[(pv.st.iter + 1) & 1];
    if (unit[ps].refresh) {
      base_iter = iter_v;
      base_lrr = rr > 0.0 ? std::log(rr) : 0.0;
    } else if (iter_v > base_iter && rr > 0.0 && std::log(rr) < base_lrr) {
      rate = (base_lrr - std::log(rr)) / (double)(iter_v - base_iter);
    }
    // the current chunk aims at the final residual: its threshold is
    // res ||b||^2, not delta^2 of the residual at its refresh
    const bool aims_final = p->res * pv.mix.nory64 >= pv.mix.delta2 * pv.mix.rr64;
    next_close = false;
    next_len = MIX_UNIT;
    if (cur.refresh) {
      if (aims_final) next_len = 4;
    } else if (pv.st.done) {
      next_close = true;  // paused during unit u - 1: unit u ran no iteration
      next_len = 0;
    } else if (rate > 0.0 && rr > 0.0 && pv.st.nory > 0.0f) {
      // iterations after unit u until the residual meets the threshold
      const double left = std::log(rr / (double)pv.st.nory) / rate - (double)cur.len;
      if (left < (double)MIX_UNIT)
        next_len = std::max<int64_t>(2, (int64_t)std::ceil(left + 0.05 * std::max(left, 0.0)) + 1);
    }
    u++;
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  MixProbe pr;
  SSF_HIP(hipMemcpy(&pr, w.probe, sizeof pr, hipMemcpyDeviceToHost));
  const MixState &ms = pr.mix;
  const PcgState<float> &st = pr.st;
  info->status = ms.final ? ms.status : 0;
  info->outer = ms.updates;
  info->iterations = st.iter + 1;
  if (info->status != 0) {
    info->ani = info->iterations;
    info->ares = ms.rr64 / ms.nory64;
  }
  info->rel_residual = std::sqrt(ms.rr64 / ms.nory64);
  return SSF_OK;
}

extern "C" int ssf_solve_dev(ssf_matrix *m, const ssf_solve_params *p, const double *rhs_dev,
                             double *lhs_dev, ssf_solve_info *info)
{
  SSF_ENTER();
  SSF_CHECK(m && p && info && rhs_dev && lhs_dev, SSF_EINVAL, "ssf_solve_dev: NULL argument");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_solve_dev: matrix values were never set");
  SSF_CHECK(p->linsolver == SSF_SOLVER_CG || p->linsolver == SSF_SOLVER_BICG, SSF_EUNSUP,
            "ssf_solve_dev: the cg (20) and bicg (30) solvers are on this path");
  SSF_CHECK(p->ni >= 0, SSF_EINVAL, "ssf_solve_dev: negative ni");
  if (p->linsolver == SSF_SOLVER_BICG) {
    // gmatrix::solve_system, GEFEL/gmatrix.cpp:1273-1289: comprow::bicg, no
    // preconditioner; symcomprow has no bicg in the reference
    SSF_CHECK(m->storage == SSF_STORAGE_COMPRESSED_ROWS, SSF_EUNSUP,
              "ssf_solve_dev: bicg needs compressed_rows (10) storage");
    SSF_CHECK(p->precision == SSF_PRECISION_FP64, SSF_EUNSUP,
              "ssf_solve_dev: bicg runs in FP64 only");
    SSF_HIP(hipSetDevice(m->ctx->device));
    int rc = ensure_ws(m, false);
    if (rc) return rc;
    if (m->n == 0) {
      info->status = 1;
      info->iterations = 0;
      return SSF_OK;
    }
    info->rel_residual = 0.0;
    rc = solve_bicg(m, p, rhs_dev, lhs_dev, info);
    if (rc) return rc;
    return true_residual(m, lhs_dev, rhs_dev, &info->rel_residual);
  }
  if (m->storage == SSF_STORAGE_SYMM_COMP_ROWS && p->prec == SSF_PREC_ILDL) {
    // gmatrix::solve_system incomdec branch (GEFEL/gmatrix.cpp:1234-1240)
    SSF_CHECK(p->precision == SSF_PRECISION_FP64, SSF_EUNSUP,
              "ssf_solve_dev: the incomplete LDL^T preconditioner runs in FP64 only");
    SSF_HIP(hipSetDevice(m->ctx->device));
    int rc = ensure_ws(m, false);
    if (rc) return rc;
    if (m->n == 0) {
      info->status = 1;
      info->iterations = 0;
      return SSF_OK;
    }
    info->rel_residual = 0.0;
    rc = solve_ildl(m, p, rhs_dev, lhs_dev, info);
    if (rc) return rc;
    return true_residual(m, lhs_dev, rhs_dev, &info->rel_residual);
  }
  int mode;
  bool jac;
  if (m->storage == SSF_STORAGE_SYMM_COMP_ROWS) {
    if (p->prec == SSF_PREC_DIAG) { mode = MODE_SCR_CG_PREC; jac = true; }
    else if (p->prec == SSF_PREC_NOPRECOND) { mode = MODE_SCR_CG; jac = false; }
    else { set_error("ssf_solve_dev: symm_comp_rows supports noprecond (0), diagprec (1) and incomdec (10) here"); return SSF_EUNSUP; }
  } else {
    if (p->prec == SSF_PREC_NOPRECOND) { mode = MODE_CR_CG_PREC; jac = false; }
    else if (p->prec == SSF_PREC_DIAG && p->precision == SSF_PRECISION_MIXED) { mode = MODE_SCR_CG_PREC; jac = true; }
    else {
      set_error("ssf_solve_dev: compressed_rows + diagprec is not implemented by the reference "
                "(precond::preconditioning, GEFEL/precond.cpp:161-183)");
      return SSF_EUNSUP;
    }
  }
  SSF_HIP(hipSetDevice(m->ctx->device));
  const bool mixed = p->precision == SSF_PRECISION_MIXED;
  SSF_CHECK(!mixed || jac, SSF_EUNSUP, "ssf_solve_dev: the mixed-precision solve is Jacobi-preconditioned (diagprec)");
  int rc = ensure_ws(m, mixed);
  if (rc) return rc;
  if (m->n == 0) {
    info->status = 1;
    info->iterations = 0;
    return SSF_OK;
  }
  info->rel_residual = 0.0;
  return mixed ? solve_mixed(m, p, rhs_dev, lhs_dev, info)
               : solve_fp64(m, mode, jac, p, rhs_dev, lhs_dev, info);
}

extern "C" int ssf_matrix_ildl_factor(ssf_matrix *m, double *incdec)
{
  SSF_ENTER();
  SSF_CHECK(m && incdec, SSF_EINVAL, "ssf_matrix_ildl_factor: NULL argument");
  SSF_CHECK(m->storage == SSF_STORAGE_SYMM_COMP_ROWS, SSF_EUNSUP,
            "ssf_matrix_ildl_factor: symm_comp_rows storage only");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_matrix_ildl_factor: matrix values were never set");
  SSF_HIP(hipSetDevice(m->ctx->device));
  if (m->n == 0) return SSF_OK;
  return ildl_factor_copy(m, incdec);
}

extern "C" int ssf_ildl_bench(ssf_matrix *m, int reps, int graph, int64_t *nfwd, int64_t *nbwd,
                              double *usec)
{
  SSF_ENTER();
  SSF_CHECK(m && nfwd && nbwd && usec && reps >= 0, SSF_EINVAL, "ssf_ildl_bench: bad argument");
  SSF_CHECK(m->storage == SSF_STORAGE_SYMM_COMP_ROWS, SSF_EUNSUP,
            "ssf_ildl_bench: symm_comp_rows storage only");
  SSF_CHECK(m->values_set && m->n > 0, SSF_ESTATE, "ssf_ildl_bench: matrix values were never set");
  SSF_HIP(hipSetDevice(m->ctx->device));
  return ildl_bench(m, reps, graph, nfwd, nbwd, usec);
}

extern "C" int ssf_ildl_apply_dev(ssf_matrix *m, const double *y_dev, double *x_dev)
{
  SSF_ENTER();
  SSF_CHECK(m && y_dev && x_dev, SSF_EINVAL, "ssf_ildl_apply_dev: NULL argument");
  SSF_CHECK(m->storage == SSF_STORAGE_SYMM_COMP_ROWS, SSF_EUNSUP,
            "ssf_ildl_apply_dev: symm_comp_rows storage only");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_ildl_apply_dev: matrix values were never set");
  SSF_CHECK(y_dev != x_dev, SSF_EINVAL, "ssf_ildl_apply_dev: x and y must not alias");
  SSF_HIP(hipSetDevice(m->ctx->device));
  if (m->n == 0) return SSF_OK;
  return ildl_apply_dev(m, y_dev, x_dev);
}

extern "C" int ssf_solve(ssf_matrix *m, const ssf_solve_params *p, const double *rhs, double *lhs,
                         ssf_solve_info *info)
{
  SSF_ENTER();
  SSF_CHECK(m && p && rhs && lhs && info, SSF_EINVAL, "ssf_solve: NULL argument");
  SSF_HIP(hipSetDevice(m->ctx->device));
  int rc = ensure_ws(m, p->precision == SSF_PRECISION_MIXED);
  if (rc) return rc;
  hipStream_t s = m->ctx->stream;
  double *y = nullptr, *x = nullptr;
  SSF_HIP(dalloc(&y, m->n, s));
  SSF_HIP(dalloc(&x, m->n, s));
  SSF_HIP(hipMemcpyAsync(y, rhs, m->n * sizeof(double), hipMemcpyHostToDevice, s));
  if (p->iv) SSF_HIP(hipMemcpyAsync(x, lhs, m->n * sizeof(double), hipMemcpyHostToDevice, s));
  rc = ssf_solve_dev(m, p, y, x, info);
  if (rc == SSF_OK) {
    SSF_HIP(hipMemcpyAsync(lhs, x, m->n * sizeof(double), hipMemcpyDeviceToHost, s));
    SSF_HIP(hipStreamSynchronize(s));
  }
  dfree(y, s);
  dfree(x, s);
  return rc;
}

// ------------------------------------------------------------ diagnostics

static int time_spmv(ssf_matrix *m, int precision, const SpmvTune &t, int reps, double *usec)
{
  hipStream_t s = m->ctx->stream;
  Workspace &w = m->ws;
  hipEvent_t a, b;
  SSF_HIP(hipEventCreate(&a));
  SSF_HIP(hipEventCreate(&b));
  auto launch = [&]() -> hipError_t {
    if (precision == 0)
      return spmv<double>(t, 0, m->csr(), m->val, w.d, w.p, nullptr, nullptr,
                          nullptr, nullptr, s);
    return spmv<float>(t, 0, m->csr(), m->val32, w.d32, w.p32, nullptr,
                       nullptr, nullptr, nullptr, s);
  };
  SSF_HIP(launch());
  SSF_HIP(hipEventRecord(a, s));
  for (int i = 0; i < reps; i++) SSF_HIP(launch());
  SSF_HIP(hipEventRecord(b, s));
  SSF_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  SSF_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *usec = 1000.0 * ms / reps;
  return SSF_OK;
}

// Mean time of one iteration of the recurrence a solve runs with tiling t: the
// FP64 drop-in's cg_prec step (product with d.p, update, direction) or the mixed
// solve's FP32 step (on the half form the product with the step's dots and the
// fused update), on the workspace vectors with a state that never stops. The
// element-wise kernels re-reduce the product's per-block partials, so the block
// count weighs on both.
static int time_iteration(ssf_matrix *m, int precision, const SpmvTune &t, int reps, double *usec)
{
  hipStream_t s = m->ctx->stream;
  Workspace &w = m->ws;
  const CsrDev A = m->csr();
  hipEvent_t a, b;
  SSF_HIP(hipEventCreate(&a));
  SSF_HIP(hipEventCreate(&b));
  if (precision == 0) {
    PcgState<double> st{};
    st.nory = 1.0;
    st.nom[0] = st.nom[1] = 1.0;
    SSF_HIP(fill(m->n, 1.0, w.r, s));
    SSF_HIP(hipMemcpyAsync(w.st, &st, sizeof st, hipMemcpyHostToDevice, s));
  } else {
    PcgState<float> st{};
    st.nory = -1.0f;  // ||r||^2 < nory never holds
    st.nom[0] = st.nom[1] = 1.0f;
    SSF_HIP(to_f32(m->n, w.d, w.r32, 1.0, s));
    SSF_HIP(hipMemcpyAsync(w.st32, &st, sizeof st, hipMemcpyHostToDevice, s));
  }
  auto launch = [&]() -> hipError_t {
    if (precision == 0)  // res = 0 and zero = 0: the tests never stop the recurrence
      return pcg_iteration<double>(t, MODE_SCR_CG_PREC, true, A, m->val, m->diag, w.x, w.r, w.d,
                                   w.p, w.h, w.part, w.st, 0.0, 0.0, s);
    return pcg_iteration<float>(t, MODE_MIXED_INNER, false, A, m->val32, nullptr, w.x32, w.r32,
                                w.d32, w.p32, nullptr, w.part32, w.st32, 0.0f, 0.0f, s);
  };
  SSF_HIP(launch());
  SSF_HIP(hipEventRecord(a, s));
  for (int i = 0; i < reps; i++) SSF_HIP(launch());
  SSF_HIP(hipEventRecord(b, s));
  SSF_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  SSF_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *usec = 1000.0 * ms / reps;
  return SSF_OK;
}

extern "C" int ssf_spmv_bench(ssf_matrix *m, int precision, int reps, double *usec)
{
  SSF_ENTER();
  SSF_CHECK(m && usec && reps > 0, SSF_EINVAL, "ssf_spmv_bench: bad argument");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_spmv_bench: matrix values were never set");
  SSF_HIP(hipSetDevice(m->ctx->device));
  int rc = ensure_ws(m, precision != 0);
  if (rc) return rc;
  SSF_HIP(fill(m->n, 1.0, m->ws.d, m->ctx->stream));
  if (precision) SSF_HIP(to_f32(m->n, m->ws.d, m->ws.d32, 1.0, m->ctx->stream));
  return time_spmv(m, precision, precision ? m->tune32 : m->tune64, reps, usec);
}

// SpMV autotuner: times every tiling candidate (row-vector kernel: lanes per row x
// blocks; stream kernel: lanes per row x blocks, with 16-bit column offsets for
// the FP32 operator when they fit; the SELL forms) on this matrix, each as one
// iteration of the recurrence its solve runs (time_iteration), and keeps the
// fastest. Solves use the chosen tiling from then on.
// tune_out = {format, lanes, blocks, idx16, map}.
extern "C" int ssf_matrix_autotune(ssf_matrix *m, int precision, int reps, int32_t *tune_out)
{
  SSF_ENTER();
  SSF_CHECK(m, SSF_EINVAL, "ssf_matrix_autotune: NULL matrix");
  SSF_CHECK(m->values_set, SSF_ESTATE, "ssf_matrix_autotune: matrix values were never set");
  SSF_HIP(hipSetDevice(m->ctx->device));
  if (reps <= 0) reps = 5;
  int rc = ensure_ws(m, precision != 0);
  if (rc) return rc;
  SSF_HIP(fill(m->n, 1.0, m->ws.d, m->ctx->stream));
  if (precision) SSF_HIP(to_f32(m->n, m->ws.d, m->ws.d32, 1.0, m->ctx->stream));
  std::vector<SpmvTune> cand;
  for (int l : {4, 8, 16, 32, 64})
    for (int b : {256, 512, 1024, 2048}) {
      if (!m->rowptr) continue;  // SELL-only matrix
      const int64_t groups = TPB / l;
      if ((int64_t)b * groups > std::max<int64_t>(m->n, groups * 8)) continue;
      SpmvTune t;
      t.format = 0;
      t.lanes = l;
      t.blocks = b;
      cand.push_back(t);
    }
  if (m->ntiles > 0)
    for (int l : {4, 8, 16})
      for (int b : {256, 512, 1024, 2048})
        for (int i16 = 0; i16 <= ((precision && m->col16) ? 1 : 0); i16++) {
          if (b > std::max(8, m->ntiles)) continue;
          SpmvTune t;
          t.format = 1;
          t.lanes = l;
          t.blocks = b;
          t.idx16 = i16;
          cand.push_back(t);
        }
  // SELL block counts: a solve's element-wise kernels re-reduce the SpMV's
  // per-block partials in every block, so more SpMV blocks than 2048 cost more
  // there than they save in the product
  if (m->nslice > 0)
    for (int b : {256, 512, 1024, 2048})
      for (int i16 = 0; i16 <= 3; i16++)
        // row patterns: value loads in flight per lane (2, 4, 8); stencil form:
        // 3 or 4 groups of slots per round, or the default (2)
        for (int dep : {2, 3, 4, 8})
          for (int map : {0, 1}) {
            if (b > std::max(8, (m->nslice + 3) / 4)) continue;
            if (i16 == 0 && !m->sell_col) continue;
            if (i16 == 1 && !m->sell_col16) continue;
            if (i16 == 2 && !m->sell_pid) continue;
            if (i16 == 3 && m->sten.empty()) continue;
            if (i16 == 2 && dep == 3) continue;
            // (rounds of 3 or 4 groups pay off for the long rows of the block
            // form; on the 27-slot heat rows they measured 1-2 us slower)
            if (i16 == 3 && (dep == 8 || (m->sten_b == 1 && dep != 2))) continue;
            if (i16 < 2 && dep != 2) continue;
            SpmvTune t;
            t.format = 2;
            t.lanes = dep;
            t.blocks = b;
            t.idx16 = i16;
            t.map = map;
            cand.push_back(t);
          }
  // symmetric half of the stencil form: blocks of 256 or 512 threads; up to 4096
  // blocks (the elasticity column's block products gain 2-4 % in the solves over
  // 2048, more than the element-wise kernels' longer re-reduction costs)
  if (m->half_nl > 0)
    for (int b : {256, 512, 1024, 2048, 4096})
      for (int tb : {1, 2})
        for (int map : {0, 1}) {
          if (b > std::max(8, (m->nslice + 3) / 4)) continue;
          SpmvTune t;
          t.format = 2;
          t.lanes = tb;
          t.blocks = b;
          t.idx16 = 4;
          t.map = map;
          cand.push_back(t);
        }
  SpmvTune best = precision ? m->tune32 : m->tune64;
  double best_t = 1e300;
  for (const SpmvTune &t : cand) {
    if (t.blocks > m->ws.nb) continue;
    double us = 0;
    rc = time_iteration(m, precision, t, reps, &us);
    if (rc) return rc;
    if (us < best_t) {
      best_t = us;
      best = t;
    }
  }
  if (precision) m->tune32 = best;
  else m->tune64 = best;
  if (tune_out) {
    tune_out[0] = best.format;
    tune_out[1] = best.lanes;
    tune_out[2] = best.blocks;
    tune_out[3] = best.idx16;
    tune_out[4] = best.map;
  }
  return SSF_OK;
}

extern "C" int ssf_matrix_set_tune(ssf_matrix *m, int precision, const int32_t *tune)
{
  SSF_ENTER();
  SSF_CHECK(m && tune, SSF_EINVAL, "ssf_matrix_set_tune: NULL argument");
  SpmvTune t;
  t.format = tune[0];
  t.lanes = tune[1];
  t.blocks = tune[2];
  t.idx16 = tune[3];
  t.map = tune[4];
  SSF_CHECK(t.map == 0 || t.map == 1, SSF_EINVAL, "ssf_matrix_set_tune: slice map must be 0 or 1");
  SSF_CHECK((t.format == 0 && m->rowptr) || (t.format == 1 && m->ntiles > 0) ||
                (t.format == 2 && m->nslice > 0),
            SSF_EINVAL,
            "ssf_matrix_set_tune: format must be 0 (row-vector), 1 (stream, needs a tiling) or 2 (SELL-64)");
  if (t.format == 0)
    SSF_CHECK(t.lanes == 4 || t.lanes == 8 || t.lanes == 16 || t.lanes == 32 || t.lanes == 64,
              SSF_EINVAL, "ssf_matrix_set_tune: row-vector lanes must be 4, 8, 16, 32 or 64");
  else if (t.format == 1)
    SSF_CHECK(t.lanes == 4 || t.lanes == 8 || t.lanes == 16, SSF_EINVAL,
              "ssf_matrix_set_tune: stream lanes must be 4, 8 or 16");
  else
    SSF_CHECK(t.lanes >= 0 && t.lanes <= 8, SSF_EINVAL,
              "ssf_matrix_set_tune: SELL depth must be in [0, 8] (value loads in flight per "
              "lane of the row-pattern form: 8 or more, 4 or more, else 2)");
  SSF_CHECK(t.blocks >= 1 && t.blocks <= 8192, SSF_EINVAL, "ssf_matrix_set_tune: blocks out of [1, 8192]");
  SSF_CHECK((t.idx16 == 0 && (t.format != 2 || m->sell_col)) ||
                (t.idx16 == 1 && precision && t.format == 1 && m->col16) ||
                (t.idx16 == 1 && t.format == 2 && m->sell_col16) ||
                (t.idx16 == 2 && t.format == 2 && m->sell_pid) ||
                (t.idx16 == 3 && t.format == 2 && !m->sten.empty()) ||
                (t.idx16 == 4 && t.format == 2 && m->half_nl > 0 && (t.lanes == 1 || t.lanes == 2)),
            SSF_EINVAL,
            "ssf_matrix_set_tune: index form 0 (int32 columns), 1 (16-bit columns) for the FP32 "
            "stream operator or the SELL operator of a matrix that has them, 2 (row patterns), "
            "3 (stencil form) or 4 (symmetric half of the stencil form, lanes 1 or 2: blocks of "
            "256 or 512 threads) for a SELL operator that has them");
  if (precision) m->tune32 = t;
  else m->tune64 = t;
  return SSF_OK;
}

extern "C" int ssf_copy_bench(ssf_ctx *ctx, int64_t bytes, int reps, double *usec)
{
  SSF_ENTER();
  SSF_CHECK(ctx && usec && reps > 0 && bytes >= 16, SSF_EINVAL, "ssf_copy_bench: bad argument");
  SSF_HIP(hipSetDevice(ctx->device));
  const int64_t n = bytes / 16 * 2;  // doubles, a multiple of 2
  hipStream_t s = ctx->stream;
  double *a = nullptr, *b = nullptr;
  SSF_HIP(dalloc(&a, n, s));
  hipError_t e = dalloc(&b, n, s);
  if (e != hipSuccess) {
    dfree(a, s);
    SSF_HIP(e);
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  chk(fill(n, 1.0, a, s));
  chk(hipEventCreate(&t0));
  chk(hipEventCreate(&t1));
  if (e == hipSuccess) {
    chk(copy_kernel(n, a, b, s));
    chk(hipEventRecord(t0, s));
    for (int i = 0; i < reps; i++) chk(copy_kernel(n, a, b, s));
    chk(hipEventRecord(t1, s));
    chk(hipEventSynchronize(t1));
    float ms = 0.f;
    chk(hipEventElapsedTime(&ms, t0, t1));
    *usec = 1000.0 * ms / reps;
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  dfree(a, s);
  dfree(b, s);
  SSF_HIP(e);
  return SSF_OK;
}

extern "C" int ssf_spmv_bytes(const ssf_matrix *m, int precision, double *canonical, double *actual)
{
  SSF_ENTER();
  SSF_CHECK(m, SSF_EINVAL, "ssf_spmv_bytes: NULL matrix");
  const double sv = precision ? 4.0 : 8.0;
  const double b = (double)m->nnz * (sv + 4.0) + 4.0 * (double)(m->n + 1) + 2.0 * sv * (double)m->n;
  const SpmvTune &t = precision ? m->tune32 : m->tune64;
  const double ib = ((t.format == 1 && t.idx16 && m->col16) ||
                     (t.format == 2 && t.idx16 && m->sell_col16)) ? 2.0 : 4.0;
  double a = (double)m->nnz * (sv + ib) + 4.0 * (double)(m->n + 1) + 2.0 * sv * (double)m->n;
  if (t.format == 1) a += 4.0 * (double)(m->ntiles + 1);
  if (t.format == 2 && m->nslice > 0) {  // padded slices, slice offsets instead of row pointers
    if (t.idx16 == 4 && m->half_nl > 0)  // half stencil form: the lower slots of every row once
      a = (double)half_stored(m) * (double)(m->n / m->half_b) * sv + 2.0 * sv * (double)m->n;
    else if (t.idx16 == 3 && !m->sten.empty())  // stencil form: the values alone
      a = (double)m->sell_tot * sv + 2.0 * sv * (double)m->n;
    else if (t.idx16 == 2 && m->sell_pid)  // a 2-byte pattern id per layout row instead of columns
      a = (double)m->sell_tot * sv + 2.0 * 64.0 * (double)m->nslice +
          4.0 * (double)(m->nslice + 1) + 2.0 * sv * (double)m->n;
    else
      a = (double)m->sell_tot * (sv + ib) + 4.0 * (double)(m->nslice + 1) + 2.0 * sv * (double)m->n;
  }
  if (canonical) *canonical = b;
  if (actual) *actual = a;
  return SSF_OK;
}

// ------------------------------------------------ matrix combinations and the
// nonstationary step

static int same_pattern(const ssf_matrix *a, const ssf_matrix *b)
{
  return a->ctx == b->ctx && a->storage == b->storage && a->n == b->n && a->negm == b->negm &&
         a->pattern_fp == b->pattern_fp;
}

extern "C" int ssf_matrix_scale(ssf_matrix *m, double c)
{
  SSF_ENTER();
  SSF_CHECK(m, SSF_EINVAL, "ssf_matrix_scale: NULL matrix");
  SSF_CHECK(m->a_ref, SSF_EUNSUP, "ssf_matrix_scale: the matrix keeps no reference-layout values");
  SSF_HIP(hipSetDevice(m->ctx->device));
  SSF_HIP(vals_scale(m->negm, m->a_ref, c, m->ctx->stream));
  int rc = values_changed(m);
  if (rc) return rc;
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  return SSF_OK;
}

extern "C" int ssf_matrix_add(ssf_matrix *m, double c, const ssf_matrix *other)
{
  SSF_ENTER();
  SSF_CHECK(m && other, SSF_EINVAL, "ssf_matrix_add: NULL matrix");
  SSF_CHECK(m->a_ref && other->a_ref, SSF_EUNSUP,
            "ssf_matrix_add: the matrices keep no reference-layout values");
  SSF_CHECK(same_pattern(m, other), SSF_EINVAL,
            "ssf_matrix_add: the matrices differ in context, storage or pattern");
  SSF_HIP(hipSetDevice(m->ctx->device));
  SSF_HIP(vals_add(m->negm, m->a_ref, c, other->a_ref, m->ctx->stream));
  int rc = values_changed(m);
  if (rc) return rc;
  SSF_HIP(hipStreamSynchronize(m->ctx->stream));
  return SSF_OK;
}

// A = K * (dt alpha) + 1.0 * C, unless A already holds it for these values of K
// and C and this dt * alpha; both time-stepping forms form it this way
// (npsolvert.cpp:432-433 and :1158-1159)
static int np_form_system(ssf_matrix *K, ssf_matrix *C, ssf_matrix *A, double alpha, double dt,
                          const char *who)
{
  const std::string w(who);
  SSF_CHECK(K != A && C != A, SSF_EINVAL, w + ": A must be a third matrix");
  SSF_CHECK(K->a_ref && C->a_ref && A->a_ref && K->values_set && C->values_set, SSF_ESTATE,
            w + ": K and C need their values");
  SSF_CHECK(same_pattern(K, C) && same_pattern(K, A), SSF_EINVAL,
            w + ": K, C and A differ in context, storage or pattern");
  SSF_HIP(hipSetDevice(K->ctx->device));
  const double sc = dt * alpha;
  if (A->np_k != K || A->np_c != C || A->np_kv != K->values_version ||
      A->np_cv != C->values_version || A->np_s != sc || !A->values_set) {
    SSF_HIP(vals_combine(A->negm, A->a_ref, K->a_ref, sc, C->a_ref, K->ctx->stream));
    int rc = values_changed(A);
    if (rc) return rc;
    A->np_k = K;
    A->np_c = C;
    A->np_kv = K->values_version;
    A->np_cv = C->values_version;
    A->np_s = sc;
  }
  return SSF_OK;
}

extern "C" int ssf_np_step_linear(ssf_matrix *K, ssf_matrix *C, ssf_matrix *A, double alpha,
                                  double dt, const ssf_solve_params *p, const double *rhs_dev,
                                  double *lhs_dev, double *tdlhs_dev, ssf_solve_info *info)
{
  SSF_ENTER();
  SSF_CHECK(K && C && A && p && rhs_dev && lhs_dev && tdlhs_dev, SSF_EINVAL,
            "ssf_np_step_linear: NULL argument");
  {
    int rc = np_form_system(K, C, A, alpha, dt, "ssf_np_step_linear");
    if (rc) return rc;
  }
  hipStream_t s = K->ctx->stream;
  const int64_t n = K->n;
  double *d = nullptr, *pv = nullptr, *f = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&d, n, s));
  c(dalloc(&pv, n, s));
  c(dalloc(&f, n, s));
  // d = lhs + (1 - alpha) dt v (:422-424)
  if (e == hipSuccess) c(axpy_rn(n, d, lhs_dev, (1.0 - alpha) * dt, tdlhs_dev, s));
  int rc = SSF_OK;
  // p = K d (:428)
  if (e == hipSuccess) rc = ssf_gmxv_dev(K, d, pv);
  // f = rhs - p (:441-443)
  if (e == hipSuccess && rc == SSF_OK) c(vec_sub(n, f, rhs_dev, pv, s));
  // A v = f from the current v (:449)
  if (e == hipSuccess && rc == SSF_OK) {
    ssf_solve_params q = *p;
    q.iv = 1;
    rc = ssf_solve_dev(A, &q, f, tdlhs_dev, info);
  }
  // lhs = d + alpha dt v (:454-456)
  if (e == hipSuccess && rc == SSF_OK) c(axpy_rn(n, lhs_dev, d, alpha * dt, tdlhs_dev, s));
  c(hipStreamSynchronize(s));
  dfree(d, s);
  dfree(pv, s);
  dfree(f, s);
  if (rc) return rc;
  SSF_HIP(e);
  return SSF_OK;
}

extern "C" int ssf_np_step_dform(ssf_matrix *K, ssf_matrix *C, ssf_matrix *A, double alpha,
                                 double dt, const ssf_solve_params *p, const double *rhs_dev,
                                 double *lhs_dev, double *tdlhs_dev, ssf_solve_info *info)
{
  SSF_ENTER();
  SSF_CHECK(K && C && A && p && rhs_dev && lhs_dev && tdlhs_dev, SSF_EINVAL,
            "ssf_np_step_dform: NULL argument");
  SSF_CHECK(alpha != 0.0 && dt != 0.0, SSF_EINVAL,
            "ssf_np_step_dform: alpha and dt must be nonzero (the rate divides by both)");
  {
    int rc = np_form_system(K, C, A, alpha, dt, "ssf_np_step_dform");
    if (rc) return rc;
  }
  hipStream_t s = K->ctx->stream;
  const int64_t n = K->n;
  double *d = nullptr, *pv = nullptr, *f = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&d, n, s));
  c(dalloc(&pv, n, s));
  c(dalloc(&f, n, s));
  // d = lhs + (1 - alpha) dt v (npsolvert.cpp:1152-1154)
  if (e == hipSuccess) c(axpy_rn(n, d, lhs_dev, (1.0 - alpha) * dt, tdlhs_dev, s));
  int rc = SSF_OK;
  // p = C d (:1156)
  if (e == hipSuccess) rc = ssf_gmxv_dev(C, d, pv);
  // f = rhs alpha dt + p (:1166-1168)
  if (e == hipSuccess && rc == SSF_OK) c(vec_sc2_add(n, f, rhs_dev, alpha, dt, pv, s));
  // A lhs = f from the current lhs (:1181)
  if (e == hipSuccess && rc == SSF_OK) {
    ssf_solve_params q = *p;
    q.iv = 1;
    rc = ssf_solve_dev(A, &q, f, lhs_dev, info);
  }
  // v = (lhs - d) / dt / alpha (:1183-1185)
  if (e == hipSuccess && rc == SSF_OK) c(vec_rate(n, tdlhs_dev, lhs_dev, d, dt, alpha, s));
  c(hipStreamSynchronize(s));
  dfree(d, s);
  dfree(pv, s);
  dfree(f, s);
  if (rc) return rc;
  SSF_HIP(e);
  return SSF_OK;
}
