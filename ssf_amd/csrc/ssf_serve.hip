// Kernels of the served solves (gfx950): one launch sequence advances every
// resident request ("slot") by one iteration of the cg_prec recurrence
// (GEFEL/scr.cpp:1142-1197), each slot through its own descriptor.
//
// Block b of a launch works on the 64-row slices [blk_s0[b], blk_s1[b]) of slot
// blk_slot[b] (the SELL-64 layout of the slot's own matrix). Every block leaves its partial sums; a small kernel with
// one block per slot then sums a slot's partials in block order into the
// slot's scalars (d.p, r.r, r.h), read by the next kernel. The summation order
// depends on the slot's own tiling only, so a request's result is the same alone
// and among others. State changes decided inside a kernel are left pending and
// applied by k_sv_advance.
#include "ssf_serve.hpp"

namespace ssf {

constexpr int SV_ADVANCE = 1;
constexpr int SV_STOP = 2;

// block partials (a, b) of the running kernel into part[2 * block]
__device__ __forceinline__ void block_partials2(double *__restrict__ part, double a, double b)
{
  const double ta = block_sum(a);
  const double tb = block_sum(b);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

// One block per slot: the slot's sums of the partials of the previous kernel,
// in block order, into out = st.sc + 3 sid + off (count values). Slots that the
// previous kernel skipped are skipped here too (done, or not fresh for init).
__global__ __launch_bounds__(TPB) void k_sv_reduce(ServeDev S, SlotState st, int off, int count,
                                                   int fresh_only)
{
  const SlotDesc D = S.slots[blockIdx.x];
  const int sid = D.sid;
  if (fresh_only ? !D.fresh : st.done[sid]) return;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < D.nblk; i += TPB) {
    a += S.part[2 * (D.blk0 + i)];
    b += S.part[2 * (D.blk0 + i) + 1];
  }
  a = block_sum(a);
  b = block_sum(b);
  if (threadIdx.x == 0) {
    st.sc[3 * sid + off] = a;
    if (count > 1) st.sc[3 * sid + off + 1] = b;
  }
}

// p = A d over the block's slices: one wavefront per slice, one lane per row,
// SELL_V entries per load (16-bit or 32-bit columns), row sums in column order
__global__ __launch_bounds__(TPB) void k_sv_spmv(ServeDev S, SlotState st)
{
  typedef double tv __attribute__((ext_vector_type(SELL_V)));
  typedef int iv __attribute__((ext_vector_type(SELL_V)));
  typedef short sv __attribute__((ext_vector_type(SELL_V)));
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const int s0 = S.blk_s0[blk], s1 = S.blk_s1[blk];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double acc = 0.0;
  for (int sl = s0 + wave; sl < s1; sl += TPB / 64) {
    const int off = D.soff[sl];
    const int w = (D.soff[sl + 1] - off) / 64;
    double sum = 0.0;
    if (D.scol16) {
      const double *__restrict__ xs = D.d + sl * 64;
      for (int q = 0; q < w; q += SELL_V) {
        const int k = off + q * 64 + lane * SELL_V;
        const tv vv = ldnt((const tv *)(D.sval + k));
        const sv cc = ldnt((const sv *)(D.scol16 + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) sum += vv[u] * xs[cc[u]];
      }
    } else {
      for (int q = 0; q < w; q += SELL_V) {
        const int k = off + q * 64 + lane * SELL_V;
        const tv vv = ldnt((const tv *)(D.sval + k));
        const iv cc = ldnt((const iv *)(D.scol + k));
#pragma unroll
        for (int u = 0; u < SELL_V; u++) sum += vv[u] * D.d[cc[u]];
      }
    }
    const int row = sl * 64 + lane;
    if (row < D.n) {
      D.p[row] = sum;
      acc += D.d[row] * sum;
    }
  }
  block_partials2(S.part, acc, 0.0);
}

// den = d.p; |den| < zero stops the slot (status 2); x += a d, r -= a p,
// h = r / diag; slot sums r.r and r.h
__global__ __launch_bounds__(TPB) void k_sv_update(ServeDev S, SlotState st, double zero)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const double den = st.sc[3 * sid];
  const double nom = st.nom[2 * sid + (st.iter[sid] & 1)];
  double a;
  if (fabs(den) < zero) {  // GEFEL/scr.cpp:1148-1152
    a = 0.0;
    if (blk == D.blk0 && threadIdx.x == 0) {
      st.brk[sid] = 1;
      st.ani[sid] = st.iter[sid];
      st.ares[sid] = nom;
    }
  } else {
    a = nom / den;
  }
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  double srr = 0.0, srh = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    if (a != 0.0) {
      D.x[i] += a * D.d[i];
      D.r[i] -= a * D.p[i];
    }
    const double ri = D.r[i];
    const double hi = ri / D.diag[i];
    D.h[i] = hi;
    srr += ri * ri;
    srh += ri * hi;
  }
  block_partials2(S.part, srr, srh);
}

// stopping test |r.r / nory| < res (GEFEL/scr.cpp:1188), else d = h + b d; the
// iteration cap ni ends the slot with status 0 after its ni-th iteration
__global__ __launch_bounds__(TPB) void k_sv_dir(ServeDev S, SlotState st, double res,
                                                long long ni)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid] || st.brk[sid]) return;
  const double rr = st.sc[3 * sid + 1];
  const double nomn = st.sc[3 * sid + 2];
  const long long it = st.iter[sid];
  const double nom = st.nom[2 * sid + (it & 1)];
  const bool stop = fabs(rr / st.nory[sid]) < res;
  if (blk == D.blk0 && threadIdx.x == 0) {
    st.nom[2 * sid + ((it + 1) & 1)] = nomn;
    if (stop) {
      st.ani[sid] = it;
      st.ares[sid] = nomn / st.nory[sid];
      st.pend[sid] = SV_STOP;
      st.pend_status[sid] = 1;
    } else if (it + 1 >= ni) {
      st.pend[sid] = SV_ADVANCE | SV_STOP;
      st.pend_status[sid] = 0;
    } else {
      st.pend[sid] = SV_ADVANCE;
    }
  }
  if (stop) return;
  const double bt = nomn / nom;
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) D.d[i] = bt * D.d[i] + D.h[i];
}

// one thread per slot: apply the pending changes, count the unfinished slots
__global__ void k_sv_advance(ServeDev S, SlotState st, int *active)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.nslot) return;
  const int sid = S.slots[k].sid;
  if (st.brk[sid]) {
    st.done[sid] = 1;
    st.status[sid] = 2;
    st.brk[sid] = 0;
    st.pend[sid] = 0;
  }
  const int pd = st.pend[sid];
  if (pd) {
    if (pd & SV_ADVANCE) st.iter[sid] += 1;
    if (pd & SV_STOP) {
      st.done[sid] = 1;
      st.status[sid] = st.pend_status[sid];
    }
    st.pend[sid] = 0;
  }
  if (active && !st.done[sid]) atomicAdd(active, 1);
}

// fresh slots: x = 0, r = y, h = r / diag, d = h; slot sums y.y and r.h
__global__ __launch_bounds__(TPB) void k_sv_init(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  if (!D.fresh) return;
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  double syy = 0.0, srh = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double yi = D.y[i];
    const double hi = yi / D.diag[i];
    D.x[i] = 0.0;
    D.r[i] = yi;
    D.h[i] = hi;
    D.d[i] = hi;
    syy += yi * yi;
    srh += yi * hi;
  }
  block_partials2(S.part, syy, srh);
}

// one thread per slot: initial scalars of the fresh slots; the fresh mark is
// cleared so the iteration kernels take the slot from now on
__global__ void k_sv_init2(ServeDev S, SlotDesc *slots, SlotState st, double zero)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.nslot) return;
  if (!slots[k].fresh) return;
  const int sid = slots[k].sid;
  const double nory = st.sc[3 * sid + 1];
  const double nom = st.sc[3 * sid + 2];
  st.nory[sid] = nory;
  st.nom[2 * sid] = nom;
  st.nom[2 * sid + 1] = nom;
  st.iter[sid] = 0;
  st.status[sid] = 0;
  st.done[sid] = 0;
  st.pend[sid] = 0;
  st.brk[sid] = 0;
  st.ani[sid] = -1;
  st.ares[sid] = 0.0;
  if (nory < zero) {  // zero right-hand side: x = 0 (GEFEL/scr.cpp:821-830)
    st.status[sid] = 3;
    st.done[sid] = 1;
    st.ani[sid] = 0;
    st.ares[sid] = nory;
  }
  slots[k].fresh = 0;
}

hipError_t serve_init(const ServeDev &S, SlotState st, double zero, hipStream_t s)
{
  if (S.nslot == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sv_init, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, 1, 2, 1);
  hipLaunchKernelGGL(k_sv_init2, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S,
                     const_cast<SlotDesc *>(S.slots), st, zero);
  return hipGetLastError();
}

hipError_t serve_iteration(const ServeDev &S, SlotState st, double res, double zero,
                           long long ni, hipStream_t s)
{
  if (S.nslot == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sv_spmv, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, 0, 1, 0);
  hipLaunchKernelGGL(k_sv_update, dim3(S.nblk), dim3(TPB), 0, s, S, st, zero);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, 1, 2, 0);
  hipLaunchKernelGGL(k_sv_dir, dim3(S.nblk), dim3(TPB), 0, s, S, st, res, ni);
  hipLaunchKernelGGL(k_sv_advance, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S, st,
                     nullptr);
  return hipGetLastError();
}

__global__ void k_sv_zero(int *p) { *p = 0; }

hipError_t serve_flags(const ServeDev &S, SlotState st, int *active, hipStream_t s)
{
  hipLaunchKernelGGL(k_sv_zero, dim3(1), dim3(1), 0, s, active);
  if (S.nslot > 0)
    hipLaunchKernelGGL(k_sv_advance, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S, st,
                       active);
  return hipGetLastError();
}

}  // namespace ssf
