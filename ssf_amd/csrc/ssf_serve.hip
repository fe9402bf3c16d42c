// Kernels of the served solves (gfx950): one launch sequence advances every
// resident request ("slot") by one iteration of the cg_prec recurrence
// (GEFEL/scr.cpp:1142-1197), each slot through its own descriptor.
//
// Block b of a launch works on rows [blk_r0[b], blk_r1[b]) of slot blk_slot[b];
// a slot's blocks are contiguous in the table and its rows are split in chunks
// of SERVE_RPB, so its partition, and with it every summation order, depends on
// its own size only: a request's result is the same alone and among others.
// Scalars of a slot are re-reduced from its blocks' partials by each of its
// blocks (no atomics, no inter-block waits); state changes decided inside a
// kernel are left pending and applied by k_sv_advance.
#include "ssf_serve.hpp"

namespace ssf {

constexpr int SV_ADVANCE = 1;
constexpr int SV_STOP = 2;

// sum of part[b0 + i], i < nb, by the whole block: thread t adds i = t, t + TPB,
// ..., then the fixed block_sum tree (depends on nb only)
__device__ __forceinline__ double slot_sum(const double *__restrict__ part, int b0, int nb)
{
  double t = 0.0;
  for (int i = threadIdx.x; i < nb; i += TPB) t += part[b0 + i];
  return block_sum(t);
}

// p = A d for the slot's rows, partial d.p
template <int G>
__global__ __launch_bounds__(TPB) void k_sv_spmv(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const int r0 = S.blk_r0[blk], r1 = S.blk_r1[blk];
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  constexpr int NG = TPB / G;
  double acc = 0.0;
  for (int row = r0 + grp; row < r1; row += NG) {
    const double s = row_dot<double, G>(D.rowptr, D.col, D.val, D.d, row, lane);
    if (lane == 0) {
      D.p[row] = s;
      acc += D.d[row] * s;
    }
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) S.part[blk] = t;
}

// den = d.p; |den| < zero stops the slot (status 2); x += a d, r -= a p,
// h = r / diag; partials r.r and r.h
__global__ __launch_bounds__(TPB) void k_sv_update(ServeDev S, SlotState st, double zero)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const double den = slot_sum(S.part, D.blk0, D.nblk);
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
  const int r0 = S.blk_r0[blk], r1 = S.blk_r1[blk];
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
  const double trr = block_sum(srr);
  const double trh = block_sum(srh);
  if (threadIdx.x == 0) {
    S.part[S.nblk + blk] = trr;
    S.part[2 * S.nblk + blk] = trh;
  }
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
  const double rr = slot_sum(S.part + S.nblk, D.blk0, D.nblk);
  const double nomn = slot_sum(S.part + 2 * S.nblk, D.blk0, D.nblk);
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
  const int r0 = S.blk_r0[blk], r1 = S.blk_r1[blk];
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

// fresh slots: x = 0, r = y, h = r / diag, d = h; partials y.y and r.h
__global__ __launch_bounds__(TPB) void k_sv_init(ServeDev S)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  if (!D.fresh) return;
  const int r0 = S.blk_r0[blk], r1 = S.blk_r1[blk];
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
  const double a = block_sum(syy);
  const double b = block_sum(srh);
  if (threadIdx.x == 0) {
    S.part[S.nblk + blk] = a;
    S.part[2 * S.nblk + blk] = b;
  }
}

// one block per slot: initial scalars of the fresh slots; the fresh mark is
// cleared so the iteration kernels take the slot from now on
__global__ __launch_bounds__(TPB) void k_sv_init2(ServeDev S, SlotDesc *slots, SlotState st,
                                                  double zero)
{
  const int k = blockIdx.x;
  const SlotDesc D = slots[k];
  if (!D.fresh) return;
  const double nory = slot_sum(S.part + S.nblk, D.blk0, D.nblk);
  const double nom = slot_sum(S.part + 2 * S.nblk, D.blk0, D.nblk);
  if (threadIdx.x != 0) return;
  const int sid = D.sid;
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
  hipLaunchKernelGGL(k_sv_init, dim3(S.nblk), dim3(TPB), 0, s, S);
  hipLaunchKernelGGL(k_sv_init2, dim3(S.nslot), dim3(TPB), 0, s, S,
                     const_cast<SlotDesc *>(S.slots), st, zero);
  return hipGetLastError();
}

hipError_t serve_iteration(const ServeDev &S, SlotState st, double res, double zero,
                           long long ni, hipStream_t s)
{
  if (S.nslot == 0) return hipSuccess;
  hipLaunchKernelGGL((k_sv_spmv<8>), dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_update, dim3(S.nblk), dim3(TPB), 0, s, S, st, zero);
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
