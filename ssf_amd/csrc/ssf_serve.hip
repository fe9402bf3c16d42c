// Kernels of the served solves (gfx950): one launch sequence advances every
// resident request ("slot") by one iteration of the cg_prec recurrence
// (GEFEL/scr.cpp:1142-1197), each slot through its own descriptor. The mixed
// form runs an FP32 CG on the symmetrically scaled operator in two passes per
// iteration (the product with four dots, then the fused step k_sv_step), and an
// FP64 refresh (reliable update) ends every chunk.
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

// block partials (a, b) of the running kernel into part[2 * block] (the FP32
// recurrence's float sums are stored exactly)
template <typename T>
__device__ __forceinline__ void block_partials2(double *__restrict__ part, T a, T b)
{
  const T ta = block_sum(a);
  const T tb = block_sum(b);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = (double)ta;
    part[2 * blockIdx.x + 1] = (double)tb;
  }
}

// block partials of four sums into part[4 * block + j]
template <typename T>
__device__ __forceinline__ void block_partials4(double *__restrict__ part, T a, T b, T c, T d)
{
  const T v[4] = {block_sum(a), block_sum(b), block_sum(c), block_sum(d)};
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 4; j++) part[4 * blockIdx.x + j] = (double)v[j];
}

// slots a kernel skips: 0 finished or paused (done), 1 not fresh, 2 final (mixed)
constexpr int SK_DONE = 0, SK_FRESH = 1, SK_FINAL = 2;
__device__ __forceinline__ bool sv_skip(const SlotDesc &D, const SlotState &st, int mode)
{
  return mode == SK_FRESH ? !D.fresh : mode == SK_FINAL ? st.final[D.sid] != 0 : st.done[D.sid] != 0;
}

// One block per slot: the slot's sums of the partials of the previous kernel,
// in block order, into oa[stride * sid] (and ob[stride * sid] when ob is given).
// Slots that the previous kernel skipped are skipped here too.
__global__ __launch_bounds__(TPB) void k_sv_reduce(ServeDev S, SlotState st, double *oa,
                                                   double *ob, int stride, int mode)
{
  const SlotDesc D = S.slots[blockIdx.x];
  const int sid = D.sid;
  if (sv_skip(D, st, mode)) return;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < D.nblk; i += TPB) {
    a += S.part[2 * (D.blk0 + i)];
    b += S.part[2 * (D.blk0 + i) + 1];
  }
  a = block_sum(a);
  b = block_sum(b);
  if (threadIdx.x == 0) {
    oa[stride * sid] = a;
    if (ob) ob[stride * sid] = b;
  }
}

// the recurrence vectors of a slot in one precision
template <typename T>
struct SvVec {
  const T *val, *diag;
  T *x, *r, *d, *p, *h;
};
template <typename T>
__device__ __forceinline__ SvVec<T> sv_vec(const SlotDesc &D);
template <>
__device__ __forceinline__ SvVec<double> sv_vec<double>(const SlotDesc &D)
{
  return {D.sval, D.diag, D.x, D.r, D.d, D.p, D.h};
}
template <>
__device__ __forceinline__ SvVec<float> sv_vec<float>(const SlotDesc &D)
{
  return {D.sval32, nullptr, D.x32, D.r32, D.d32, D.p32, D.r32};
}

// one row sum of the SELL layout: sum_k val * x[col] in column order
template <typename T>
__device__ __forceinline__ T sv_row(const SlotDesc &D, const T *__restrict__ val,
                                    const T *__restrict__ x, int sl, int lane)
{
  typedef T tv __attribute__((ext_vector_type(SELL_V)));
  typedef int iv __attribute__((ext_vector_type(SELL_V)));
  typedef short sv __attribute__((ext_vector_type(SELL_V)));
  const int off = D.soff[sl];
  const int w = (D.soff[sl + 1] - off) / 64;
  T sum = T(0);
  if (D.sten) {  // stencil form: the values alone; slots in column order, 4 groups a round
    const int r = sl * 64 + lane;
    const int base = r - r % D.sten_b, nm1 = D.n - 1;
    const unsigned vb = (unsigned)(off + lane * SELL_V) * sizeof(T);
    int q = 0;
    for (; q + 4 * SELL_V <= w; q += 4 * SELL_V)
      sum = sten_groups<T, 4>(val, vb + q * 64 * sizeof(T), D.sten + q, x, base, nm1, sum);
    for (; q < w; q += SELL_V)
      sum = sten_groups<T, 1>(val, vb + q * 64 * sizeof(T), D.sten + q, x, base, nm1, sum);
  } else if (D.scol16) {
    const T *__restrict__ xs = x + sl * 64;
    for (int q = 0; q < w; q += SELL_V) {
      const int k = off + q * 64 + lane * SELL_V;
      const tv vv = ldnt((const tv *)(val + k));
      const sv cc = ldnt((const sv *)(D.scol16 + k));
#pragma unroll
      for (int u = 0; u < SELL_V; u++) sum += vv[u] * xs[cc[u]];
    }
  } else {
    for (int q = 0; q < w; q += SELL_V) {
      const int k = off + q * 64 + lane * SELL_V;
      const tv vv = ldnt((const tv *)(val + k));
      const iv cc = ldnt((const iv *)(D.scol + k));
#pragma unroll
      for (int u = 0; u < SELL_V; u++) sum += vv[u] * x[cc[u]];
    }
  }
  return sum;
}

template <typename T>
__device__ __forceinline__ const T *sv_half(const SlotDesc &D);
template <>
__device__ __forceinline__ const double *sv_half<double>(const SlotDesc &D)
{
  return D.half_v;
}
template <>
__device__ __forceinline__ const float *sv_half<float>(const SlotDesc &D)
{
  return D.half_v32;
}

// y = A x over the block's part of the slot in the half form; f(row, (A x)_row,
// x_row) for every row of the matrix. Row variant: the block's 64-row slices;
// block variant (3 unknowns per node): the 64-node slices t whose first row
// slice 3 t is the block's, so the slot's blocks cover every node once.
template <typename T, typename F>
__device__ __forceinline__ void sv_half_rows(const SlotDesc &D, const T *__restrict__ hv,
                                             const T *__restrict__ x, int s0, int s1, F f)
{
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  auto off = [&](int k) { return D.half_off[k]; };
  if (D.half_b == 1) {
    for (int sl = s0 + wave; sl < s1; sl += TPB / 64) {
      const int r = sl * 64 + lane;
      T xr;
      const T sum = half_row<T, 13>(hv, D.half_pad, off, x, r, D.n - 1, xr);
      if (r < D.n) f(r, sum, xr);
    }
  } else {
    const int nn = D.n / 3;
    for (int t = (s0 + 2) / 3 + wave; t < (s1 + 2) / 3; t += TPB / 64) {
      const int q = t * 64 + lane;
      T sum[3], xd[3];
      half_node<T, 3, 13>(hv, off, x, q, nn, sum, xd);
      if (q < nn)
#pragma unroll
        for (int u = 0; u < 3; u++) f(3 * q + u, sum[u], xd[u]);
    }
  }
}

// one block per slot: the slot's four sums of block_partials4, in block order,
// into st.sc4[4 sid + j] (slots the previous kernel skipped are skipped)
__global__ __launch_bounds__(TPB) void k_sv_reduce4(ServeDev S, SlotState st)
{
  const SlotDesc D = S.slots[blockIdx.x];
  const int sid = D.sid;
  if (st.done[sid]) return;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < D.nblk; i += TPB)
#pragma unroll
    for (int j = 0; j < 4; j++) a[j] += S.part[4 * (D.blk0 + i) + j];
#pragma unroll
  for (int j = 0; j < 4; j++) a[j] = block_sum(a[j]);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 4; j++) st.sc4[4 * sid + j] = a[j];
}

// p = A d over the block's slices: one wavefront per slice, one lane per row,
// SELL_V entries per load (16-bit or 32-bit columns), row sums in column order;
// or in the half form
// DOTS4: also r.r, r.p and p.p (block_partials4), for the fused step of the
// mixed form's FP32 recurrence (k_sv_step)
template <typename T, bool DOTS4>
__global__ __launch_bounds__(TPB) void k_sv_spmv(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const SvVec<T> V = sv_vec<T>(D);
  const int s0 = S.blk_s0[blk], s1 = S.blk_s1[blk];
  T acc = T(0), rr = T(0), rp = T(0), pp = T(0);
  auto row_done = [&](int row, T sum, T dr) {
    V.p[row] = sum;
    acc += dr * sum;
    if (DOTS4) {
      const T rv = V.r[row];
      rr += rv * rv;
      rp += rv * sum;
      pp += sum * sum;
    }
  };
  if (const T *hv = sv_half<T>(D)) {
    sv_half_rows<T>(D, hv, V.d, s0, s1, row_done);
  } else {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int sl = s0 + wave; sl < s1; sl += TPB / 64) {
      const T sum = sv_row<T>(D, V.val, V.d, sl, lane);
      const int row = sl * 64 + lane;
      if (row < D.n) row_done(row, sum, V.d[row]);
    }
  }
  if (DOTS4) block_partials4<T>(S.part, acc, rr, rp, pp);
  else block_partials2<T>(S.part, acc, T(0));
}

// The mixed form's FP32 step after k_sv_spmv<T, true> and k_sv_reduce4: alpha =
// r.r / d.p, x += alpha d, r -= alpha p and, with ||r - alpha p||^2 = r.r -
// 2 alpha r.p + alpha^2 p.p from the same dots, d = r + beta d in the same pass
// (as the single-system k_cg_fused_step); then ||r||^2 < nory pauses the slot
// until the chunk's refresh. |d.p| < zero stops it (status 2, GEFEL/scr.cpp:1148).
template <typename T>
__global__ __launch_bounds__(TPB) void k_sv_step(ServeDev S, SlotState st, double zero)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const SvVec<T> V = sv_vec<T>(D);
  const T dp = (T)st.sc4[4 * sid], rr = (T)st.sc4[4 * sid + 1];
  const T rp = (T)st.sc4[4 * sid + 2], pp = (T)st.sc4[4 * sid + 3];
  const long long it = st.iter[sid];
  if (fabs(dp) < (T)zero) {
    if (blk == D.blk0 && threadIdx.x == 0) {
      st.brk[sid] = 1;
      st.ani[sid] = it;
      st.ares[sid] = (double)rr;
    }
    return;
  }
  const T a = rr / dp;
  const T rrn = rr - T(2) * a * rp + a * a * pp;
  const T bt = rrn / rr;
  if (blk == D.blk0 && threadIdx.x == 0) {
    st.nom[2 * sid + ((it + 1) & 1)] = (double)rrn;
    st.pend[sid] = SV_ADVANCE | (rrn < (T)st.nory[sid] ? SV_STOP : 0);
    st.pend_status[sid] = 1;
  }
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    V.x[i] += a * V.d[i];
    const T ri = V.r[i] - a * V.p[i];
    V.r[i] = ri;
    V.d[i] = ri + bt * V.d[i];
  }
}

// den = d.p; |den| < zero stops the slot (status 2); x += a d, r -= a p,
// h = r / diag; slot sums r.r and r.h (the FP64 recurrence)
template <typename T>
__global__ __launch_bounds__(TPB) void k_sv_update(ServeDev S, SlotState st, double zero)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid]) return;
  const SvVec<T> V = sv_vec<T>(D);
  const T den = (T)st.sc[3 * sid];
  const T nom = (T)st.nom[2 * sid + (st.iter[sid] & 1)];
  T a;
  if (fabs(den) < (T)zero) {  // GEFEL/scr.cpp:1148-1152
    a = T(0);
    if (blk == D.blk0 && threadIdx.x == 0) {
      st.brk[sid] = 1;
      st.ani[sid] = st.iter[sid];
      st.ares[sid] = (double)nom;
    }
  } else {
    a = nom / den;
  }
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  T srr = T(0), srh = T(0);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    if (a != T(0)) {
      V.x[i] += a * V.d[i];
      V.r[i] -= a * V.p[i];
    }
    const T ri = V.r[i];
    srr += ri * ri;
    if (sizeof(T) == sizeof(double)) {
      const T hi = ri / V.diag[i];
      V.h[i] = hi;
      srh += ri * hi;
    }
  }
  block_partials2<T>(S.part, srr, sizeof(T) == sizeof(double) ? srh : srr);
}

// stopping test |r.r / nory| < res (GEFEL/scr.cpp:1188), else d = h + b d; the
// iteration cap ni ends the slot with status 0 after its ni-th iteration
template <typename T>
__global__ __launch_bounds__(TPB) void k_sv_dir(ServeDev S, SlotState st, double res,
                                                long long ni)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.done[sid] || st.brk[sid]) return;
  const SvVec<T> V = sv_vec<T>(D);
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
  const T bt = (T)nomn / (T)nom;
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) V.d[i] = bt * V.d[i] + V.h[i];
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
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.sc + 1, st.sc + 2, 3,
                     SK_FRESH);
  hipLaunchKernelGGL(k_sv_init2, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S,
                     const_cast<SlotDesc *>(S.slots), st, zero);
  return hipGetLastError();
}

hipError_t serve_iteration(const ServeDev &S, SlotState st, double res, double zero,
                           long long ni, hipStream_t s)
{
  if (S.nslot == 0) return hipSuccess;
  hipLaunchKernelGGL((k_sv_spmv<double, false>), dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.sc, nullptr, 3, SK_DONE);
  hipLaunchKernelGGL(k_sv_update<double>, dim3(S.nblk), dim3(TPB), 0, s, S, st, zero);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.sc + 1, st.sc + 2, 3,
                     SK_DONE);
  hipLaunchKernelGGL(k_sv_dir<double>, dim3(S.nblk), dim3(TPB), 0, s, S, st, res, ni);
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

// ----------------------------------------------- mixed form: reliable updates
//
// Per slot, the scheme of the single-system mixed solve: FP32 CG on
// float(S A S), S = diag(1/sqrt|a_ii|) (Jacobi-PCG on A in exact arithmetic),
// for the residual scaled by s = 1/||r|| of the last refresh; at the end of every
// chunk x64 += S x32 / s, r64 = b - A x64 in FP64, the FP64 test, and the FP32
// recurrence restarts on s' S r64 with its direction rescaled by s'/s.

// fresh slots: x64 = 0, x32 = 0, r64 = y; block partials of y.y
__global__ __launch_bounds__(TPB) void k_svm_init(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  if (!D.fresh) return;
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  double syy = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double yi = D.y[i];
    D.x[i] = 0.0;
    D.r[i] = yi;
    D.x32[i] = 0.0f;
    syy += yi * yi;
  }
  block_partials2<double>(S.part, syy, 0.0);
}

__device__ __forceinline__ double svm_scale(double rr, double old)
{
  return rr > 0.0 ? 1.0 / sqrt(rr) : old;
}

// fresh slots: the FP32 recurrence on S y / ||y||: r32 = d32; block partials of
// r32.r32 (twice: it is also the recurrence's r.h)
__global__ __launch_bounds__(TPB) void k_svm_init32(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  if (!D.fresh) return;
  const double sc = svm_scale(st.nory64[D.sid], 1.0);
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  float srr = 0.f;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const float ri = (float)(D.y[i] * sc * D.isd[i]);
    D.r32[i] = ri;
    D.d32[i] = ri;
    srr += ri * ri;
  }
  block_partials2<float>(S.part, srr, srr);
}

// one thread per fresh slot: the scalars of both recurrences; zero rhs: x = 0
__global__ void k_svm_init2(ServeDev S, SlotDesc *slots, SlotState st, double res, double delta,
                            double zero)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.nslot || !slots[k].fresh) return;
  const int sid = slots[k].sid;
  const double yy = st.nory64[sid];
  const double sc = svm_scale(yy, 1.0);
  st.rr64[sid] = yy;
  st.best[sid] = yy;
  st.scale[sid] = sc;
  st.delta2[sid] = delta * delta;
  st.bad[sid] = 0;
  st.since_best[sid] = 0;
  st.updates[sid] = 0;
  st.final[sid] = 0;
  st.mstatus[sid] = 0;
  // refresh once the FP32 residual falls by delta, or to the target if closer
  const double nom = st.sc[3 * sid + 2];
  st.nory[sid] = (double)(float)(nom * fmax(delta * delta, res));
  st.nom[2 * sid] = nom;
  st.nom[2 * sid + 1] = nom;
  st.iter[sid] = 0;
  st.status[sid] = 0;
  st.done[sid] = 0;
  st.pend[sid] = 0;
  st.brk[sid] = 0;
  st.ani[sid] = -1;
  st.ares[sid] = 0.0;
  if (yy < zero) {  // zero right-hand side: x = 0 (GEFEL/scr.cpp:821-830)
    st.final[sid] = 1;
    st.mstatus[sid] = 3;
    st.done[sid] = 1;
  }
  slots[k].fresh = 0;
}

// x64 += S x32 / s, x32 = 0; block partials of ||x64||^2
__global__ __launch_bounds__(TPB) void k_svm_xupdate(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  if (st.final[D.sid]) return;
  const double inv = 1.0 / st.scale[D.sid];
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  double xx = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const double xi = D.x[i] + (double)D.x32[i] * inv * D.isd[i];
    D.x[i] = xi;
    D.x32[i] = 0.0f;
    xx += xi * xi;
  }
  block_partials2<double>(S.part, xx, 0.0);
}

// r64 = y - A x64 in FP64; block partials of ||r64||^2
__global__ __launch_bounds__(TPB) void k_svm_resid(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  if (st.final[D.sid]) return;
  const int s0 = S.blk_s0[blk], s1 = S.blk_s1[blk];
  double acc = 0.0;
  if (D.half_v) {
    sv_half_rows<double>(D, D.half_v, D.x, s0, s1, [&](int row, double ax, double) {
      const double ri = D.y[row] - ax;
      D.r[row] = ri;
      acc += ri * ri;
    });
  } else {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int sl = s0 + wave; sl < s1; sl += TPB / 64) {
      const double ax = sv_row<double>(D, D.sval, D.x, sl, lane);
      const int row = sl * 64 + lane;
      if (row < D.n) {
        const double ri = D.y[row] - ax;
        D.r[row] = ri;
        acc += ri * ri;
      }
    }
  }
  block_partials2<double>(S.part, acc, 0.0);
}

// the FP32 recurrence restarts on s' S r64 (s' = 1/||r64||), its direction
// rescaled by s'/s, or restarted from r32 when ||r64|| rose 1e4-fold above its
// best; block partials of r32.r32 (twice). rr64 of the refresh is st.sc[3 sid].
__global__ __launch_bounds__(TPB) void k_svm_refresh(ServeDev S, SlotState st)
{
  const int blk = blockIdx.x;
  const SlotDesc D = S.slots[S.blk_slot[blk]];
  const int sid = D.sid;
  if (st.final[sid]) return;
  const double rr64 = st.sc[3 * sid];
  const double old = st.scale[sid];
  const double sc = svm_scale(rr64, old);
  const float c = (float)(sc / old);
  const bool restart = rr64 > 1e4 * st.best[sid];
  const int r0 = S.blk_s0[blk] * 64, r1 = min(D.n, S.blk_s1[blk] * 64);
  float srr = 0.f;
  for (int i = r0 + threadIdx.x; i < r1; i += TPB) {
    const float ri = (float)(D.r[i] * sc * D.isd[i]);
    D.r32[i] = ri;
    D.d32[i] = restart ? ri : D.d32[i] * c;
    srr += ri * ri;
  }
  block_partials2<float>(S.part, srr, srr);
}

// below this the FP64 residual is within the rounding error of its own
// computation: about 64 eps (||b|| + ||A||_inf ||x||) in 2-norm
constexpr double SVM_FLOOR = 64.0 * 2.220446049250313e-16;

// one thread per slot: the FP64 test and the FP32 recurrence's new scalars;
// counts the slots not yet final into *active
__global__ void k_svm_check(ServeDev S, SlotState st, double res, long long ni, int *active)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.nslot) return;
  const SlotDesc &D = S.slots[k];
  const int sid = D.sid;
  if (!st.final[sid]) {
    const double rr64 = st.sc[3 * sid];
    const double nomn = st.sc[3 * sid + 2];
    const double prev = st.rr64[sid];
    st.rr64[sid] = rr64;
    st.updates[sid] += 1;
    const bool met = st.done[sid] && st.status[sid] == 1;
    if (rr64 < res * st.nory64[sid]) {
      st.final[sid] = 1;
      st.mstatus[sid] = 1;
      st.done[sid] = 1;
    } else {
      st.bad[sid] = (met && rr64 > 0.25 * prev) ? st.bad[sid] + 1 : 0;
      if (rr64 < 0.81 * st.best[sid]) st.since_best[sid] = 0;
      else st.since_best[sid] += 1;
      if (rr64 < st.best[sid]) st.best[sid] = rr64;
      const double fl = SVM_FLOOR * (sqrt(st.nory64[sid]) + D.norm_a[0] * sqrt(st.xx[sid]));
      if (rr64 < fl * fl && (st.bad[sid] >= 3 || st.since_best[sid] >= 10)) {
        st.final[sid] = 1;  // FP64 floor reached above res
        st.mstatus[sid] = 4;
        st.done[sid] = 1;
      } else if (st.iter[sid] >= ni) {
        st.final[sid] = 1;  // ni FP32 iterations spent
        st.mstatus[sid] = 0;
        st.done[sid] = 1;
      } else {
        st.scale[sid] = svm_scale(rr64, st.scale[sid]);
        // next refresh once the FP32 residual falls by delta from its value
        // now (nomn), or to the final target if that is closer
        st.nory[sid] = (double)(float)(nomn * fmax(st.delta2[sid], res * st.nory64[sid] / rr64));
        st.nom[2 * sid + (st.iter[sid] & 1)] = nomn;
        st.done[sid] = 0;
        st.status[sid] = 0;
        st.pend[sid] = 0;
        st.brk[sid] = 0;
      }
    }
  }
  if (active && !st.final[sid]) atomicAdd(active, 1);
}

hipError_t serve_mixed_init(const ServeDev &S, SlotState st, double res, double delta,
                            double zero, hipStream_t s)
{
  if (S.nslot == 0) return hipSuccess;
  hipLaunchKernelGGL(k_svm_init, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.nory64, nullptr, 1,
                     SK_FRESH);
  hipLaunchKernelGGL(k_svm_init32, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.sc + 1, st.sc + 2, 3,
                     SK_FRESH);
  hipLaunchKernelGGL(k_svm_init2, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S,
                     const_cast<SlotDesc *>(S.slots), st, res, delta, zero);
  return hipGetLastError();
}

hipError_t serve_mixed_iteration(const ServeDev &S, SlotState st, hipStream_t s)
{
  if (S.nslot == 0) return hipSuccess;
  // two passes over the vectors: the product with its four dots, the fused step
  hipLaunchKernelGGL((k_sv_spmv<float, true>), dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce4, dim3(S.nslot), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_step<float>, dim3(S.nblk), dim3(TPB), 0, s, S, st, 1e-30);
  hipLaunchKernelGGL(k_sv_advance, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S, st,
                     nullptr);
  return hipGetLastError();
}

hipError_t serve_mixed_refresh(const ServeDev &S, SlotState st, double res, long long ni,
                               int *active, hipStream_t s)
{
  hipLaunchKernelGGL(k_sv_zero, dim3(1), dim3(1), 0, s, active);
  if (S.nslot == 0) return hipGetLastError();
  hipLaunchKernelGGL(k_svm_xupdate, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.xx, nullptr, 1,
                     SK_FINAL);
  hipLaunchKernelGGL(k_svm_resid, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.sc, nullptr, 3,
                     SK_FINAL);
  hipLaunchKernelGGL(k_svm_refresh, dim3(S.nblk), dim3(TPB), 0, s, S, st);
  hipLaunchKernelGGL(k_sv_reduce, dim3(S.nslot), dim3(TPB), 0, s, S, st, st.sc + 1, st.sc + 2, 3,
                     SK_FINAL);
  hipLaunchKernelGGL(k_svm_check, dim3((S.nslot + TPB - 1) / TPB), dim3(TPB), 0, s, S, st, res, ni,
                     active);
  return hipGetLastError();
}

}  // namespace ssf
