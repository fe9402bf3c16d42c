// gfx950 kernels of the SIFEL linear-system path: CSR SpMV with fused dots, the
// preconditioned CG recurrence of GEFEL/scr.cpp:1093 (and its cg / comprow
// variants), reference-layout <-> full-CSR value maps, and the batched element
// kernels, which write their matrices for the assembly by entry (ssf_assemble.hip).
//
// Reductions are deterministic: every block writes one partial per quantity, and
// every consumer block re-reduces the whole partial array in the same fixed order,
// so each block holds bit-identical scalars without inter-block synchronisation
// and a system gives the same result whatever else runs on the device.
#include <algorithm>

#include "ssf_device.hpp"

namespace ssf {

// ------------------------------------------------------- element-wise helpers

__global__ void k_expand_ref(int64_t negm, const int *__restrict__ lo, const int *__restrict__ up,
                             const double *__restrict__ a, double *__restrict__ val)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= negm) return;
  const double v = a[k];
  val[lo[k]] = v;
  const int u = up[k];
  if (u >= 0) val[u] = v;
}

hipError_t expand_ref(int64_t negm, const int *map_lo, const int *map_up, const double *a,
                      double *val, hipStream_t s)
{
  if (negm == 0) return hipSuccess;
  const int64_t nb = (negm + TPB - 1) / TPB;
  hipLaunchKernelGGL(k_expand_ref, dim3((unsigned)nb), dim3(TPB), 0, s, negm, map_lo, map_up, a, val);
  return hipGetLastError();
}

__global__ void k_extract_diag(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                               const double *__restrict__ val, double *__restrict__ diag,
                               float *__restrict__ diag32)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // columns ascending: binary search for the diagonal
  int lo = rowptr[i], hi = rowptr[i + 1] - 1;
  double v = 0.0;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = col[mid];
    if (c == i) {
      v = val[mid];
      break;
    }
    if (c < i) lo = mid + 1;
    else hi = mid - 1;
  }
  diag[i] = v;
  if (diag32) diag32[i] = (float)v;
}

hipError_t extract_diag(int n, const int *rowptr, const int *col, const double *val,
                        double *diag, float *diag32, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_extract_diag, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, s, n, rowptr, col, val, diag, diag32);
  return hipGetLastError();
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b, double scale)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] * scale);
}

hipError_t to_f32(int64_t n, const double *a, float *b, double scale, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, a, b, scale);
  return hipGetLastError();
}

__global__ void k_axpy_f32(int64_t n, double alpha, const float *__restrict__ e, double *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += alpha * (double)e[i];
}

hipError_t axpy_f32(int64_t n, double alpha, const float *e, double *x, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy_f32, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, alpha, e, x);
  return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_dot_partial(int64_t n, const double *__restrict__ a,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ part)
{
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    acc += a[i] * b[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

hipError_t dot_partial(int64_t n, const double *a, const double *b, double *part, int nb,
                       hipStream_t s)
{
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(TPB), 0, s, n, a, b, part);
  return hipGetLastError();
}

__global__ void k_fill(int64_t n, double v, double *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

hipError_t fill(int64_t n, double v, double *x, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, v, x);
  return hipGetLastError();
}

__global__ void k_reduce_slot(const double *__restrict__ part, int nb, double *out)
{
  const double v = reduce_partials(part, nb);
  if (threadIdx.x == 0) *out = v;
}

hipError_t reduce_to_host_slot(const double *part, int nb, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce_slot, dim3(1), dim3(TPB), 0, s, part, nb, out);
  return hipGetLastError();
}

// plain streaming copy (16-byte accesses): the HBM reference point of the
// roofline and of the counter calibration
typedef double dvec2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(TPB) void k_copy(int64_t n2, const dvec2 *__restrict__ a,
                                              dvec2 *__restrict__ b)
{
  // 4 independent 16-byte loads in flight per thread before the stores
  const int64_t stride = (int64_t)gridDim.x * TPB;
  int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    const dvec2 v0 = __builtin_nontemporal_load(a + i);
    const dvec2 v1 = __builtin_nontemporal_load(a + i + stride);
    const dvec2 v2 = __builtin_nontemporal_load(a + i + 2 * stride);
    const dvec2 v3 = __builtin_nontemporal_load(a + i + 3 * stride);
    __builtin_nontemporal_store(v0, b + i);
    __builtin_nontemporal_store(v1, b + i + stride);
    __builtin_nontemporal_store(v2, b + i + 2 * stride);
    __builtin_nontemporal_store(v3, b + i + 3 * stride);
  }
  for (; i < n2; i += stride) b[i] = __builtin_nontemporal_load(a + i);
}

__global__ __launch_bounds__(TPB) void k_sell_gather(int64_t tot, const int *__restrict__ src,
                                                     const double *__restrict__ val,
                                                     double *__restrict__ out)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= tot) return;
  const int j = src[k];
  out[k] = j >= 0 ? val[j] : 0.0;
}

hipError_t sell_gather(int64_t tot, const int *src, const double *val, double *out, hipStream_t s)
{
  if (tot == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sell_gather, dim3((unsigned)((tot + TPB - 1) / TPB)), dim3(TPB), 0, s, tot,
                     src, val, out);
  return hipGetLastError();
}

hipError_t copy_kernel(int64_t n, const double *a, double *b, hipStream_t s)
{
  const int64_t n2 = n / 2;
  if (n2 == 0) return hipSuccess;
  const int64_t nb = std::min<int64_t>((n2 + TPB - 1) / TPB, 4096);
  hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(TPB), 0, s, n2, (const dvec2 *)a, (dvec2 *)b);
  return hipGetLastError();
}

// ----------------------------------------------------------------- assembly

// natural derivatives of the trilinear hexahedron (GEFEL/basefun.cpp:2317-2375)
__device__ __forceinline__ void hex_nat(double xi, double eta, double zeta, double *nx,
                                        double *ny, double *nz)
{
  const double xp = 1.0 + xi, xm = 1.0 - xi, yp = 1.0 + eta, ym = 1.0 - eta;
  const double zp = 1.0 + zeta, zm = 1.0 - zeta;
  nx[0] = yp * zp * 0.125; nx[1] = -(yp * zp * 0.125); nx[2] = -(ym * zp * 0.125); nx[3] = ym * zp * 0.125;
  nx[4] = yp * zm * 0.125; nx[5] = -(yp * zm * 0.125); nx[6] = -(ym * zm * 0.125); nx[7] = ym * zm * 0.125;
  ny[0] = xp * zp * 0.125; ny[1] = xm * zp * 0.125; ny[2] = -(xm * zp * 0.125); ny[3] = -(xp * zp * 0.125);
  ny[4] = xp * zm * 0.125; ny[5] = xm * zm * 0.125; ny[6] = -(xm * zm * 0.125); ny[7] = -(xp * zm * 0.125);
  nz[0] = xp * yp * 0.125; nz[1] = xm * yp * 0.125; nz[2] = xm * ym * 0.125; nz[3] = xp * ym * 0.125;
  nz[4] = -(xp * yp * 0.125); nz[5] = -(xm * yp * 0.125); nz[6] = -(xm * ym * 0.125); nz[7] = -(xp * ym * 0.125);
}

// global derivatives and |J| (GEFEL/difcalc.cpp:487-571); |J| sign fixed as
// MEFEL/SRC/linhex.cpp:619-623 / TRFEL/SRC/linhext.cpp:809-812
__device__ __forceinline__ double hex_global(const double *x, const double *y, const double *z,
                                             double xi, double eta, double zeta, double *dx,
                                             double *dy, double *dz)
{
  double nx[8], ny[8], nz[8];
  hex_nat(xi, eta, zeta, nx, ny, nz);
  double d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0, d7 = 0, d8 = 0, d9 = 0;
#pragma unroll
  for (int a = 0; a < 8; a++) {
    d1 += nx[a] * x[a]; d2 += ny[a] * x[a]; d3 += nz[a] * x[a];
    d4 += nx[a] * y[a]; d5 += ny[a] * y[a]; d6 += nz[a] * y[a];
    d7 += nx[a] * z[a]; d8 += ny[a] * z[a]; d9 += nz[a] * z[a];
  }
  const double jac = d1 * d5 * d9 + d4 * d8 * d3 + d7 * d2 * d6 - d7 * d5 * d3 - d8 * d6 * d1 -
                     d9 * d4 * d2;
  const double ij = 1.0 / jac;
  const double c00 = d5 * d9 - d8 * d6, c01 = d7 * d6 - d9 * d4, c02 = d4 * d8 - d7 * d5;
  const double c10 = d8 * d3 - d9 * d2, c11 = d1 * d9 - d7 * d3, c12 = d7 * d2 - d8 * d1;
  const double c20 = d2 * d6 - d5 * d3, c21 = d4 * d3 - d6 * d1, c22 = d1 * d5 - d4 * d2;
#pragma unroll
  for (int a = 0; a < 8; a++) {
    dx[a] = (nx[a] * c00 + ny[a] * c01 + nz[a] * c02) * ij;
    dy[a] = (nx[a] * c10 + ny[a] * c11 + nz[a] * c12) * ij;
    dz[a] = (nx[a] * c20 + ny[a] * c21 + nz[a] * c22) * ij;
  }
  return fabs(jac);
}

#define GPT 0.577350269189626  /* GEFEL/intp.cpp:26-27, weights 1 */

// hex8 heat conduction (TRFEL/SRC/linhext.cpp:774-840, isotrmat::matcond3d):
// K_e = sum_gp k |J| B^T B with B = grad N (3 x 8).
__device__ void ke_hex8_heat(const double *x, const double *y, const double *z, double kc,
                             double *K)
{
#pragma unroll
  for (int q = 0; q < 64; q++) K[q] = 0.0;
  for (int g = 0; g < 8; g++) {
    const double xi = (g & 4) ? GPT : -GPT, eta = (g & 2) ? GPT : -GPT, zeta = (g & 1) ? GPT : -GPT;
    double dx[8], dy[8], dz[8];
    const double jac = hex_global(x, y, z, xi, eta, zeta, dx, dy, dz);
    const double w = kc * jac;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = i; j < 8; j++) {
        const double s = (dx[i] * dx[j] + dy[i] * dy[j] + dz[i] * dz[j]) * w;
        K[i * 8 + j] += s;
        if (j != i) K[j * 8 + i] += s;
      }
  }
}

// quad4 plane stress (MEFEL/SRC/plelemlq.cpp:611-690, elastisomat::matstiff_plstress)
__device__ void ke_quad4_pstress(const double *x, const double *y, double e, double nu,
                                 double t, double *K)
{
  const double c = e / (1.0 - nu * nu);
  const double d00 = c, d01 = c * nu, d22 = e / 2.0 / (1.0 + nu);
#pragma unroll
  for (int q = 0; q < 64; q++) K[q] = 0.0;
  for (int g = 0; g < 4; g++) {
    const double xi = (g & 2) ? GPT : -GPT, eta = (g & 1) ? GPT : -GPT;
    double nx[4], ny[4];
    nx[0] = 0.25 * (1.0 + eta); nx[1] = 0.25 * (-1.0 - eta); nx[2] = 0.25 * (-1.0 + eta); nx[3] = 0.25 * (1.0 - eta);
    ny[0] = 0.25 * (1.0 + xi); ny[1] = 0.25 * (1.0 - xi); ny[2] = 0.25 * (-1.0 + xi); ny[3] = 0.25 * (-1.0 - xi);
    double d1 = 0, d2 = 0, d3 = 0, d4 = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      d1 += nx[a] * x[a]; d2 += ny[a] * x[a]; d3 += nx[a] * y[a]; d4 += ny[a] * y[a];
    }
    const double jac = d1 * d4 - d2 * d3;
    const double ij = 1.0 / jac;
    double bx[4], by[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      bx[a] = (nx[a] * d4 - ny[a] * d3) * ij;
      by[a] = (ny[a] * d1 - nx[a] * d2) * ij;
    }
    const double w = fabs(jac * t);  // thickness approx: bilinear of equal nodal values
    // B = [[bx0 0 bx1 0 ..], [0 by0 0 by1 ..], [by0 bx0 by1 bx1 ..]]
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        // 2x2 block (a,b) of B^T D B
        const double kxx = bx[a] * d00 * bx[b] + by[a] * d22 * by[b];
        const double kxy = bx[a] * d01 * by[b] + by[a] * d22 * bx[b];
        const double kyx = by[a] * d01 * bx[b] + bx[a] * d22 * by[b];
        const double kyy = by[a] * d00 * by[b] + bx[a] * d22 * bx[b];
        K[(2 * a) * 8 + 2 * b] += kxx * w;
        K[(2 * a) * 8 + 2 * b + 1] += kxy * w;
        K[(2 * a + 1) * 8 + 2 * b] += kyx * w;
        K[(2 * a + 1) * 8 + 2 * b + 1] += kyy * w;
      }
  }
}

// hex8 heat capacity (TRFEL/SRC/linhext.cpp:979-1029): C_e = sum_gp N^T N |J| rho c
// with rho the bilinear interpolation of equal nodal densities (linhext::approx)
__device__ void ke_hex8_cap(const double *x, const double *y, const double *z, double c,
                            double rho_e, double *K)
{
#pragma unroll
  for (int q = 0; q < 64; q++) K[q] = 0.0;
  for (int g = 0; g < 8; g++) {
    const double xi = (g & 4) ? GPT : -GPT, eta = (g & 2) ? GPT : -GPT, zeta = (g & 1) ? GPT : -GPT;
    double dx[8], dy[8], dz[8];
    const double jac = hex_global(x, y, z, xi, eta, zeta, dx, dy, dz);
    double n[8];
    n[0] = (1.0 + xi) * (1.0 + eta) * (1.0 + zeta) / 8.0;
    n[1] = (1.0 - xi) * (1.0 + eta) * (1.0 + zeta) / 8.0;
    n[2] = (1.0 - xi) * (1.0 - eta) * (1.0 + zeta) / 8.0;
    n[3] = (1.0 + xi) * (1.0 - eta) * (1.0 + zeta) / 8.0;
    n[4] = (1.0 + xi) * (1.0 + eta) * (1.0 - zeta) / 8.0;
    n[5] = (1.0 - xi) * (1.0 + eta) * (1.0 - zeta) / 8.0;
    n[6] = (1.0 - xi) * (1.0 - eta) * (1.0 - zeta) / 8.0;
    n[7] = (1.0 + xi) * (1.0 - eta) * (1.0 - zeta) / 8.0;
    double rho = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) rho += n[a] * rho_e;
    const double w = jac * (rho * c);  // hex_global returns |J|
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = i; j < 8; j++) {
        const double v = (n[i] * n[j]) * w;
        K[i * 8 + j] += v;
        if (j != i) K[j * 8 + i] += v;
      }
  }
}

// one thread per element: K_e (symmetric) into kbuf as its lower triangle (entry
// (i, j), j <= i, is value t = i (i + 1) / 2 + j), in groups of 64 elements: value t
// of element e at (e / 64) 64 36 + 64 t + e % 64, so that every store of a
// wavefront writes 64 consecutive doubles; and, for an element with prescribed
// values (pidx[e] >= 0), its lift values -sum_j K_ij u_j over the prescribed j into
// fbuf[pidx[e]][8] (rows with a code number >= 1)
template <int KIND>
__global__ __launch_bounds__(TPB) void k_assemble8(int64_t ne, const int *__restrict__ conn,
                                                   const int *__restrict__ cn,
                                                   const double *__restrict__ xyz,
                                                   const double *__restrict__ ma,
                                                   const double *__restrict__ mb,
                                                   const double *__restrict__ th,
                                                   double *__restrict__ kbuf,
                                                   const int *__restrict__ pidx,
                                                   const double *__restrict__ pv,
                                                   double *__restrict__ fbuf)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double K[64];
  if (KIND == SSF_ELEM_HEX8_HEAT || KIND == SSF_ELEM_HEX8_HEAT_CAP) {
    double x[8], y[8], z[8];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int nd = conn[(int64_t)e * 8 + a];
      x[a] = xyz[3 * (int64_t)nd]; y[a] = xyz[3 * (int64_t)nd + 1]; z[a] = xyz[3 * (int64_t)nd + 2];
    }
    if (KIND == SSF_ELEM_HEX8_HEAT) ke_hex8_heat(x, y, z, ma[e], K);
    else ke_hex8_cap(x, y, z, ma[e], mb ? mb[e] : 1.0, K);
  } else {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int nd = conn[(int64_t)e * 4 + a];
      x[a] = xyz[3 * (int64_t)nd]; y[a] = xyz[3 * (int64_t)nd + 1];
    }
    ke_quad4_pstress(x, y, ma[e], mb[e], th ? th[e] : 1.0, K);
  }
  double *kb = kbuf + (e >> 6) * (64 * 36) + (e & 63);
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) kb[64 * (i * (i + 1) / 2 + j)] = K[i * 8 + j];
  if (fbuf && pidx[e] >= 0) {
    int c[8];
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = cn[e * 8 + i];
    double *fb = fbuf + (int64_t)pidx[e] * 8;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (c[i] < 1) continue;
      double f = 0.0;
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (c[j] < 0) f += K[i * 8 + j] * pv[-c[j] - 1];
      fb[i] = -f;
    }
  }
}

hipError_t assemble_hex8_elast(int64_t ne, const int *conn, const int *cn, const double *xyz,
                               const double *ma, const double *mb, double *kbuf, const int *pidx,
                               const double *pv, double *fbuf, hipStream_t s);

hipError_t assemble_elements(int kind, int64_t ne, const int *conn, const int *cn,
                             const double *xyz, const double *mat_a, const double *mat_b,
                             const double *thick, double *kbuf, const int *pidx,
                             const double *pv, double *fbuf, hipStream_t s)
{
  if (ne <= 0) return hipSuccess;
  if (kind == SSF_ELEM_HEX8_ELAST)
    return assemble_hex8_elast(ne, conn, cn, xyz, mat_a, mat_b, kbuf, pidx, pv, fbuf, s);
  const unsigned nb = (unsigned)((ne + TPB - 1) / TPB);
  if (kind == SSF_ELEM_HEX8_HEAT)
    hipLaunchKernelGGL((k_assemble8<SSF_ELEM_HEX8_HEAT>), dim3(nb), dim3(TPB), 0, s, ne, conn, cn, xyz, mat_a, mat_b, thick, kbuf, pidx, pv, fbuf);
  else if (kind == SSF_ELEM_HEX8_HEAT_CAP)
    hipLaunchKernelGGL((k_assemble8<SSF_ELEM_HEX8_HEAT_CAP>), dim3(nb), dim3(TPB), 0, s, ne, conn, cn, xyz, mat_a, mat_b, thick, kbuf, pidx, pv, fbuf);
  else
    hipLaunchKernelGGL((k_assemble8<SSF_ELEM_QUAD4_PSTRESS>), dim3(nb), dim3(TPB), 0, s, ne, conn, cn, xyz, mat_a, mat_b, thick, kbuf, pidx, pv, fbuf);
  return hipGetLastError();
}

}  // namespace ssf

namespace ssf {

// ------------------------------------------------- value and vector updates
// Each operation rounds as the reference's loops do (no contraction): the
// nonstationary step and the matrix combinations stay bit-identical to it.

__global__ void k_vals_scale(int64_t n, double *__restrict__ a, double c)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) a[i] *= c;
}

__global__ void k_vals_add(int64_t n, double *__restrict__ a, double c, const double *__restrict__ b)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) a[i] += c * b[i];
}

// out = k * s, then out += 1.0 * cm (scalmat then addmat)
__global__ void k_vals_combine(int64_t n, double *__restrict__ out, const double *__restrict__ k,
                               double s, const double *__restrict__ cm)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  double v = k[i];
  v *= s;
  v += 1.0 * cm[i];
  out[i] = v;
}

// y = x + c * v
__global__ void k_axpy_rn(int64_t n, double *__restrict__ y, const double *__restrict__ x, double c,
                          const double *__restrict__ v)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) y[i] = x[i] + c * v[i];
}

// f = a - b
__global__ void k_sub(int64_t n, double *__restrict__ f, const double *__restrict__ a,
                      const double *__restrict__ b)
{
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) f[i] = a[i] - b[i];
}

// f = rhs alpha dt + p, the products left to right and unfused
__global__ void k_sc2_add(int64_t n, double *__restrict__ f, const double *__restrict__ rhs,
                          double alpha, double dt, const double *__restrict__ p)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) f[i] = rhs[i] * alpha * dt + p[i];
}

// v = (x - d) / dt / alpha, two correctly rounded divisions
__global__ void k_rate(int64_t n, double *__restrict__ v, const double *__restrict__ x,
                       const double *__restrict__ d, double dt, double alpha)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (i < n) v[i] = (x[i] - d[i]) / dt / alpha;
}

static unsigned grid_for(int64_t n) { return (unsigned)((std::max<int64_t>(n, 1) + TPB - 1) / TPB); }

hipError_t vals_scale(int64_t n, double *a, double c, hipStream_t s)
{
  hipLaunchKernelGGL(k_vals_scale, dim3(grid_for(n)), dim3(TPB), 0, s, n, a, c);
  return hipGetLastError();
}
hipError_t vals_add(int64_t n, double *a, double c, const double *b, hipStream_t s)
{
  hipLaunchKernelGGL(k_vals_add, dim3(grid_for(n)), dim3(TPB), 0, s, n, a, c, b);
  return hipGetLastError();
}
hipError_t vals_combine(int64_t n, double *out, const double *k, double sc, const double *cm,
                        hipStream_t s)
{
  hipLaunchKernelGGL(k_vals_combine, dim3(grid_for(n)), dim3(TPB), 0, s, n, out, k, sc, cm);
  return hipGetLastError();
}
hipError_t axpy_rn(int64_t n, double *y, const double *x, double c, const double *v, hipStream_t s)
{
  hipLaunchKernelGGL(k_axpy_rn, dim3(grid_for(n)), dim3(TPB), 0, s, n, y, x, c, v);
  return hipGetLastError();
}
hipError_t vec_sub(int64_t n, double *f, const double *a, const double *b, hipStream_t s)
{
  hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(TPB), 0, s, n, f, a, b);
  return hipGetLastError();
}
hipError_t vec_sc2_add(int64_t n, double *f, const double *rhs, double alpha, double dt,
                       const double *p, hipStream_t s)
{
  hipLaunchKernelGGL(k_sc2_add, dim3(grid_for(n)), dim3(TPB), 0, s, n, f, rhs, alpha, dt, p);
  return hipGetLastError();
}
hipError_t vec_rate(int64_t n, double *v, const double *x, const double *d, double dt,
                    double alpha, hipStream_t s)
{
  hipLaunchKernelGGL(k_rate, dim3(grid_for(n)), dim3(TPB), 0, s, n, v, x, d, dt, alpha);
  return hipGetLastError();
}

}  // namespace ssf
