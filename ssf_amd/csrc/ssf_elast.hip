// Batched hex8 linear-elasticity element kernel on FP64 MFMA (gfx950).
//
// MEFEL/SRC/linhex.cpp:583-640 computes K_e = sum over the 2x2x2 Gauss points of
// B^T D B |J| (geom_matrix :153-186, D of elastisomat::matstiff_spacestr,
// MEFEL/SRC/elastisomat.cpp:395-410). Stacking the eight 6x24 strain matrices
// gives G (48x24) and H = blockdiag(|J_g| D) G (48x24), so K_e = G^T H: a
// 24x24x48 product per element. One wavefront owns one element:
//   1. lane (g, a) = (lane/8, lane%8) computes the global derivatives of shape
//      function a at Gauss point g (every lane of a Gauss point recomputes the
//      3x3 Jacobian) and leaves them in LDS;
//   2. K_e, padded to 32x32, is three 16x16 tiles (0,0), (0,1), (1,1) of
//      v_mfma_f64_16x16x4f64, 12 K-steps each; operands are formed on the fly
//      from the LDS derivatives (G) and D (H);
//   3. each lane writes its tile entries of the lower triangle (the mirror of
//      tile (0,1) for the entries above the diagonal) to the element's 300
//      values of the matrix buffer, which the assembly by entry adds into the
//      reference layout in localize's order (ssf_assemble.hip); an element with
//      prescribed values also writes its lift values -K_e u_p.
#include "ssf_device.hpp"

namespace ssf {

typedef double double4_t __attribute__((ext_vector_type(4)));

#define GPT 0.577350269189626

__device__ __forceinline__ void hex_nat1(double xi, double eta, double zeta, int a, double &nx,
                                         double &ny, double &nz)
{
  // GEFEL/basefun.cpp:2317-2375: node a at (sx, sy, sz) in {+1,-1}^3 with the
  // ordering (+,+,+), (-,+,+), (-,-,+), (+,-,+), then the same at zeta = -1
  const double sx = (a == 0 || a == 3 || a == 4 || a == 7) ? 1.0 : -1.0;
  const double sy = (a == 0 || a == 1 || a == 4 || a == 5) ? 1.0 : -1.0;
  const double sz = (a < 4) ? 1.0 : -1.0;
  nx = sx * (1.0 + sy * eta) * (1.0 + sz * zeta) * 0.125;
  ny = sy * (1.0 + sx * xi) * (1.0 + sz * zeta) * 0.125;
  nz = sz * (1.0 + sx * xi) * (1.0 + sy * eta) * 0.125;
}

constexpr int EPB = TPB / 64;  // elements (wavefronts) per block

__global__ __launch_bounds__(TPB) void k_hex8_elast(int64_t ne, const int *__restrict__ conn,
                                                    const int *__restrict__ cn,
                                                    const double *__restrict__ xyz,
                                                    const double *__restrict__ E,
                                                    const double *__restrict__ NU,
                                                    double *__restrict__ kbuf,
                                                    const int *__restrict__ pidx,
                                                    const double *__restrict__ pv,
                                                    double *__restrict__ fbuf)
{
  __shared__ double s_xyz[EPB][3][8];
  __shared__ double s_b[EPB][8][8][3];  // [gp][node][d/dx, d/dy, d/dz]
  __shared__ double s_w[EPB][8];        // |J| of each Gauss point (weights are 1)
  __shared__ double s_k[EPB][24 * 24];  // K_e for the lift (only if needed)
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ep = (int64_t)blockIdx.x * EPB + wv;
  const bool active = ep < ne;
  const int64_t e = active ? ep : 0;
  if (active && lane < 24) {
    const int a = lane & 7, c = lane >> 3;
    const int nd = conn[(int64_t)e * 8 + a];
    s_xyz[wv][c][a] = xyz[3 * (int64_t)nd + c];
  }
  __syncthreads();
  // 1. derivatives: lane -> (gauss point g = lane / 8, node a = lane % 8)
  {
    const int g = lane >> 3, a = lane & 7;
    // Gauss point order of linhex.cpp:598-606: xi outer, eta, zeta inner
    const double xi = (g & 4) ? GPT : -GPT, eta = (g & 2) ? GPT : -GPT, zeta = (g & 1) ? GPT : -GPT;
    double d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0, d7 = 0, d8 = 0, d9 = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      double nx, ny, nz;
      hex_nat1(xi, eta, zeta, b, nx, ny, nz);
      const double x = s_xyz[wv][0][b], y = s_xyz[wv][1][b], z = s_xyz[wv][2][b];
      d1 += nx * x; d2 += ny * x; d3 += nz * x;
      d4 += nx * y; d5 += ny * y; d6 += nz * y;
      d7 += nx * z; d8 += ny * z; d9 += nz * z;
    }
    const double jac = d1 * d5 * d9 + d4 * d8 * d3 + d7 * d2 * d6 - d7 * d5 * d3 -
                       d8 * d6 * d1 - d9 * d4 * d2;
    const double ij = 1.0 / jac;
    double nx, ny, nz;
    hex_nat1(xi, eta, zeta, a, nx, ny, nz);
    s_b[wv][g][a][0] = (nx * (d5 * d9 - d8 * d6) + ny * (d7 * d6 - d9 * d4) + nz * (d4 * d8 - d7 * d5)) * ij;
    s_b[wv][g][a][1] = (nx * (d8 * d3 - d9 * d2) + ny * (d1 * d9 - d7 * d3) + nz * (d7 * d2 - d8 * d1)) * ij;
    s_b[wv][g][a][2] = (nx * (d2 * d6 - d5 * d3) + ny * (d4 * d3 - d6 * d1) + nz * (d1 * d5 - d4 * d2)) * ij;
    if (a == 0) s_w[wv][g] = fabs(jac);
  }
  __syncthreads();
  // material (elastisomat::matstiff_spacestr)
  const double em = active ? E[e] : 1.0, nu = active ? NU[e] : 0.25;
  const double gm = em / 2.0 / (1.0 + nu);
  const double sm = em / (1.0 + nu) / (1.0 - 2.0 * nu);
  const double dd0 = sm * (1.0 - nu), dd1 = sm * nu;

  // 2. MFMA tiles. Operand lane layout of v_mfma_f64_16x16x4f64:
  //    A[i][k]: i = lane % 16, k = lane / 16;  B[k][j]: j = lane % 16, k = lane / 16;
  //    D[i][j] (4 values per lane, one 64-bit value per row group):
  //    i = 4 * v + lane / 16, j = lane % 16.
  const int li = lane & 15, lk = lane >> 4;
  double4_t acc[3];
#pragma unroll
  for (int t = 0; t < 3; t++) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};
  // columns handled by this lane in tile row / column 0 and 1
  const int c0 = li, c1 = 16 + li;  // K_e column index of the operand
#pragma unroll
  for (int s = 0; s < 12; s++) {
    const int k = 4 * s + lk;  // row of G / H: Gauss point g = k / 6, strain c = k % 6
    const int g = k / 6, c = k % 6;
    const double w = s_w[wv][g];
    double gv[2], hv[2];
#pragma unroll
    for (int h2 = 0; h2 < 2; h2++) {
      const int col = h2 ? c1 : c0;
      double gval = 0.0, hval = 0.0;
      if (col < 24) {
        const int a = col / 3, comp = col % 3;
        const double bx = s_b[wv][g][a][0], by = s_b[wv][g][a][1], bz = s_b[wv][g][a][2];
        // strain-displacement entry B[c][col] (linhex.cpp:167-185)
        if (comp == 0) gval = (c == 0) ? bx : (c == 4) ? bz : (c == 5) ? by : 0.0;
        else if (comp == 1) gval = (c == 1) ? by : (c == 3) ? bz : (c == 5) ? bx : 0.0;
        else gval = (c == 2) ? bz : (c == 3) ? by : (c == 4) ? bx : 0.0;
        // (D B)[c][col]
        double db;
        if (comp == 0) db = (c == 0) ? dd0 * bx : (c < 3) ? dd1 * bx : (c == 3) ? 0.0 : (c == 4) ? gm * bz : gm * by;
        else if (comp == 1) db = (c == 1) ? dd0 * by : (c < 3) ? dd1 * by : (c == 3) ? gm * bz : (c == 4) ? 0.0 : gm * bx;
        else db = (c == 2) ? dd0 * bz : (c < 3) ? dd1 * bz : (c == 3) ? gm * by : (c == 4) ? gm * bx : 0.0;
        hval = db * w;
      }
      gv[h2] = gval;
      hv[h2] = hval;
    }
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[0], hv[0], acc[0], 0, 0, 0);  // (0,0)
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[0], hv[1], acc[1], 0, 0, 0);  // (0,1)
    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[1], hv[1], acc[2], 0, 0, 0);  // (1,1)
  }
  if (!active) return;
  // 3. the lower triangle (i >= j) into the element's kbuf values at
  //    i (i + 1) / 2 + j; the lift needs whole rows, from LDS
  double *kb = kbuf + e * 300;
  const bool want_lift = fbuf != nullptr && pidx[e] >= 0;  // wave-uniform
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const int ti = (t == 2) ? 1 : 0, tj = (t == 0) ? 0 : 1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = 16 * ti + 4 * v + lk, j = 16 * tj + li;
      if (i >= 24 || j >= 24) continue;
      const double kv = acc[t][v];
      if (t == 1) kb[j * (j + 1) / 2 + i] = kv;  // tile (0,1): i < j, stored as (j, i)
      else if (i >= j) kb[i * (i + 1) / 2 + j] = kv;
      if (want_lift) {
        s_k[wv][i * 24 + j] = kv;
        if (t == 1) s_k[wv][j * 24 + i] = kv;  // mirror of tile (0,1) is tile (1,0)
      }
    }
  }
  if (!want_lift) return;
  // s_k is written and read by lanes of this wavefront only
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
  if (lane < 24) {
    const int ci = cn[e * 24 + lane];
    if (ci >= 1) {
      double f = 0.0;
      for (int j = 0; j < 24; j++) {
        const int cj = cn[e * 24 + j];
        if (cj < 0) f += s_k[wv][lane * 24 + j] * pv[-cj - 1];
      }
      fbuf[(int64_t)pidx[e] * 24 + lane] = -f;
    }
  }
}

hipError_t assemble_hex8_elast(int64_t ne, const int *conn, const int *cn, const double *xyz,
                               const double *ma, const double *mb, double *kbuf, const int *pidx,
                               const double *pv, double *fbuf, hipStream_t s)
{
  if (ne <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((ne + EPB - 1) / EPB);
  hipLaunchKernelGGL(k_hex8_elast, dim3(nb), dim3(TPB), 0, s, ne, conn, cn, xyz, ma, mb, kbuf,
                     pidx, pv, fbuf);
  return hipGetLastError();
}

}  // namespace ssf
