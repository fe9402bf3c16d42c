// Batched hex8 linear-elasticity element kernel on FP64 MFMA (gfx950).
//
// MEFEL/SRC/linhex.cpp:583-640 computes K_e = sum over the 2x2x2 Gauss points of
// B^T D B |J| (geom_matrix :153-186, D of elastisomat::matstiff_spacestr,
// MEFEL/SRC/elastisomat.cpp:395-410). With b_g the 24 global shape-function
// derivatives dN_a/dx_p of Gauss point g (index 3a + p) and the isotropic D
// (normal diagonal dd0, normal off-diagonal dd1, shear gm), every entry of B^T D B
// is a combination of the products b_g[3a+p] b_g[3b+q]:
//   K[(a,p),(b,q)] = dd1 M[3a+p][3b+q] + gm M[3a+q][3b+p]                 (p != q)
//   K[(a,p),(b,p)] = dd0 M[3a+p][3b+p] + gm (M[3a+r][3b+r] + M[3a+s][3b+s])  (r, s != p)
// with M = sum_g |J_g| b_g b_g^T, a 24x24x8 product instead of B^T D B's
// 24x24x48. One wavefront owns one element:
//   1. lane (g, a) = (lane/8, lane%8) computes the global derivatives of shape
//      function a at Gauss point g (every lane of a Gauss point recomputes the
//      3x3 Jacobian) and leaves them in LDS;
//   2. M, padded to 32x32, is three 16x16 tiles (0,0), (0,1), (1,1) of
//      v_mfma_f64_16x16x4f64, 2 K-steps (Gauss points 0-3, 4-7) each; its 3x3
//      blocks (a, b) with a >= b go to LDS;
//   3. lane l forms the lower-triangle entries l, l + 64, ... of K_e from M and
//      stores them to the element's 300 values of the matrix buffer (64
//      consecutive doubles per store), which the assembly by entry adds into the
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

// the 3x3 blocks (a, b), a >= b, of the symmetric M: block a (a + 1) / 2 + b,
// entry (r, c) at 9 blk + 3 r + c
constexpr int MBLK = 36 * 9;

__device__ __forceinline__ int mpos(int i, int j)  // M[i][j] with i / 3 >= j / 3
{
  const int a = i / 3, b = j / 3;
  return 9 * (a * (a + 1) / 2 + b) + 3 * (i - 3 * a) + (j - 3 * b);
}

// K_e[I][J], I >= J (so I / 3 >= J / 3), from the blocks of M
__device__ __forceinline__ double ke_entry(const double *m, int I, int J, double dd0, double dd1,
                                           double gm)
{
  const int a = I / 3, p = I - 3 * a, b = J / 3, q = J - 3 * b;
  const double *bl = m + 9 * (a * (a + 1) / 2 + b);
  if (p != q) return dd1 * bl[3 * p + q] + gm * bl[3 * q + p];
  const int r = p == 0 ? 1 : 0, t = p == 2 ? 1 : 2;
  return dd0 * bl[4 * p] + gm * (bl[4 * r] + bl[4 * t]);
}

// lower-triangle position idx = i (i + 1) / 2 + j, j <= i < 24
__device__ __forceinline__ void tri_ij(int idx, int &i, int &j)
{
  int r = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
  if ((r + 1) * (r + 2) / 2 <= idx) r++;
  else if (r * (r + 1) / 2 > idx) r--;
  i = r;
  j = idx - r * (r + 1) / 2;
}


// wavefronts of the element loop: every wavefront assembles elements w, w + nw, ...
// so that its lane constants (shape-function factors, LDS positions of M and of
// the K_e entries) are formed once
constexpr int ELAST_WAVES = 8192;

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
  __shared__ double s_b[EPB][8][24];   // [gp][3a + p]: dN_a/dx_p
  __shared__ double s_w[EPB][8];       // |J| of each Gauss point (weights are 1)
  __shared__ double s_m[EPB][MBLK + 1];  // the lower 3x3 blocks of M, then a zero
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t w0 = (int64_t)blockIdx.x * EPB + wv, nw = (int64_t)gridDim.x * EPB;
  if (w0 >= ne) return;  // wave-uniform; no block-wide barrier follows
  double *sm_m = s_m[wv];
  if (lane == 0) sm_m[MBLK] = 0.0;

  // --- lane constants
  // lane (g, a): Gauss point g (order of linhex.cpp:598-606: xi outer, eta, zeta
  // inner), node a
  const int g = lane >> 3, a = lane & 7;
  const double xi = (g & 4) ? GPT : -GPT, eta = (g & 2) ? GPT : -GPT, zeta = (g & 1) ? GPT : -GPT;
  // natural derivatives of node b (signs sx, sy, sz, GEFEL/basefun.cpp:2317-2375):
  // dN/dxi = sx fyz, dN/deta = sy fxz, dN/dzeta = sz fxy with
  // fyz = (1 + sy eta)(1 + sz zeta)/8 etc.; index 2 (s1 < 0) + (s2 < 0)
  double fyz[4], fxz[4], fxy[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double s1 = (k & 2) ? -1.0 : 1.0, s2 = (k & 1) ? -1.0 : 1.0;
    fyz[k] = (1.0 + s1 * eta) * (1.0 + s2 * zeta) * 0.125;
    fxz[k] = (1.0 + s1 * xi) * (1.0 + s2 * zeta) * 0.125;
    fxy[k] = (1.0 + s1 * xi) * (1.0 + s2 * eta) * 0.125;
  }
  double nxa, nya, nza;  // the lane's own node a
  hex_nat1(xi, eta, zeta, a, nxa, nya, nza);
  const int li = lane & 15, lk = lane >> 4;
  // LDS positions of the MFMA results (tile t, value v): row i = 16 ti + 4 v + lk,
  // column j = 16 tj + li; -1 where nothing is stored
  int mp[3][4], mp2[4];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const int ti = (t == 2) ? 1 : 0, tj = (t == 0) ? 0 : 1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = 16 * ti + 4 * v + lk, j = 16 * tj + li;
      int p1 = -1, p2 = -1;
      if (i < 24 && j < 24) {
        if (t == 1) {
          p1 = mpos(j, i);                   // tile (0,1) lies above the diagonal: its mirror
          if (i / 3 == j / 3) p2 = mpos(i, j);  // block (5, 5) straddles the tiles
        } else if (i / 3 >= j / 3) {
          p1 = mpos(i, j);
        }
      }
      mp[t][v] = p1;
      if (t == 1) mp2[v] = p2;
    }
  }
  // K_e entry idx = lane + 64 r of the lower triangle, (I, J) = ((a, p), (b, q)):
  // c(r) M[o1] + gm (M[o2] + M[o3]) with c = dd0 on the diagonal blocks' p = q
  // entries (o3 then the third term), else c = dd1 and o3 the zero
  int o1[5], o2[5], o3[5];
  unsigned dmask = 0;
#pragma unroll
  for (int r = 0; r < 5; r++) {
    const int idx = lane + 64 * r;
    int I = 0, J = 0;
    if (idx < 300) tri_ij(idx, I, J);
    const int ea = I / 3, p = I - 3 * ea, eb = J / 3, q = J - 3 * eb;
    const int blk = 9 * (ea * (ea + 1) / 2 + eb);
    if (p != q) {
      o1[r] = blk + 3 * p + q;
      o2[r] = blk + 3 * q + p;
      o3[r] = MBLK;
    } else {
      const int r1 = p == 0 ? 1 : 0, r2 = p == 2 ? 1 : 2;
      o1[r] = blk + 4 * p;
      o2[r] = blk + 4 * r1;
      o3[r] = blk + 4 * r2;
      dmask |= 1u << r;
    }
  }

  // coordinates of the first element (lanes 0..23: coordinate lane / 8 of node lane % 8)
  double xv = 0.0;
  if (lane < 24) xv = xyz[3 * (int64_t)conn[w0 * 8 + (lane & 7)] + (lane >> 3)];
  for (int64_t e = w0; e < ne; e += nw) {
    if (lane < 24) s_xyz[wv][lane >> 3][lane & 7] = xv;
    wave_lds_sync();
    const int64_t en = e + nw;  // the next element's coordinates, in flight meanwhile
    if (lane < 24 && en < ne) xv = xyz[3 * (int64_t)conn[en * 8 + (lane & 7)] + (lane >> 3)];
    // 1. Jacobian of Gauss point g and the derivatives of node a
    double d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0, d7 = 0, d8 = 0, d9 = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool px = (b == 0 || b == 3 || b == 4 || b == 7), py = (b == 0 || b == 1 || b == 4 || b == 5),
                 pz = b < 4;
      const double nx = (px ? 1.0 : -1.0) * fyz[(py ? 0 : 2) + (pz ? 0 : 1)];
      const double ny = (py ? 1.0 : -1.0) * fxz[(px ? 0 : 2) + (pz ? 0 : 1)];
      const double nz = (pz ? 1.0 : -1.0) * fxy[(px ? 0 : 2) + (py ? 0 : 1)];
      const double x = s_xyz[wv][0][b], y = s_xyz[wv][1][b], z = s_xyz[wv][2][b];
      d1 += nx * x; d2 += ny * x; d3 += nz * x;
      d4 += nx * y; d5 += ny * y; d6 += nz * y;
      d7 += nx * z; d8 += ny * z; d9 += nz * z;
    }
    const double jac = d1 * d5 * d9 + d4 * d8 * d3 + d7 * d2 * d6 - d7 * d5 * d3 -
                       d8 * d6 * d1 - d9 * d4 * d2;
    const double ij = 1.0 / jac;
    s_b[wv][g][3 * a + 0] = (nxa * (d5 * d9 - d8 * d6) + nya * (d7 * d6 - d9 * d4) + nza * (d4 * d8 - d7 * d5)) * ij;
    s_b[wv][g][3 * a + 1] = (nxa * (d8 * d3 - d9 * d2) + nya * (d1 * d9 - d7 * d3) + nza * (d7 * d2 - d8 * d1)) * ij;
    s_b[wv][g][3 * a + 2] = (nxa * (d2 * d6 - d5 * d3) + nya * (d4 * d3 - d6 * d1) + nza * (d1 * d5 - d4 * d2)) * ij;
    if (a == 0) s_w[wv][g] = fabs(jac);
    wave_lds_sync();
    // material (elastisomat::matstiff_spacestr)
    const double em = E[e], nu = NU[e];
    const double gm = em / 2.0 / (1.0 + nu);
    const double sm = em / (1.0 + nu) / (1.0 - 2.0 * nu);
    const double dd0 = sm * (1.0 - nu), dd1 = sm * nu;

    // 2. M = sum_g |J_g| b_g b_g^T on MFMA. Operand lane layout of
    //    v_mfma_f64_16x16x4f64: A[i][k]: i = lane % 16, k = lane / 16;
    //    B[k][j]: j = lane % 16, k = lane / 16; D[i][j] (4 values per lane):
    //    i = 4 * v + lane / 16, j = lane % 16. Here k is the Gauss point.
    double4_t acc[3];
#pragma unroll
    for (int t = 0; t < 3; t++) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int st = 0; st < 2; st++) {
      const int gk = 4 * st + lk;
      const double w = s_w[wv][gk];
      const double b0 = s_b[wv][gk][li];                      // column li
      const double b1 = li < 8 ? s_b[wv][gk][16 + li] : 0.0;  // column 16 + li (< 24)
      const double a0 = b0 * w, a1 = b1 * w;
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0], 0, 0, 0);  // (0,0)
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[1], 0, 0, 0);  // (0,1)
      acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[2], 0, 0, 0);  // (1,1)
    }
#pragma unroll
    for (int t = 0; t < 3; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        if (mp[t][v] >= 0) sm_m[mp[t][v]] = acc[t][v];
        if (t == 1 && mp2[v] >= 0) sm_m[mp2[v]] = acc[t][v];
      }
    wave_lds_sync();
    // 3. the lower triangle of K_e, entry idx = i (i + 1) / 2 + j at kb[idx]
    double *kb = kbuf + e * 300;
#pragma unroll
    for (int r = 0; r < 5; r++) {
      if (r < 4 || lane < 300 - 256) {
        const double c = ((dmask >> r) & 1u) ? dd0 : dd1;
        kb[lane + 64 * r] = c * sm_m[o1[r]] + gm * (sm_m[o2[r]] + sm_m[o3[r]]);
      }
    }
    if (fbuf != nullptr && pidx[e] >= 0 && lane < 24) {  // pidx[e] is wave-uniform
      // the lift -K_e u_p of the rows with a code number >= 1, from the same entries
      const int ci = cn[e * 24 + lane];
      if (ci >= 1) {
        double f = 0.0;
        for (int j = 0; j < 24; j++) {
          const int cj = cn[e * 24 + j];
          if (cj < 0) {
            const double kij = lane >= j ? ke_entry(sm_m, lane, j, dd0, dd1, gm)
                                         : ke_entry(sm_m, j, lane, dd0, dd1, gm);
            f += kij * pv[-cj - 1];
          }
        }
        fbuf[(int64_t)pidx[e] * 24 + lane] = -f;
      }
    }
    wave_lds_sync();  // this element's LDS reads before the next one's writes
  }
}

hipError_t assemble_hex8_elast(int64_t ne, const int *conn, const int *cn, const double *xyz,
                               const double *ma, const double *mb, double *kbuf, const int *pidx,
                               const double *pv, double *fbuf, hipStream_t s)
{
  if (ne <= 0) return hipSuccess;
  const int64_t waves = std::min<int64_t>(ne, ELAST_WAVES);
  const unsigned nb = (unsigned)((waves + EPB - 1) / EPB);
  hipLaunchKernelGGL(k_hex8_elast, dim3(nb), dim3(TPB), 0, s, ne, conn, cn, xyz, ma, mb, kbuf,
                     pidx, pv, fbuf);
  return hipGetLastError();
}

}  // namespace ssf
