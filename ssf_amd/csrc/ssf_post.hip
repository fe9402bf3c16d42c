// Post-solve element pass on the device: integration-point strains and stresses
// (gradients and fluxes for heat) and the internal forces (internal fluxes)
// assembled into the global vector.
//
// Reference path, per element in element order:
//   eldispl                 MEFEL/SRC/globmat.cpp:1207-1311 (select_nodalval,
//                           TRFEL/SRC/globmatt.cpp:342-363, for heat)
//   strains                 linhex::gl_ip_strains MEFEL/SRC/linhex.cpp:1110-1145,
//                           planeelemlq::gl_ip_strains MEFEL/SRC/plelemlq.cpp:1173-1205,
//                           linhext::intpointgrad TRFEL/SRC/linhext.cpp:432-498
//   stresses                elastisomat::nlstresses MEFEL/SRC/elastisomat.cpp:627-651,
//                           transmat::computenlfluxes TRFEL/SRC/transmat.cpp:1123-1148
//   element forces          linhex::elem_integration MEFEL/SRC/linhex.cpp:2157-2191,
//                           planeelemlq::elem_integration MEFEL/SRC/plelemlq.cpp:2307-2359,
//                           linhext::internal_fluxes TRFEL/SRC/linhext.cpp:1086-1146
//   assembly                loc_internal_forces MEFEL/SRC/globmat.cpp:506-650 /
//                           internal_fluxes TRFEL/SRC/elemswitcht.cpp:1845, locglob
//                           GEFEL/matrix.cpp:3500
//
// Layout of the work: one lane per integration point, NGP lanes per element
// (8 for the hexahedra, 4 for the quadrilateral), elements contiguous within a
// wavefront. Lane g loads node g (coordinates, and the nodal values picked by the
// code numbers) into LDS; every lane then evaluates its own point from the
// element's LDS copy and writes its strain / stress records, which are contiguous
// over the wavefront. The element forces are summed over the points in point
// order through LDS and written per element; a second kernel adds them per
// unknown in element order, from a table of (element, local DOF) entries sorted
// by code number (a stable radix sort of the element code numbers, built once).
//
// Floating-point contraction is off in the arithmetic below and every sum runs
// in the reference's order (the dense mxv / mtxv loops minus their exact-zero
// terms), so the results equal the reference's x86-64 arithmetic bit for bit,
// up to the sign of exact zeros.
#include <hip/hip_ext.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "ssf_device.hpp"

namespace ssf {
namespace {

template <int KIND>
struct Shape;
template <>
struct Shape<SSF_ELEM_HEX8_ELAST> {
  static constexpr int NNE = 8, NGP = 8, DPN = 3, NC = 6;
};
template <>
struct Shape<SSF_ELEM_HEX8_HEAT> {
  static constexpr int NNE = 8, NGP = 8, DPN = 1, NC = 3;
};
template <>
struct Shape<SSF_ELEM_QUAD4_PSTRESS> {
  static constexpr int NNE = 4, NGP = 4, DPN = 2, NC = 4;
};

constexpr double kGP = 0.577350269189626;  // GEFEL/intp.cpp:26-27, weights 1

// The NGP lanes of an element are consecutive lanes of one wavefront and only
// they touch the element's LDS records, so a wavefront-local LDS sync orders them.
__device__ __forceinline__ void wave_sync() { wave_lds_sync(); }

// Natural derivatives of the trilinear hexahedron (GEFEL/basefun.cpp:2317-2375),
// e.g. dN_0/dxi = (1 + eta)(1 + zeta) / 8 for node 0 at (+,+,+). At a Gauss point
// (+-g, +-g, +-g) each factor is 1 + g or 1 - g, so the product / 8 is one of three
// constants P, M, Q, folded at compile time with the same IEEE operations
// ((1+g)(1+g)/8 and so on; the / (-8.0) forms are their exact negatives): the
// selected constant equals the reference's runtime value bit for bit.
constexpr double kP = (1.0 + kGP) * (1.0 + kGP) / 8.0;
constexpr double kM = (1.0 + kGP) * (1.0 - kGP) / 8.0;
constexpr double kQ = (1.0 - kGP) * (1.0 - kGP) / 8.0;

__device__ __forceinline__ double pick(bool p1, bool p2)
{
  return p1 ? (p2 ? kP : kM) : (p2 ? kM : kQ);
}

// node a sits at (sx, sy, sz) in {+1,-1}^3: (+,+,+), (-,+,+), (-,-,+), (+,-,+),
// then the same at zeta = -1; Gauss point g has xi > 0 iff g & 4, eta > 0 iff
// g & 2, zeta > 0 iff g & 1 (MEFEL/SRC/linhex.cpp:598-606 loop order)
__device__ __forceinline__ void hex_nat(int g, double *nx, double *ny, double *nz)
{
  const bool gx = g & 4, gy = g & 2, gz = g & 1;
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const bool px = (a == 0 || a == 3 || a == 4 || a == 7);
    const bool py = (a == 0 || a == 1 || a == 4 || a == 5);
    const bool pz = a < 4;
    const double vx = pick(py == gy, pz == gz), vy = pick(px == gx, pz == gz),
                 vz = pick(px == gx, py == gy);
    nx[a] = px ? vx : -vx;
    ny[a] = py ? vy : -vy;
    nz[a] = pz ? vz : -vz;
  }
}

// a / b rounded to nearest, given y = RN(1/b): Markstein's division method with
// FMA remainders. q = RN(a y) is within 2 ulp of a/b; one correction
// q + (a - b q) y leaves a faithful quotient, for which the remainder a - b q is
// exact and, by Markstein's theorem (y the correctly rounded reciprocal, q
// faithful), the second correction rounds to RN(a / b). Valid while a - b q does
// not underflow (|a| above ~2^-960), far below the values of element geometry.
__device__ __forceinline__ double div_rn(double a, double b, double y)
{
  double q = a * y;
  double r = __builtin_fma(-b, q, a);
  q = __builtin_fma(r, y, q);
  r = __builtin_fma(-b, q, a);
  return __builtin_fma(r, y, q);
}

// GEFEL/difcalc.cpp:487-571 derivatives_3d: natural -> global derivatives in
// place; returns the signed Jacobian. The 24 quotients by the Jacobian share one
// IEEE reciprocal (div_rn). A degenerate element (zero Jacobian) gives NaN
// derivatives where the reference's division gives infinities.
__device__ __forceinline__ double hex_global(double *dx, double *dy, double *dz,
                                             const double *x, const double *y,
                                             const double *z)
{
#pragma clang fp contract(off)
  double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0, d5 = 0.0, d6 = 0.0, d7 = 0.0, d8 = 0.0,
         d9 = 0.0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    d1 += dx[i] * x[i];
    d2 += dy[i] * x[i];
    d3 += dz[i] * x[i];
    d4 += dx[i] * y[i];
    d5 += dy[i] * y[i];
    d6 += dz[i] * y[i];
    d7 += dx[i] * z[i];
    d8 += dy[i] * z[i];
    d9 += dz[i] * z[i];
  }
  const double jac = d1 * d5 * d9 + d4 * d8 * d3 + d7 * d2 * d6 - d7 * d5 * d3 - d8 * d6 * d1 -
                     d9 * d4 * d2;
  const double c00 = d5 * d9 - d8 * d6, c01 = d7 * d6 - d9 * d4, c02 = d4 * d8 - d7 * d5;
  const double c10 = d8 * d3 - d9 * d2, c11 = d1 * d9 - d7 * d3, c12 = d7 * d2 - d8 * d1;
  const double c20 = d2 * d6 - d5 * d3, c21 = d4 * d3 - d6 * d1, c22 = d1 * d5 - d4 * d2;
  const double rj = 1.0 / jac;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const double ddx = dx[i], ddy = dy[i], ddz = dz[i];
    dx[i] = div_rn(ddx * c00 + ddy * c01 + ddz * c02, jac, rj);
    dy[i] = div_rn(ddx * c10 + ddy * c11 + ddz * c12, jac, rj);
    dz[i] = div_rn(ddx * c20 + ddy * c21 + ddz * c22, jac, rj);
  }
  return jac;
}

// One integration point of an element. eval() writes the point's NC strain and
// stress components and keeps what its force contributions need; force(c, a) is
// the contribution to DOF c of node a, already times the point's Jacobian factor.
template <int KIND>
struct Point;

// hex8 elasticity: rows of the strain-displacement matrix (linhex.cpp:167-183)
// eps_x, eps_y, eps_z, gamma_yz, gamma_xz, gamma_xy; D of matstiff_spacestr
// (elastisomat.cpp:395-410); forces B^T sigma jac (signed, linhex.cpp:2181-2184)
template <>
struct Point<SSF_ELEM_HEX8_ELAST> {
  double bx[8], by[8], bz[8];
  double s0, s1, s2, s3, s4, s5, w;

  __device__ void eval(int g, const double *x, const double *y, const double *z,
                       const double *r, double em, double nu, double th, double *eps,
                       double *sig)
  {
#pragma clang fp contract(off)
    (void)th;
    hex_nat(g, bx, by, bz);
    w = hex_global(bx, by, bz, x, y, z);
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) e0 += r[3 * a] * bx[a];
#pragma unroll
    for (int a = 0; a < 8; a++) e1 += r[3 * a + 1] * by[a];
#pragma unroll
    for (int a = 0; a < 8; a++) e2 += r[3 * a + 2] * bz[a];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      e3 += r[3 * a + 1] * bz[a];
      e3 += r[3 * a + 2] * by[a];
    }
#pragma unroll
    for (int a = 0; a < 8; a++) {
      e4 += r[3 * a] * bz[a];
      e4 += r[3 * a + 2] * bx[a];
    }
#pragma unroll
    for (int a = 0; a < 8; a++) {
      e5 += r[3 * a] * by[a];
      e5 += r[3 * a + 1] * bx[a];
    }
    const double gm = em / 2.0 / (1.0 + nu);
    const double s = em / (1.0 + nu) / (1.0 - 2.0 * nu);
    const double d00 = s * (1.0 - nu), d01 = s * nu;
    s0 = 0.0;
    s1 = 0.0;
    s2 = 0.0;
    s0 += e0 * d00; s0 += e1 * d01; s0 += e2 * d01;
    s1 += e0 * d01; s1 += e1 * d00; s1 += e2 * d01;
    s2 += e0 * d01; s2 += e1 * d01; s2 += e2 * d00;
    s3 = e3 * gm;
    s4 = e4 * gm;
    s5 = e5 * gm;
    eps[0] = e0; eps[1] = e1; eps[2] = e2; eps[3] = e3; eps[4] = e4; eps[5] = e5;
    sig[0] = s0; sig[1] = s1; sig[2] = s2; sig[3] = s3; sig[4] = s4; sig[5] = s5;
  }
  // column 3a + c of B^T sigma: the nonzero rows of that column in row order
  __device__ double force(int c, int a) const
  {
#pragma clang fp contract(off)
    double q = 0.0;
    if (c == 0) { q += s0 * bx[a]; q += s4 * bz[a]; q += s5 * by[a]; }
    else if (c == 1) { q += s1 * by[a]; q += s3 * bz[a]; q += s5 * bx[a]; }
    else { q += s2 * bz[a]; q += s3 * by[a]; q += s4 * bx[a]; }
    return q * w;
  }
};

// hex8 heat: gradient = grad N t, flux q = -k grad (isotrmat), internal fluxes
// with |J| (linhext.cpp:1124-1132)
template <>
struct Point<SSF_ELEM_HEX8_HEAT> {
  double bx[8], by[8], bz[8];
  double q0, q1, q2, w;

  __device__ void eval(int g, const double *x, const double *y, const double *z,
                       const double *r, double kc, double nu, double th, double *eps,
                       double *sig)
  {
#pragma clang fp contract(off)
    (void)nu;
    (void)th;
    hex_nat(g, bx, by, bz);
    const double jac = hex_global(bx, by, bz, x, y, z);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) g0 += r[a] * bx[a];
#pragma unroll
    for (int a = 0; a < 8; a++) g1 += r[a] * by[a];
#pragma unroll
    for (int a = 0; a < 8; a++) g2 += r[a] * bz[a];
    q0 = 0.0 - kc * g0;
    q1 = 0.0 - kc * g1;
    q2 = 0.0 - kc * g2;
    eps[0] = g0; eps[1] = g1; eps[2] = g2;
    sig[0] = q0; sig[1] = q1; sig[2] = q2;
    w = jac < 0.0 ? fabs(jac) : jac;
  }
  __device__ double force(int c, int a) const
  {
#pragma clang fp contract(off)
    (void)c;
    double q = 0.0;
    q += q0 * bx[a]; q += q1 * by[a]; q += q2 * bz[a];
    return q * w;
  }
};

// quad4 plane stress: rows eps_x, eps_y, gamma_xy (plelemlq.cpp:294-314); the
// stored strain gains eps_z = -nu/(1-nu)(eps_x+eps_y) and the stress a zero
// fourth component (elastisomat.cpp:640-643); thickness interpolated at the
// point (bf_lin_4_2d, GEFEL/basefun.cpp:609-615); forces B^T sigma jac thick
template <>
struct Point<SSF_ELEM_QUAD4_PSTRESS> {
  double bx[4], by[4];
  double s0, s1, s2, w;

  __device__ void eval(int g, const double *x, const double *y, const double *z,
                       const double *r, double em, double nu, double th, double *eps,
                       double *sig)
  {
#pragma clang fp contract(off)
    (void)z;
    const double xi = (g & 2) ? kGP : -kGP, eta = (g & 1) ? kGP : -kGP;
    double bf[4];
    bf[0] = 0.25 * (1.0 + xi) * (1.0 + eta);
    bf[1] = 0.25 * (1.0 - xi) * (1.0 + eta);
    bf[2] = 0.25 * (1.0 - xi) * (1.0 - eta);
    bf[3] = 0.25 * (1.0 + xi) * (1.0 - eta);
    double thick = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++) thick += bf[a] * th;
    double nx[4], ny[4];
    nx[0] = 0.25 * (1.0 + eta); nx[1] = 0.25 * (-1.0 - eta);
    nx[2] = 0.25 * (-1.0 + eta); nx[3] = 0.25 * (1.0 - eta);
    ny[0] = 0.25 * (1.0 + xi); ny[1] = 0.25 * (1.0 - xi);
    ny[2] = 0.25 * (-1.0 + xi); ny[3] = 0.25 * (-1.0 - xi);
    double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      d1 += nx[a] * x[a];
      d2 += ny[a] * x[a];
      d3 += nx[a] * y[a];
      d4 += ny[a] * y[a];
    }
    const double jac = d1 * d4 - d2 * d3;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      bx[a] = (nx[a] * d4 - ny[a] * d3) / jac;
      by[a] = (ny[a] * d1 - nx[a] * d2) / jac;
    }
    double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++) e0 += r[2 * a] * bx[a];
#pragma unroll
    for (int a = 0; a < 4; a++) e1 += r[2 * a + 1] * by[a];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      e2 += r[2 * a] * by[a];
      e2 += r[2 * a + 1] * bx[a];
    }
    const double c = em / (1.0 - nu * nu);
    const double cn = c * nu, d22 = em / 2.0 / (1.0 + nu);
    s0 = 0.0;
    s1 = 0.0;
    s0 += e0 * c; s0 += e1 * cn;
    s1 += e0 * cn; s1 += e1 * c;
    s2 = e2 * d22;
    eps[0] = e0; eps[1] = e1; eps[2] = e2; eps[3] = -nu / (1.0 - nu) * (e0 + e1);
    sig[0] = s0; sig[1] = s1; sig[2] = s2; sig[3] = 0.0;
    w = jac * thick;
  }
  __device__ double force(int c, int a) const
  {
#pragma clang fp contract(off)
    double q = 0.0;
    if (c == 0) { q += s0 * bx[a]; q += s2 * by[a]; }
    else { q += s1 * by[a]; q += s2 * bx[a]; }
    return q * w;
  }
};

// Element forces are kept per element in component-major order,
// fe[e][c][a] = force on DOF c of node a (local DOF a * DPN + c).
template <int KIND>
__global__ __launch_bounds__(TPB) void k_post(int64_t ne, const int *__restrict__ conn,
                                              const int *__restrict__ cn,
                                              const double *__restrict__ xyz,
                                              const double *__restrict__ ma,
                                              const double *__restrict__ mb,
                                              const double *__restrict__ th,
                                              const double *__restrict__ lhs,
                                              const double *__restrict__ pv,
                                              double *__restrict__ eps_out,
                                              double *__restrict__ sig_out,
                                              double *__restrict__ fe)
{
  using S = Shape<KIND>;
  constexpr int NNE = S::NNE, NGP = S::NGP, DPN = S::DPN, NC = S::NC, ND = NNE * DPN;
  constexpr int EPB = TPB / NGP;
  static_assert(NNE == NGP && 64 % NGP == 0, "one node per lane, elements within a wavefront");
  __shared__ double s_x[EPB][NNE], s_y[EPB][NNE], s_z[EPB][NNE];
  __shared__ double s_r[EPB][ND];
  __shared__ double s_c[EPB][NGP][NNE + 1];
  const int le = threadIdx.x / NGP, g = threadIdx.x % NGP;
  const int64_t e = (int64_t)xcd_block(blockIdx.x, gridDim.x) * EPB + le;
  const bool active = e < ne;
  if (active) {
    const int64_t nd = conn[e * NNE + g];
    s_x[le][g] = xyz[3 * nd];
    s_y[le][g] = xyz[3 * nd + 1];
    s_z[le][g] = xyz[3 * nd + 2];
#pragma unroll
    for (int c = 0; c < DPN; c++) {
      const int k = cn[e * ND + g * DPN + c];
      s_r[le][g * DPN + c] = k > 0 ? lhs[k - 1] : (k < 0 ? pv[-k - 1] : 0.0);
    }
  }
  wave_sync();
  Point<KIND> pt;
  if (active) {
    double eps[NC], sig[NC];
    pt.eval(g, s_x[le], s_y[le], s_z[le], s_r[le], ma[e],
            KIND == SSF_ELEM_HEX8_HEAT ? 0.0 : mb[e], th ? th[e] : 1.0, eps, sig);
    const int64_t o = (e * NGP + g) * NC;
    if (eps_out) {
#pragma unroll
      for (int c = 0; c < NC; c++) __builtin_nontemporal_store(eps[c], eps_out + o + c);
    }
    if (sig_out) {
#pragma unroll
      for (int c = 0; c < NC; c++) __builtin_nontemporal_store(sig[c], sig_out + o + c);
    }
  }
  if (!fe) return;
  // element forces, one DOF component per pass: f = contr_0 + 0, then
  // contr_g + f in point order (elem_integration's addv)
#pragma unroll
  for (int c = 0; c < DPN; c++) {
    if (active) {
#pragma unroll
      for (int a = 0; a < NNE; a++) s_c[le][g][a] = pt.force(c, a);
    }
    wave_sync();
    if (active) {
      double f = 0.0;
#pragma unroll
      for (int h = 0; h < NGP; h++) f = s_c[le][h][g] + f;
      fe[e * ND + c * NNE + g] = f;
    }
    wave_sync();
  }
}

// Element volume loads (hex8 elasticity: loadel::volumeload MEFEL/SRC/loadel.cpp:438
// with linhex::load_matrix MEFEL/SRC/linhex.cpp:969, signed Jacobian) and heat
// sources (linhext::quantity_source_vector TRFEL/SRC/linhext.cpp:1039, |J|), one
// thread per element: lm = sum over the points of N^T N jac (the scalar 8x8
// block; the 3-row N of elasticity repeats it per component and adds exact zeros
// elsewhere), nf = lm nodval in column order. Written to fe in the component-major
// order of the force entries.
template <int KIND>
__global__ __launch_bounds__(TPB) void k_vload(int64_t ne, const int *__restrict__ conn,
                                               const double *__restrict__ xyz,
                                               const double *__restrict__ nodval,
                                               double *__restrict__ fe)
{
#pragma clang fp contract(off)
  constexpr int DPN = KIND == SSF_ELEM_HEX8_ELAST ? 3 : 1, ND = 8 * DPN;
  const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (e >= ne) return;
  double x[8], y[8], z[8];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int64_t nd = conn[e * 8 + a];
    x[a] = xyz[3 * nd];
    y[a] = xyz[3 * nd + 1];
    z[a] = xyz[3 * nd + 2];
  }
  double lm[8][8];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) lm[a][b] = 0.0;
  for (int g = 0; g < 8; g++) {
    const double xi = (g & 4) ? kGP : -kGP, eta = (g & 2) ? kGP : -kGP, zeta = (g & 1) ? kGP : -kGP;
    double nx[8], ny[8], nz[8];
    hex_nat(g, nx, ny, nz);
    double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0, d5 = 0.0, d6 = 0.0, d7 = 0.0, d8 = 0.0,
           d9 = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      d1 += nx[i] * x[i];
      d2 += ny[i] * x[i];
      d3 += nz[i] * x[i];
      d4 += nx[i] * y[i];
      d5 += ny[i] * y[i];
      d6 += nz[i] * y[i];
      d7 += nx[i] * z[i];
      d8 += ny[i] * z[i];
      d9 += nz[i] * z[i];
    }
    double jac = d1 * d5 * d9 + d4 * d8 * d3 + d7 * d2 * d6 - d7 * d5 * d3 - d8 * d6 * d1 -
                 d9 * d4 * d2;
    if (KIND == SSF_ELEM_HEX8_HEAT && jac < 0.0) jac = fabs(jac);
    double bf[8];
    bf[0] = (1.0 + xi) * (1.0 + eta) * (1.0 + zeta) / 8.0;
    bf[1] = (1.0 - xi) * (1.0 + eta) * (1.0 + zeta) / 8.0;
    bf[2] = (1.0 - xi) * (1.0 - eta) * (1.0 + zeta) / 8.0;
    bf[3] = (1.0 + xi) * (1.0 - eta) * (1.0 + zeta) / 8.0;
    bf[4] = (1.0 + xi) * (1.0 + eta) * (1.0 - zeta) / 8.0;
    bf[5] = (1.0 - xi) * (1.0 + eta) * (1.0 - zeta) / 8.0;
    bf[6] = (1.0 - xi) * (1.0 - eta) * (1.0 - zeta) / 8.0;
    bf[7] = (1.0 + xi) * (1.0 - eta) * (1.0 - zeta) / 8.0;
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int b = a; b < 8; b++) {
        const double v = (bf[a] * bf[b]) * jac;
        lm[a][b] += v;
        if (b != a) lm[b][a] += v;
      }
  }
  const double *nv = nodval + e * ND;
#pragma unroll
  for (int c = 0; c < DPN; c++)
#pragma unroll
    for (int a = 0; a < 8; a++) {
      double sacc = 0.0;
#pragma unroll
      for (int b = 0; b < 8; b++) sacc += lm[a][b] * nv[b * DPN + c];
      fe[e * ND + c * 8 + a] = KIND == SSF_ELEM_HEX8_ELAST ? sacc : 0.0 + sacc;
    }
}

// ------------------------------------------------------------ surface loads
// linhex::node_forces_surf (MEFEL/SRC/linhex.cpp:2543-3047), one thread per load
// record (element, face flags ls[6], nodal values nv[72] = 6 faces x 4 nodes x 3).
// Per loaded face: M_ab = sum over 2x2 points of (N_a N_b) |J_face| for the face's
// four nodes (the nonzero entries of nnj's 24x24 matrix, the same for the three
// components), the nodal values in global components (ls 1) or rotated from the
// face's axes (ls 2, locglob_nodeval), v = M av summed in ascending DOF order (the
// order of mxv's nonzero terms) and nf = v + nf. Other entries of the reference's
// products are exact zeros and change no sum.
struct SurfFace {
  int fix, outer, inner;
  double val;
  int node[4];  // jac2d_3d's face order, also the order of the nodal values
  int asc[4];   // positions in node[] by ascending element node
  int g[4];     // g1 = X[g[1]] - X[g[0]], g2 = X[g[3]] - X[g[2]]
};
__constant__ SurfFace kSurf[6] = {
    {0, 1, 2, 1.0, {0, 3, 7, 4}, {0, 1, 3, 2}, {7, 4, 7, 3}},
    {1, 0, 2, 1.0, {1, 0, 4, 5}, {1, 0, 2, 3}, {4, 5, 4, 0}},
    {0, 1, 2, -1.0, {2, 1, 5, 6}, {1, 0, 2, 3}, {5, 6, 5, 1}},
    {1, 0, 2, -1.0, {3, 2, 6, 7}, {1, 0, 2, 3}, {6, 7, 6, 2}},
    {2, 0, 1, 1.0, {0, 1, 2, 3}, {0, 1, 2, 3}, {2, 3, 2, 1}},
    {2, 0, 1, -1.0, {4, 7, 6, 5}, {0, 3, 2, 1}, {6, 5, 6, 7}},
};

__device__ double surf_bf(int a, double xi, double eta, double zeta)
{
#pragma clang fp contract(off)
  switch (a) {  // bf_lin_hex_3d (GEFEL/basefun.cpp)
    case 0: return (1.0 + xi) * (1.0 + eta) * (1.0 + zeta) / 8.0;
    case 1: return (1.0 - xi) * (1.0 + eta) * (1.0 + zeta) / 8.0;
    case 2: return (1.0 - xi) * (1.0 - eta) * (1.0 + zeta) / 8.0;
    case 3: return (1.0 + xi) * (1.0 - eta) * (1.0 + zeta) / 8.0;
    case 4: return (1.0 + xi) * (1.0 + eta) * (1.0 - zeta) / 8.0;
    case 5: return (1.0 - xi) * (1.0 + eta) * (1.0 - zeta) / 8.0;
    case 6: return (1.0 - xi) * (1.0 - eta) * (1.0 - zeta) / 8.0;
    default: return (1.0 + xi) * (1.0 - eta) * (1.0 - zeta) / 8.0;
  }
}

__global__ __launch_bounds__(TPB) void k_sload(int64_t nrec, const int *__restrict__ eid,
                                               const int *__restrict__ ls,
                                               const double *__restrict__ nodvals,
                                               const int *__restrict__ conn,
                                               const double *__restrict__ xyz,
                                               double *__restrict__ fe)
{
#pragma clang fp contract(off)
  const int64_t r = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (r >= nrec) return;
  const int e = eid[r];
  double x[8], y[8], z[8], nf[8][3];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int nd = conn[e * 8 + a];
    x[a] = xyz[3 * nd];
    y[a] = xyz[3 * nd + 1];
    z[a] = xyz[3 * nd + 2];
    nf[a][0] = nf[a][1] = nf[a][2] = 0.0;
  }
  for (int f = 0; f < 6; f++) {
    const int lsf = ls[r * 6 + f];
    if (lsf <= 0) continue;
    const SurfFace &F = kSurf[f];
    double lx[4], ly[4], lz[4], M[4][4];
    for (int m = 0; m < 4; m++) {
      lx[m] = x[F.node[m]];
      ly[m] = y[F.node[m]];
      lz[m] = z[F.node[m]];
      for (int q = 0; q < 4; q++) M[m][q] = 0.0;
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        double c[3];
        c[F.fix] = F.val;
        c[F.outer] = i ? kGP : -kGP;
        c[F.inner] = j ? kGP : -kGP;
        // jac2d3d (GEFEL/difcalc.cpp:819-866), derivatives of basefun.cpp:626-649
        const double u = c[F.outer], w = c[F.inner];
        const double nu[4] = {0.25 * (1.0 + w), 0.25 * (-1.0 - w), 0.25 * (-1.0 + w),
                              0.25 * (1.0 - w)};
        const double nv[4] = {0.25 * (1.0 + u), 0.25 * (1.0 - u), 0.25 * (-1.0 + u),
                              0.25 * (-1.0 - u)};
        double dxu = 0.0, dyu = 0.0, dzu = 0.0, dxv = 0.0, dyv = 0.0, dzv = 0.0;
        for (int m = 0; m < 4; m++) {
          dxu += nu[m] * lx[m];
          dyu += nu[m] * ly[m];
          dzu += nu[m] * lz[m];
          dxv += nv[m] * lx[m];
          dyv += nv[m] * ly[m];
          dzv += nv[m] * lz[m];
        }
        const double g11 = dxu * dxu + dyu * dyu + dzu * dzu;
        const double g22 = dxv * dxv + dyv * dyv + dzv * dzv;
        const double g12 = dxu * dxv + dyu * dyv + dzu * dzv;
        const double jac = sqrt(g11 * g22 - g12 * g12) * 1.0 * 1.0;  // * w[i] * w[j]
        double N[4];
        for (int m = 0; m < 4; m++) N[m] = surf_bf(F.node[m], c[0], c[1], c[2]);
        for (int m = 0; m < 4; m++)
          for (int q = 0; q < 4; q++) M[m][q] += (N[m] * N[q]) * jac;
      }
    const double *fv = nodvals + r * 72 + 12 * f;
    double av[4][3];
    if (lsf == 2) {
      // locglob_nodeval (MEFEL/SRC/linhex.cpp:3049-3143)
      double g1[3] = {x[F.g[1]] - x[F.g[0]], y[F.g[1]] - y[F.g[0]], z[F.g[1]] - z[F.g[0]]};
      double g2[3] = {x[F.g[3]] - x[F.g[2]], y[F.g[3]] - y[F.g[2]], z[F.g[3]] - z[F.g[2]]};
      double norm = 0.0;
      for (int k = 0; k < 3; k++) norm += g1[k] * g1[k];
      norm = 1.0 / sqrt(norm);
      for (int k = 0; k < 3; k++) g1[k] = norm * g1[k];
      norm = 0.0;
      for (int k = 0; k < 3; k++) norm += g1[k] * g2[k];
      for (int k = 0; k < 3; k++) g2[k] = g2[k] - norm * g1[k];
      norm = 0.0;
      for (int k = 0; k < 3; k++) norm += g2[k] * g2[k];
      norm = 1.0 / sqrt(norm);
      for (int k = 0; k < 3; k++) g2[k] = norm * g2[k];
      const double g3[3] = {g1[1] * g2[2] - g1[2] * g2[1], g1[2] * g2[0] - g1[0] * g2[2],
                            g1[0] * g2[1] - g1[1] * g2[0]};
      for (int m = 0; m < 4; m++)
        for (int k = 0; k < 3; k++) {
          double sacc = 0.0;
          sacc += g1[k] * fv[3 * m];
          sacc += g2[k] * fv[3 * m + 1];
          sacc += g3[k] * fv[3 * m + 2];
          av[m][k] = sacc;
        }
    } else {
      for (int m = 0; m < 4; m++)
        for (int k = 0; k < 3; k++) av[m][k] = fv[3 * m + k];
    }
    for (int m = 0; m < 4; m++)
      for (int k = 0; k < 3; k++) {
        double v = 0.0;
        for (int q = 0; q < 4; q++) v += M[m][F.asc[q]] * av[F.asc[q]][k];
        nf[F.node[m]][k] = v + nf[F.node[m]][k];
      }
  }
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int k = 0; k < 3; k++) fe[r * 24 + k * 8 + a] = nf[a][k];
}

// code numbers of each record's element; records with an element id outside
// [0, ne) get no DOFs and raise *bad
__global__ void k_rec_cn(int64_t nrec, int64_t ne, const int *__restrict__ eid,
                         const int *__restrict__ cn, int *__restrict__ cnr, int *bad)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= nrec * 24) return;
  const int64_t r = k / 24;
  const int e = eid[r];
  const bool ok = e >= 0 && e < ne;
  if (!ok && k % 24 == 0) *bad = 1;
  cnr[k] = ok ? cn[(int64_t)e * 24 + k % 24] : 0;
}

// ifor[d] = sum over the element entries of unknown d, in element order
// (locglob into a cleared vector). The entry indices and values are loaded eight
// at a time so that their loads overlap; the sum keeps the entry order.
__global__ __launch_bounds__(TPB) void k_post_gather(int n, const int *__restrict__ start,
                                                     const int *__restrict__ src,
                                                     const double *__restrict__ fe,
                                                     double *__restrict__ ifor, int accumulate)
{
  const int d = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
  if (d >= n) return;
  const int k1 = start[d + 1];
  double s = accumulate ? ifor[d] : 0.0;
  for (int k = start[d]; k < k1; k += 8) {
    int idx[8];
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) idx[j] = k + j < k1 ? src[k + j] : -1;
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = idx[j] >= 0 ? fe[idx[j]] : 0.0;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (idx[j] >= 0) s = s + v[j];
  }
  ifor[d] = s;
}

// sort keys of the (element, local DOF) entries: unknown (code number - 1), or n
// for a fixed / prescribed DOF (sorted last, not summed); values: the entry's
// place in the component-major element forces
__global__ void k_post_keys(int64_t nent, int n, int nne, int dpn, const int *__restrict__ cn,
                            int *__restrict__ key, int *__restrict__ val)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k >= nent) return;
  const int c = cn[k];
  const int nd = nne * dpn;
  const int64_t e = k / nd;
  const int i = (int)(k - e * nd);
  key[k] = c > 0 ? c - 1 : n;
  val[k] = (int)(e * nd + (i % dpn) * nne + i / dpn);
}

// start[d] = first sorted entry with key >= d, for d = 0..n
__global__ void k_post_start(int64_t nent, int n, const int *__restrict__ skey,
                             int *__restrict__ start)
{
  const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= nent) return;
  const int kp = skey[p];
  const int kprev = p > 0 ? skey[p - 1] : -1;
  for (int d = kprev + 1; d <= kp && d <= n; d++) start[d] = (int)p;
  if (p == nent - 1)
    for (int d = kp + 1; d <= n; d++) start[d] = (int)nent;
}

__global__ void k_post_fill_int(int64_t m, int v, int *__restrict__ a)
{
  const int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x;
  if (k < m) a[k] = v;
}

inline unsigned blocks_for(int64_t m, int64_t per) { return (unsigned)((m + per - 1) / per); }

}  // namespace
}  // namespace ssf

struct ssf_elements {
  ssf_ctx *ctx = nullptr;
  int kind = 0;
  int64_t n = 0, nn = 0, ne = 0;
  int nne = 0, nd = 0, ngp = 0, nc = 0;
  int64_t maxpv = 0;  // largest prescribed-value index the code numbers refer to
  int *conn = nullptr, *cn = nullptr;
  double *xyz = nullptr, *ma = nullptr, *mb = nullptr, *th = nullptr;
  int *start = nullptr, *src = nullptr;  // per-unknown entry lists
  double *fe = nullptr;                  // element forces [ne][dpn][nne]
};

using namespace ssf;

void ssf::elements_shape(const ssf_elements *el, ssf_ctx **ctx, int *kind, int64_t *n)
{
  *ctx = el->ctx;
  *kind = el->kind;
  *n = el->n;
}

static void post_shape(int kind, int &nne, int &nd, int &ngp, int &nc)
{
  nne = kind == SSF_ELEM_QUAD4_PSTRESS ? 4 : 8;
  nd = kind == SSF_ELEM_HEX8_ELAST ? 24 : 8;
  ngp = kind == SSF_ELEM_QUAD4_PSTRESS ? 4 : 8;
  nc = kind == SSF_ELEM_HEX8_ELAST ? 6 : (kind == SSF_ELEM_HEX8_HEAT ? 3 : 4);
}

extern "C" int ssf_elements_destroy(ssf_elements *el)
{
  SSF_ENTER();
  if (!el) return SSF_OK;
  (void)hipSetDevice(el->ctx->device);
  hipStream_t s = el->ctx->stream;
  dfree(el->conn, s);
  dfree(el->cn, s);
  dfree(el->xyz, s);
  dfree(el->ma, s);
  dfree(el->mb, s);
  dfree(el->th, s);
  dfree(el->start, s);
  dfree(el->src, s);
  dfree(el->fe, s);
  delete el;
  return SSF_OK;
}

extern "C" int ssf_elements_create(ssf_ctx *ctx, int kind, int64_t n, int64_t nn,
                                   const double *xyz, int64_t ne, const int64_t *conn,
                                   const int64_t *cn, const double *mat_a, const double *mat_b,
                                   const double *thick, ssf_elements **out)
{
  SSF_ENTER();
  SSF_CHECK(out, SSF_EINVAL, "ssf_elements_create: NULL output");
  *out = nullptr;
  SSF_CHECK(ctx && xyz && conn && cn && mat_a, SSF_EINVAL, "ssf_elements_create: NULL argument");
  SSF_CHECK(kind >= 0 && kind <= 2, SSF_EINVAL, "ssf_elements_create: unknown element kind");
  SSF_CHECK(kind == SSF_ELEM_HEX8_HEAT || mat_b, SSF_EINVAL,
            "ssf_elements_create: Poisson ratios missing");
  int nne, nd, ngp, nc;
  post_shape(kind, nne, nd, ngp, nc);
  SSF_CHECK(n >= 0 && nn > 0 && ne > 0 && n < (1LL << 31) - 1 && nn < (1LL << 31) &&
                ne * nd < (1LL << 31) && ne * ngp * nc < (1LL << 40),
            SSF_EINVAL, "ssf_elements_create: bad sizes");
  int64_t maxpv = 0;
  std::vector<int> conn32(ne * nne), cn32(ne * nd);
  for (int64_t k = 0; k < ne * nne; k++) {
    SSF_CHECK(conn[k] >= 0 && conn[k] < nn, SSF_EINVAL, "ssf_elements_create: node id out of range");
    conn32[k] = (int)conn[k];
  }
  for (int64_t k = 0; k < ne * nd; k++) {
    SSF_CHECK(cn[k] <= n && cn[k] > -(1LL << 31), SSF_EINVAL,
              "ssf_elements_create: code number out of range");
    maxpv = std::max<int64_t>(maxpv, -cn[k]);
    cn32[k] = (int)cn[k];
  }
  SSF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  ssf_elements *el = new ssf_elements();
  el->ctx = ctx;
  el->kind = kind;
  el->n = n;
  el->nn = nn;
  el->ne = ne;
  el->nne = nne;
  el->nd = nd;
  el->ngp = ngp;
  el->nc = nc;
  el->maxpv = maxpv;
  const int64_t nent = ne * nd;
  int *key = nullptr, *skey = nullptr, *val = nullptr;
  void *tmp = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&el->conn, ne * nne, s));
  c(dalloc(&el->cn, nent, s));
  c(dalloc(&el->xyz, nn * 3, s));
  c(dalloc(&el->ma, ne, s));
  if (mat_b && kind != SSF_ELEM_HEX8_HEAT) c(dalloc(&el->mb, ne, s));
  if (thick && kind == SSF_ELEM_QUAD4_PSTRESS) c(dalloc(&el->th, ne, s));
  c(dalloc(&el->start, n + 1, s));
  c(dalloc(&el->src, nent, s));
  c(dalloc(&el->fe, nent, s));
  c(dalloc(&key, nent, s));
  c(dalloc(&skey, nent, s));
  c(dalloc(&val, nent, s));
  if (e == hipSuccess) {
    c(hipMemcpyAsync(el->conn, conn32.data(), ne * nne * sizeof(int), hipMemcpyHostToDevice, s));
    c(hipMemcpyAsync(el->cn, cn32.data(), nent * sizeof(int), hipMemcpyHostToDevice, s));
    c(hipMemcpyAsync(el->xyz, xyz, nn * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    c(hipMemcpyAsync(el->ma, mat_a, ne * sizeof(double), hipMemcpyHostToDevice, s));
    if (el->mb) c(hipMemcpyAsync(el->mb, mat_b, ne * sizeof(double), hipMemcpyHostToDevice, s));
    if (el->th) c(hipMemcpyAsync(el->th, thick, ne * sizeof(double), hipMemcpyHostToDevice, s));
  }
  // entries of each unknown in element order: stable sort by code number
  int bits = 1;
  while (bits < 31 && (1LL << bits) <= n) bits++;
  size_t tb = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_post_keys, dim3(blocks_for(nent, TPB)), dim3(TPB), 0, s, nent, (int)n,
                       nne, nd / nne, el->cn, key, val);
    c(hipGetLastError());
  }
  if (e == hipSuccess)
    c(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, val, el->src, (int)nent, 0, bits,
                                         s));
  if (e == hipSuccess) c(hipMallocAsync(&tmp, std::max<size_t>(tb, 1), s));
  if (e == hipSuccess)
    c(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, skey, val, el->src, (int)nent, 0, bits, s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_post_fill_int, dim3(blocks_for(n + 1, TPB)), dim3(TPB), 0, s, n + 1,
                       (int)nent, el->start);
    c(hipGetLastError());
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_post_start, dim3(blocks_for(nent, TPB)), dim3(TPB), 0, s, nent, (int)n,
                       skey, el->start);
    c(hipGetLastError());
  }
  c(hipStreamSynchronize(s));
  if (tmp) (void)hipFreeAsync(tmp, s);
  dfree(key, s);
  dfree(skey, s);
  dfree(val, s);
  if (e != hipSuccess) {
    ssf_elements_destroy(el);
    SSF_HIP(e);
  }
  *out = el;
  return SSF_OK;
}

extern "C" int ssf_elements_sizes(const ssf_elements *el, int32_t *ngp, int32_t *ncomp,
                                  int32_t *ndofe)
{
  SSF_ENTER();
  SSF_CHECK(el, SSF_EINVAL, "ssf_elements_sizes: NULL element set");
  if (ngp) *ngp = el->ngp;
  if (ncomp) *ncomp = el->nc;
  if (ndofe) *ndofe = el->nd;
  return SSF_OK;
}

static hipError_t post_launch(ssf_elements *el, const double *lhs, const double *pv,
                              double *eps, double *sig, double *ifor, hipEvent_t ev0,
                              hipEvent_t ev1)
{
  hipStream_t s = el->ctx->stream;
  double *fe = ifor ? el->fe : nullptr;
  const int epb = TPB / el->ngp;
  const dim3 grid(blocks_for(el->ne, epb)), block(TPB);
#define SSF_POST(K)                                                                          \
  do {                                                                                       \
    if (ev0)                                                                                 \
      hipExtLaunchKernelGGL((k_post<K>), grid, block, 0, s, ev0, ev1, 0, el->ne, el->conn,   \
                            el->cn, el->xyz, el->ma, el->mb, el->th, lhs, pv, eps, sig, fe); \
    else                                                                                     \
      hipLaunchKernelGGL((k_post<K>), grid, block, 0, s, el->ne, el->conn, el->cn, el->xyz,  \
                         el->ma, el->mb, el->th, lhs, pv, eps, sig, fe);                     \
  } while (0)
  if (el->kind == SSF_ELEM_HEX8_ELAST) SSF_POST(SSF_ELEM_HEX8_ELAST);
  else if (el->kind == SSF_ELEM_HEX8_HEAT) SSF_POST(SSF_ELEM_HEX8_HEAT);
  else SSF_POST(SSF_ELEM_QUAD4_PSTRESS);
#undef SSF_POST
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !ifor || el->n == 0) return e;
  hipLaunchKernelGGL(k_post_gather, dim3(blocks_for(el->n, TPB)), dim3(TPB), 0, s, (int)el->n,
                     el->start, el->src, el->fe, ifor, 0);
  return hipGetLastError();
}

extern "C" int ssf_elements_results_dev(ssf_elements *el, const double *lhs_dev,
                                        const double *pv_dev, int64_t npv, double *strain_dev,
                                        double *stress_dev, double *ifor_dev)
{
  SSF_ENTER();
  SSF_CHECK(el, SSF_EINVAL, "ssf_elements_results_dev: NULL element set");
  SSF_CHECK(lhs_dev || el->n == 0, SSF_EINVAL, "ssf_elements_results_dev: NULL solution");
  SSF_CHECK(npv >= el->maxpv && (pv_dev || el->maxpv == 0), SSF_EINVAL,
            "ssf_elements_results_dev: prescribed values missing for the code numbers");
  SSF_HIP(hipSetDevice(el->ctx->device));
  SSF_HIP(post_launch(el, lhs_dev, pv_dev, strain_dev, stress_dev, ifor_dev, nullptr, nullptr));
  return SSF_OK;
}

extern "C" int ssf_elements_results(ssf_elements *el, const double *lhs, const double *pv,
                                    int64_t npv, double *strain, double *stress, double *ifor)
{
  SSF_ENTER();
  SSF_CHECK(el, SSF_EINVAL, "ssf_elements_results: NULL element set");
  SSF_CHECK(lhs || el->n == 0, SSF_EINVAL, "ssf_elements_results: NULL solution");
  SSF_CHECK(npv >= el->maxpv && (pv || el->maxpv == 0), SSF_EINVAL,
            "ssf_elements_results: prescribed values missing for the code numbers");
  SSF_HIP(hipSetDevice(el->ctx->device));
  hipStream_t s = el->ctx->stream;
  const int64_t nip = el->ne * el->ngp * el->nc;
  double *d_lhs = nullptr, *d_pv = nullptr, *d_eps = nullptr, *d_sig = nullptr, *d_f = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&d_lhs, el->n, s));
  if (pv && npv) c(dalloc(&d_pv, npv, s));
  if (strain) c(dalloc(&d_eps, nip, s));
  if (stress) c(dalloc(&d_sig, nip, s));
  if (ifor) c(dalloc(&d_f, el->n, s));
  if (e == hipSuccess && el->n)
    c(hipMemcpyAsync(d_lhs, lhs, el->n * sizeof(double), hipMemcpyHostToDevice, s));
  if (e == hipSuccess && d_pv)
    c(hipMemcpyAsync(d_pv, pv, npv * sizeof(double), hipMemcpyHostToDevice, s));
  if (e == hipSuccess) c(post_launch(el, d_lhs, d_pv, d_eps, d_sig, d_f, nullptr, nullptr));
  if (e == hipSuccess && strain)
    c(hipMemcpyAsync(strain, d_eps, nip * sizeof(double), hipMemcpyDeviceToHost, s));
  if (e == hipSuccess && stress)
    c(hipMemcpyAsync(stress, d_sig, nip * sizeof(double), hipMemcpyDeviceToHost, s));
  if (e == hipSuccess && ifor && el->n)
    c(hipMemcpyAsync(ifor, d_f, el->n * sizeof(double), hipMemcpyDeviceToHost, s));
  c(hipStreamSynchronize(s));
  dfree(d_lhs, s);
  dfree(d_pv, s);
  dfree(d_eps, s);
  dfree(d_sig, s);
  dfree(d_f, s);
  SSF_HIP(e);
  return SSF_OK;
}

extern "C" int ssf_elements_bench(ssf_elements *el, const double *lhs_dev, const double *pv_dev,
                                  int64_t npv, double *strain_dev, double *stress_dev,
                                  double *ifor_dev, int reps, double *usec)
{
  SSF_ENTER();
  SSF_CHECK(el && usec && reps > 0, SSF_EINVAL, "ssf_elements_bench: bad argument");
  SSF_CHECK(lhs_dev || el->n == 0, SSF_EINVAL, "ssf_elements_bench: NULL solution");
  SSF_CHECK(npv >= el->maxpv && (pv_dev || el->maxpv == 0), SSF_EINVAL,
            "ssf_elements_bench: prescribed values missing for the code numbers");
  SSF_HIP(hipSetDevice(el->ctx->device));
  hipStream_t s = el->ctx->stream;
  hipEvent_t k0, k1, c0, c1;
  SSF_HIP(hipEventCreate(&k0));
  SSF_HIP(hipEventCreate(&k1));
  SSF_HIP(hipEventCreate(&c0));
  SSF_HIP(hipEventCreate(&c1));
  hipError_t e = post_launch(el, lhs_dev, pv_dev, strain_dev, stress_dev, ifor_dev, nullptr,
                             nullptr);  // warm-up
  double tk = 0.0, tc = 0.0;
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    if (e == hipSuccess) e = hipEventRecord(c0, s);
    if (e == hipSuccess) e = post_launch(el, lhs_dev, pv_dev, strain_dev, stress_dev, ifor_dev, k0, k1);
    if (e == hipSuccess) e = hipEventRecord(c1, s);
    if (e == hipSuccess) e = hipEventSynchronize(c1);
    float a = 0.f, b = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&a, k0, k1);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, c0, c1);
    tk += a;
    tc += b;
  }
  (void)hipEventDestroy(k0);
  (void)hipEventDestroy(k1);
  (void)hipEventDestroy(c0);
  (void)hipEventDestroy(c1);
  SSF_HIP(e);
  usec[0] = 1e3 * tk / reps;
  usec[1] = 1e3 * tc / reps;
  return SSF_OK;
}

static hipError_t vload_launch(ssf_elements *el, const double *nodval, double *rhs)
{
  hipStream_t s = el->ctx->stream;
  const dim3 grid(blocks_for(el->ne, TPB)), block(TPB);
  if (el->kind == SSF_ELEM_HEX8_ELAST)
    hipLaunchKernelGGL(k_vload<SSF_ELEM_HEX8_ELAST>, grid, block, 0, s, el->ne, el->conn, el->xyz,
                       nodval, el->fe);
  else
    hipLaunchKernelGGL(k_vload<SSF_ELEM_HEX8_HEAT>, grid, block, 0, s, el->ne, el->conn, el->xyz,
                       nodval, el->fe);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || el->n == 0) return e;
  hipLaunchKernelGGL(k_post_gather, dim3(blocks_for(el->n, TPB)), block, 0, s, (int)el->n,
                     el->start, el->src, el->fe, rhs, 1);
  return hipGetLastError();
}

// surface loads of nrec records added into rhs in record order: per-record forces,
// then a table of the records' (record, local DOF) entries sorted stably by code
// number, summed per unknown (locglob record by record, loadcase.cpp:572-597)
static int sload_launch(ssf_elements *el, int64_t nrec, const int *eid, const int *ls,
                        const double *nodvals, double *rhs)
{
  hipStream_t s = el->ctx->stream;
  const int64_t nent = nrec * 24, n = el->n;
  double *fe = nullptr;
  int *cnr = nullptr, *key = nullptr, *skey = nullptr, *val = nullptr, *src = nullptr,
      *start = nullptr, *bad = nullptr;
  void *tmp = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&fe, nent, s));
  c(dalloc(&cnr, nent, s));
  c(dalloc(&key, nent, s));
  c(dalloc(&skey, nent, s));
  c(dalloc(&val, nent, s));
  c(dalloc(&src, nent, s));
  c(dalloc(&start, n + 1, s));
  c(dalloc(&bad, 1, s));
  int hbad = 0;
  if (e == hipSuccess) c(hipMemsetAsync(bad, 0, sizeof(int), s));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rec_cn, dim3(blocks_for(nent, TPB)), dim3(TPB), 0, s, nrec, el->ne, eid,
                       el->cn, cnr, bad);
    c(hipGetLastError());
  }
  if (e == hipSuccess) c(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  if (e == hipSuccess) c(hipStreamSynchronize(s));
  int bits = 1;
  while (bits < 31 && (1LL << bits) <= n) bits++;
  size_t tb = 0;
  if (e == hipSuccess && !hbad) {
    hipLaunchKernelGGL(k_sload, dim3(blocks_for(nrec, TPB)), dim3(TPB), 0, s, nrec, eid, ls,
                       nodvals, el->conn, el->xyz, fe);
    hipLaunchKernelGGL(k_post_keys, dim3(blocks_for(nent, TPB)), dim3(TPB), 0, s, nent, (int)n,
                       8, 3, cnr, key, val);
    c(hipGetLastError());
    if (e == hipSuccess)
      c(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, val, src, (int)nent, 0, bits,
                                           s));
    if (e == hipSuccess) c(hipMallocAsync(&tmp, std::max<size_t>(tb, 1), s));
    if (e == hipSuccess)
      c(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, skey, val, src, (int)nent, 0, bits, s));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_post_fill_int, dim3(blocks_for(n + 1, TPB)), dim3(TPB), 0, s, n + 1,
                         (int)nent, start);
      hipLaunchKernelGGL(k_post_start, dim3(blocks_for(nent, TPB)), dim3(TPB), 0, s, nent,
                         (int)n, skey, start);
      hipLaunchKernelGGL(k_post_gather, dim3(blocks_for(n, TPB)), dim3(TPB), 0, s, (int)n, start,
                         src, fe, rhs, 1);
      c(hipGetLastError());
    }
  }
  c(hipStreamSynchronize(s));
  if (tmp) (void)hipFreeAsync(tmp, s);
  dfree(fe, s); dfree(cnr, s); dfree(key, s); dfree(skey, s); dfree(val, s); dfree(src, s);
  dfree(start, s); dfree(bad, s);
  SSF_HIP(e);
  SSF_CHECK(!hbad, SSF_EINVAL, "ssf_elements_surface_load: element id out of range");
  return SSF_OK;
}

extern "C" int ssf_elements_surface_load_dev(ssf_elements *el, int64_t nrec, const int32_t *eid_dev,
                                             const int32_t *ls_dev, const double *nodvals_dev,
                                             double *rhs_dev)
{
  SSF_ENTER();
  SSF_CHECK(el && nrec >= 0 && (nrec == 0 || (eid_dev && ls_dev && nodvals_dev)) &&
                (rhs_dev || el->n == 0),
            SSF_EINVAL, "ssf_elements_surface_load_dev: NULL argument");
  SSF_CHECK(el->kind == SSF_ELEM_HEX8_ELAST, SSF_EUNSUP,
            "ssf_elements_surface_load: hex8 elasticity only (linhex::node_forces_surf)");
  SSF_CHECK(nrec * 24 < (1LL << 31), SSF_EINVAL, "ssf_elements_surface_load: too many records");
  SSF_HIP(hipSetDevice(el->ctx->device));
  if (nrec == 0 || el->n == 0) return SSF_OK;
  return sload_launch(el, nrec, eid_dev, ls_dev, nodvals_dev, rhs_dev);
}

extern "C" int ssf_elements_surface_load(ssf_elements *el, int64_t nrec, const int64_t *eid,
                                         const int64_t *ls, const double *nodvals, double *rhs)
{
  SSF_ENTER();
  SSF_CHECK(el && nrec >= 0 && (nrec == 0 || (eid && ls && nodvals)) && (rhs || el->n == 0),
            SSF_EINVAL, "ssf_elements_surface_load: NULL argument");
  SSF_CHECK(el->kind == SSF_ELEM_HEX8_ELAST, SSF_EUNSUP,
            "ssf_elements_surface_load: hex8 elasticity only (linhex::node_forces_surf)");
  SSF_CHECK(nrec * 24 < (1LL << 31), SSF_EINVAL, "ssf_elements_surface_load: too many records");
  for (int64_t r = 0; r < nrec; r++)
    SSF_CHECK(eid[r] >= 0 && eid[r] < el->ne, SSF_EINVAL,
              "ssf_elements_surface_load: element id out of range");
  SSF_HIP(hipSetDevice(el->ctx->device));
  if (nrec == 0 || el->n == 0) return SSF_OK;
  hipStream_t s = el->ctx->stream;
  std::vector<int> e32(nrec), l32(nrec * 6);
  for (int64_t r = 0; r < nrec; r++) e32[r] = (int)eid[r];
  for (int64_t k = 0; k < nrec * 6; k++) l32[k] = (int)ls[k];
  int *d_eid = nullptr, *d_ls = nullptr;
  double *d_nv = nullptr, *d_rhs = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&d_eid, nrec, s));
  c(dalloc(&d_ls, nrec * 6, s));
  c(dalloc(&d_nv, nrec * 72, s));
  c(dalloc(&d_rhs, el->n, s));
  if (e == hipSuccess) {
    c(hipMemcpyAsync(d_eid, e32.data(), nrec * sizeof(int), hipMemcpyHostToDevice, s));
    c(hipMemcpyAsync(d_ls, l32.data(), nrec * 6 * sizeof(int), hipMemcpyHostToDevice, s));
    c(hipMemcpyAsync(d_nv, nodvals, nrec * 72 * sizeof(double), hipMemcpyHostToDevice, s));
    c(hipMemcpyAsync(d_rhs, rhs, el->n * sizeof(double), hipMemcpyHostToDevice, s));
  }
  int rc = SSF_OK;
  if (e == hipSuccess) rc = sload_launch(el, nrec, d_eid, d_ls, d_nv, d_rhs);
  if (e == hipSuccess && rc == SSF_OK)
    c(hipMemcpyAsync(rhs, d_rhs, el->n * sizeof(double), hipMemcpyDeviceToHost, s));
  c(hipStreamSynchronize(s));
  dfree(d_eid, s);
  dfree(d_ls, s);
  dfree(d_nv, s);
  dfree(d_rhs, s);
  if (rc) return rc;
  SSF_HIP(e);
  return SSF_OK;
}

extern "C" int ssf_elements_volume_load_dev(ssf_elements *el, const double *nodval_dev,
                                            double *rhs_dev)
{
  SSF_ENTER();
  SSF_CHECK(el && nodval_dev && (rhs_dev || el->n == 0), SSF_EINVAL,
            "ssf_elements_volume_load_dev: NULL argument");
  SSF_CHECK(el->kind != SSF_ELEM_QUAD4_PSTRESS, SSF_EUNSUP,
            "ssf_elements_volume_load: hex8 elasticity and heat only");
  SSF_HIP(hipSetDevice(el->ctx->device));
  SSF_HIP(vload_launch(el, nodval_dev, rhs_dev));
  return SSF_OK;
}

extern "C" int ssf_elements_volume_load(ssf_elements *el, const double *nodval, double *rhs)
{
  SSF_ENTER();
  SSF_CHECK(el && nodval && (rhs || el->n == 0), SSF_EINVAL,
            "ssf_elements_volume_load: NULL argument");
  SSF_CHECK(el->kind != SSF_ELEM_QUAD4_PSTRESS, SSF_EUNSUP,
            "ssf_elements_volume_load: hex8 elasticity and heat only");
  SSF_HIP(hipSetDevice(el->ctx->device));
  hipStream_t s = el->ctx->stream;
  double *d_nv = nullptr, *d_rhs = nullptr;
  hipError_t e = hipSuccess;
  auto c = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  c(dalloc(&d_nv, el->ne * el->nd, s));
  c(dalloc(&d_rhs, el->n, s));
  if (e == hipSuccess)
    c(hipMemcpyAsync(d_nv, nodval, el->ne * el->nd * sizeof(double), hipMemcpyHostToDevice, s));
  if (e == hipSuccess && el->n)
    c(hipMemcpyAsync(d_rhs, rhs, el->n * sizeof(double), hipMemcpyHostToDevice, s));
  if (e == hipSuccess) c(vload_launch(el, d_nv, d_rhs));
  if (e == hipSuccess && el->n)
    c(hipMemcpyAsync(rhs, d_rhs, el->n * sizeof(double), hipMemcpyDeviceToHost, s));
  c(hipStreamSynchronize(s));
  dfree(d_nv, s);
  dfree(d_rhs, s);
  SSF_HIP(e);
  return SSF_OK;
}
