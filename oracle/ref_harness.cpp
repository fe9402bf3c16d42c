// Test infrastructure only: a thin extern "C" harness around the reference's own
// GEFEL sources (compiled from /root/reference by oracle/Makefile into oracle/_ref/).
// Nothing in the product links this; tests use it to pin the oracle restatement.
//
// Storage, SpMV and CG calls go straight into the reference classes
// (GEFEL/scr.cpp, GEFEL/cr.cpp). The element routines compose the reference's
// GEFEL building blocks (gauss_points, dx/dy/dz_bf_lin_hex_3d, derivatives_3d,
// derivatives_2d, bf_lin_4_2d, bdbjac, bdbj) in the loop order of
// MEFEL/SRC/linhex.cpp:583-640, MEFEL/SRC/plelemlq.cpp:611-690 and
// TRFEL/SRC/linhext.cpp:774-840, because those element classes read MEFEL/TRFEL
// globals (Mt, Mm, Tt, Tm) and cannot be called in isolation. The isotropic
// material matrices (MEFEL/SRC/elastisomat.cpp:255-266,395-410 and
// TRFEL/SRC/isotrmat.cpp:124-135) are restated here for the same reason.
#include <math.h>
#include <string.h>
#include "scr.h"
#include "cr.h"
#include "precond.h"
#include "matrix.h"
#include "vector.h"
#include "basefun.h"
#include "difcalc.h"
#include "intp.h"
#include "gtopology.h"
#include "gelement.h"

extern "C" {

// ---------------------------------------------------------------- symcomprow
void *ref_scr_create(long n, long ne, long ndofe, const long *cn)
{
  symcomprow *s = new symcomprow();
  long *c = new long[ndofe];
  s->allocadr(n);
  for (long e = 0; e < ne; e++) {
    memcpy(c, cn + e * ndofe, ndofe * sizeof(long));
    s->numcontr_elem(c, ndofe);
  }
  s->addresses();
  for (long e = 0; e < ne; e++) {
    memcpy(c, cn + e * ndofe, ndofe * sizeof(long));
    s->fillarray_elem(c, ndofe);
  }
  s->sort_and_colindex_tko();
  delete[] c;
  return s;
}
long ref_scr_negm(void *h) { return ((symcomprow *)h)->negm; }
void ref_scr_get(void *h, long *adr, long *ci, double *a)
{
  symcomprow *s = (symcomprow *)h;
  if (adr) memcpy(adr, s->adr, (s->n + 1) * sizeof(long));
  if (ci) memcpy(ci, s->ci, s->negm * sizeof(long));
  if (a) memcpy(a, s->a, s->negm * sizeof(double));
}
void ref_scr_set_a(void *h, const double *a)
{
  symcomprow *s = (symcomprow *)h;
  memcpy(s->a, a, s->negm * sizeof(double));
}
void ref_scr_localized(void *h, double *lm, long *cn, long nc)
{
  ((symcomprow *)h)->localized(lm, cn, nc);
}
void ref_scr_mxv(void *h, double *b, double *c) { ((symcomprow *)h)->mxv_scr(b, c); }
void ref_scr_cg_prec(void *h, double *x, double *y, long ni, double res, long *ani,
                     double *ares, double zero, long iv, long tprec)
{
  long a = *ani;
  double r = *ares;
  ((symcomprow *)h)->cg_prec(x, y, ni, res, a, r, zero, iv, tprec, 0.0, NULL);
  *ani = a;
  *ares = r;
}
void ref_scr_cg(void *h, double *x, double *y, long ni, double res, long *ani,
                double *ares, double zero, long iv)
{
  long a = *ani;
  double r = *ares;
  ((symcomprow *)h)->cg(x, y, ni, res, a, r, zero, iv);
  *ani = a;
  *ares = r;
}
// symcomprow::incomplete_ldl (GEFEL/scr.cpp:990) on the object's values; the
// factor incdec[negm] is copied out
void ref_scr_incomplete_ldl(void *h, double gamma, double *out)
{
  symcomprow *s = (symcomprow *)h;
  delete[] s->incdec;
  s->incdec = NULL;
  s->incomplete_ldl(gamma);
  memcpy(out, s->incdec, s->negm * sizeof(double));
}
// symcomprow::prec_ildl_scr (GEFEL/scr.cpp:1031) with the factor of the last
// ref_scr_incomplete_ldl
void ref_scr_prec_ildl(void *h, double *x, double *y) { ((symcomprow *)h)->prec_ildl_scr(x, y); }
void ref_scr_free(void *h) { delete (symcomprow *)h; }

// ------------------------------------------------------------------ comprow
void *ref_cr_create(long n, long ne, long ndofe, const long *cn)
{
  comprow *s = new comprow();
  long *c = new long[ndofe];
  s->allocadr(n);
  for (long e = 0; e < ne; e++) {
    memcpy(c, cn + e * ndofe, ndofe * sizeof(long));
    s->numcontr_elem(c, ndofe);
  }
  s->addresses();
  for (long e = 0; e < ne; e++) {
    memcpy(c, cn + e * ndofe, ndofe * sizeof(long));
    s->fillarray_elem(c, ndofe);
  }
  s->sort_and_colindex();
  delete[] c;
  return s;
}
long ref_cr_negm(void *h) { return ((comprow *)h)->negm; }
void ref_cr_get(void *h, long *adr, long *ci, double *a)
{
  comprow *s = (comprow *)h;
  if (adr) memcpy(adr, s->adr, (s->n + 1) * sizeof(long));
  if (ci) memcpy(ci, s->ci, s->negm * sizeof(long));
  if (a) memcpy(a, s->a, s->negm * sizeof(double));
}
void ref_cr_localized(void *h, double *lm, long *cn, long nc)
{
  ((comprow *)h)->localized(lm, cn, nc);
}
void ref_cr_mxv(void *h, double *b, double *c) { ((comprow *)h)->mxv_cr(b, c); }
void ref_cr_cg_prec_noprec(void *h, double *x, double *y, long ni, double res, long *ani,
                           double *ares, double zero, long iv)
{
  comprow *s = (comprow *)h;
  precond pr;
  pr.pt = noprecond;
  pr.n = s->n;
  long a = *ani;
  double r = *ares;
  s->cg_prec(pr, x, y, ni, res, a, r, zero, iv);
  *ani = a;
  *ares = r;
}
void ref_cr_set_a(void *h, const double *a)
{
  comprow *s = (comprow *)h;
  memcpy(s->a, a, s->negm * sizeof(double));
}
void ref_cr_mtxv(void *h, double *b, double *c) { ((comprow *)h)->mtxv_cr(b, c); }
// comprow::bicg (GEFEL/cr.cpp:1353-1419) itself
void ref_cr_bicg(void *h, double *x, double *y, long ni, double res, long *ani, double *ares,
                 double zero, long iv)
{
  long a = *ani;
  double r = *ares;
  ((comprow *)h)->bicg(x, y, ni, res, a, r, zero, iv);
  *ani = a;
  *ares = r;
}
void ref_cr_free(void *h) { delete (comprow *)h; }

// ------------------------------------------------------------ element routines
// hex8 linear elasticity, order of MEFEL/SRC/linhex.cpp:583-640 with geom_matrix :153-186
void ref_hex8_elast_ke(const double *xc, const double *yc, const double *zc, double e,
                       double nu, double *ke)
{
  const long nne = 8, ndofe = 24;
  vector x(nne), y(nne), z(nne), w(2), gp(2);
  vector dx(nne), dy(nne), dz(nne);
  matrix gm(6, ndofe), d(6, 6), sm(ndofe, ndofe);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; }
  // elastisomat::matstiff_spacestr
  double g = e / 2.0 / (1.0 + nu);
  double s = e / (1.0 + nu) / (1.0 - 2.0 * nu);
  fillm(0.0, d);
  d[0][0] = s * (1 - nu); d[0][1] = s * nu; d[0][2] = s * nu;
  d[1][0] = d[0][1]; d[1][1] = d[0][0]; d[1][2] = d[0][1];
  d[2][0] = d[0][1]; d[2][1] = d[0][1]; d[2][2] = d[0][0];
  d[3][3] = g; d[4][4] = g; d[5][5] = g;

  nullm(sm);
  gauss_points(gp.a, w.a, 2);
  for (long i = 0; i < 2; i++) {
    double xi = gp[i];
    for (long j = 0; j < 2; j++) {
      double eta = gp[j];
      for (long k = 0; k < 2; k++) {
        double zeta = gp[k], jac;
        dx_bf_lin_hex_3d(dx.a, eta, zeta);
        dy_bf_lin_hex_3d(dy.a, xi, zeta);
        dz_bf_lin_hex_3d(dz.a, xi, eta);
        derivatives_3d(dx, dy, dz, jac, x, y, z, xi, eta, zeta);
        nullm(gm);
        long jj = 0, kk = 1, ll = 2;
        for (long a = 0; a < nne; a++) {
          gm[0][jj] = dx[a]; gm[1][kk] = dy[a]; gm[2][ll] = dz[a];
          gm[3][kk] = dz[a]; gm[3][ll] = dy[a];
          gm[4][jj] = dz[a]; gm[4][ll] = dx[a];
          gm[5][jj] = dy[a]; gm[5][kk] = dx[a];
          jj += 3; kk += 3; ll += 3;
        }
        jac *= w[i] * w[j] * w[k];
        if (jac < 0.0) jac = fabs(jac);
        bdbjac(sm, gm, d, gm, jac);
      }
    }
  }
  for (long i = 0; i < ndofe; i++)
    for (long j = 0; j < ndofe; j++) ke[i * ndofe + j] = sm[i][j];
}

// quad4 plane stress, order of MEFEL/SRC/plelemlq.cpp:611-690 with geom_matrix :294-314
void ref_quad4_pstress_ke(const double *xc, const double *yc, double e, double nu,
                          const double *thick_nodes, double *ke)
{
  const long nne = 4, ndofe = 8;
  vector x(nne), y(nne), t(nne), w(2), gp(2), dx(nne), dy(nne), bf(nne);
  matrix gm(3, ndofe), d(3, 3), sm(ndofe, ndofe);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; t[i] = thick_nodes[i]; }
  // elastisomat::matstiff_plstress
  double c = e / (1.0 - nu * nu);
  fillm(0.0, d);
  d[0][0] = c; d[0][1] = c * nu; d[0][2] = 0.0;
  d[1][0] = c * nu; d[1][1] = c; d[1][2] = 0.0;
  d[2][0] = 0.0; d[2][1] = 0.0; d[2][2] = e / 2.0 / (1.0 + nu);

  nullm(sm);
  gauss_points(gp.a, w.a, 2);
  for (long i = 0; i < 2; i++) {
    double xi = gp[i];
    for (long j = 0; j < 2; j++) {
      double eta = gp[j], jac, thick;
      dx_bf_lin_4_2d(dx.a, eta);
      dy_bf_lin_4_2d(dy.a, xi);
      derivatives_2d(dx, dy, jac, x, y, xi, eta);
      nullm(gm);
      long i1 = 0, i2 = 1;
      for (long a = 0; a < nne; a++) {
        gm[0][i1] = dx[a]; gm[1][i2] = dy[a];
        gm[2][i1] = dy[a]; gm[2][i2] = dx[a];
        i1 += 2; i2 += 2;
      }
      // planeelemlq::approx = scprd(bf_lin_4_2d, nodal thickness)
      bf_lin_4_2d(bf.a, xi, eta);
      scprd(bf, t, thick);
      jac *= thick * w[i] * w[j];
      if (jac < 0.0) jac = fabs(jac);
      bdbjac(sm, gm, d, gm, jac);
    }
  }
  for (long i = 0; i < ndofe; i++)
    for (long j = 0; j < ndofe; j++) ke[i * ndofe + j] = sm[i][j];
}

// hex8 heat conduction, order of TRFEL/SRC/linhext.cpp:774-840 with grad_matrix :743-761
void ref_hex8_heat_ke(const double *xc, const double *yc, const double *zc, double kk,
                      double *ke)
{
  const long nne = 8;
  vector x(nne), y(nne), z(nne), w(2), gp(2), dx(nne), dy(nne), dz(nne);
  matrix gm(3, nne), d(3, 3), km(nne, nne), dc(1, 3), nmat(1, nne);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; }
  nullm(km);
  gauss_points(gp.a, w.a, 2);
  for (long i = 0; i < 2; i++) {
    double xi = gp[i], ww1 = w[i];
    for (long j = 0; j < 2; j++) {
      double eta = gp[j], ww2 = w[j];
      for (long k = 0; k < 2; k++) {
        double zeta = gp[k], ww3 = w[k], jac;
        dx_bf_lin_hex_3d(dx.a, eta, zeta);
        dy_bf_lin_hex_3d(dy.a, xi, zeta);
        dz_bf_lin_hex_3d(dz.a, xi, eta);
        derivatives_3d(dx, dy, dz, jac, x, y, z, xi, eta, zeta);
        nullm(gm);
        for (long a = 0; a < nne; a++) { gm[0][a] = dx[a]; gm[1][a] = dy[a]; gm[2][a] = dz[a]; }
        // isotrmat::matcond3d
        fillm(0.0, d);
        d[0][0] = kk; d[1][1] = kk; d[2][2] = kk;
        if (jac < 0.0) jac = fabs(jac);
        jac *= ww1 * ww2 * ww3;
        bdbj(km.a, gm.a, d.a, jac, gm.m, gm.n);
        // convective term: matcond2 leaves the zeroed 1x3 matrix untouched for an
        // isotropic medium, so this adds exact zeros (kept for identical rounding)
        nullm(dc);
        bf_lin_hex_3d(nmat.a, xi, eta, zeta);
        bdbjac(km, nmat, dc, gm, jac);
      }
    }
  }
  for (long i = 0; i < nne; i++)
    for (long j = 0; j < nne; j++) ke[i * nne + j] = km[i][j];
}

// ------------------------------------------------------- post-solve element pass
// Integration-point values and element forces, composed from GEFEL's mxv, mtxv,
// cmulv, addv, scprd and the shape-function routines in the loop orders of
// linhex::gl_ip_strains / elem_integration (MEFEL/SRC/linhex.cpp:1110-1145,
// 2157-2191), planeelemlq::gl_ip_strains / elem_integration (MEFEL/SRC/plelemlq.cpp:
// 1173-1205, 2307-2359), linhext::intpointgrad / internal_fluxes (TRFEL/SRC/
// linhext.cpp:432-498, 1086-1146). The material laws of elastisomat::nlstresses
// (MEFEL/SRC/elastisomat.cpp:627-651) and transmat::computenlfluxes (TRFEL/SRC/
// transmat.cpp:1123-1148) are restated, as the material classes read MEFEL/TRFEL
// globals.
static void ref_hex_gm(matrix &gm, vector &x, vector &y, vector &z, double xi, double eta,
                       double zeta, double &jac)
{
  const long nne = 8;
  vector dx(nne), dy(nne), dz(nne);
  dx_bf_lin_hex_3d(dx.a, eta, zeta);
  dy_bf_lin_hex_3d(dy.a, xi, zeta);
  dz_bf_lin_hex_3d(dz.a, xi, eta);
  derivatives_3d(dx, dy, dz, jac, x, y, z, xi, eta, zeta);
  nullm(gm);
  long jj = 0, kk = 1, ll = 2;
  for (long a = 0; a < nne; a++) {
    gm[0][jj] = dx[a]; gm[1][kk] = dy[a]; gm[2][ll] = dz[a];
    gm[3][kk] = dz[a]; gm[3][ll] = dy[a];
    gm[4][jj] = dz[a]; gm[4][ll] = dx[a];
    gm[5][jj] = dy[a]; gm[5][kk] = dx[a];
    jj += 3; kk += 3; ll += 3;
  }
}

void ref_hex8_elast_ip(const double *xc, const double *yc, const double *zc, double e,
                       double nu, const double *rr, double *eps_out, double *sig_out,
                       double *f_out)
{
  const long nne = 8, ndofe = 24;
  vector x(nne), y(nne), z(nne), r(ndofe), w(2), gp(2), eps(6), sig(6), contr(ndofe),
      nv(ndofe);
  matrix gm(6, ndofe), d(6, 6);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; }
  for (long i = 0; i < ndofe; i++) r[i] = rr[i];
  double g = e / 2.0 / (1.0 + nu);
  double s = e / (1.0 + nu) / (1.0 - 2.0 * nu);
  fillm(0.0, d);
  d[0][0] = s * (1 - nu); d[0][1] = s * nu; d[0][2] = s * nu;
  d[1][0] = d[0][1]; d[1][1] = d[0][0]; d[1][2] = d[0][1];
  d[2][0] = d[0][1]; d[2][1] = d[0][1]; d[2][2] = d[0][0];
  d[3][3] = g; d[4][4] = g; d[5][5] = g;
  gauss_points(gp.a, w.a, 2);
  nullv(nv);
  long ip = 0;
  for (long i = 0; i < 2; i++)
    for (long j = 0; j < 2; j++)
      for (long k = 0; k < 2; k++) {
        double jac;
        ref_hex_gm(gm, x, y, z, gp[i], gp[j], gp[k], jac);
        mxv(gm, r, eps);
        mxv(d, eps, sig);
        mtxv(gm, sig, contr);
        cmulv(jac * w[i] * w[j] * w[k], contr);
        addv(contr, nv, nv);
        for (long c = 0; c < 6; c++) { eps_out[6 * ip + c] = eps[c]; sig_out[6 * ip + c] = sig[c]; }
        ip++;
      }
  for (long i = 0; i < ndofe; i++) f_out[i] = nv[i];
}

void ref_quad4_pstress_ip(const double *xc, const double *yc, double e, double nu,
                          const double *thick_nodes, const double *rr, double *eps_out,
                          double *sig_out, double *f_out)
{
  const long nne = 4, ndofe = 8;
  vector x(nne), y(nne), t(nne), r(ndofe), w(2), gp(2), dx(nne), dy(nne), bf(nne), eps3(3),
      eps(4), sig(4), ipv(3), contr(ndofe), nv(ndofe);
  matrix gm(3, ndofe), d(4, 4);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; t[i] = thick_nodes[i]; }
  for (long i = 0; i < ndofe; i++) r[i] = rr[i];
  double c = e / (1.0 - nu * nu);
  fillm(0.0, d);
  d[0][0] = c; d[0][1] = c * nu; d[0][2] = 0.0;
  d[1][0] = c * nu; d[1][1] = c; d[1][2] = 0.0;
  d[2][0] = 0.0; d[2][1] = 0.0; d[2][2] = e / 2.0 / (1.0 + nu);
  gauss_points(gp.a, w.a, 2);
  nullv(nv);
  long ip = 0;
  for (long i = 0; i < 2; i++) {
    double xi = gp[i];
    for (long j = 0; j < 2; j++) {
      double eta = gp[j], jac, thick;
      bf_lin_4_2d(bf.a, xi, eta);
      scprd(bf, t, thick);
      dx_bf_lin_4_2d(dx.a, eta);
      dy_bf_lin_4_2d(dy.a, xi);
      derivatives_2d(dx, dy, jac, x, y, xi, eta);
      nullm(gm);
      long i1 = 0, i2 = 1;
      for (long a = 0; a < nne; a++) {
        gm[0][i1] = dx[a]; gm[1][i2] = dy[a];
        gm[2][i1] = dy[a]; gm[2][i2] = dx[a];
        i1 += 2; i2 += 2;
      }
      mxv(gm, r, eps3);
      eps[0] = eps3[0]; eps[1] = eps3[1]; eps[2] = eps3[2]; eps[3] = 0.0;
      mxv(d, eps, sig);
      eps[3] = -nu / (1.0 - nu) * (eps[0] + eps[1]);
      ipv[0] = sig[0]; ipv[1] = sig[1]; ipv[2] = sig[2];
      mtxv(gm, ipv, contr);
      cmulv(jac * w[i] * w[j] * thick, contr);
      addv(contr, nv, nv);
      for (long q = 0; q < 4; q++) { eps_out[4 * ip + q] = eps[q]; sig_out[4 * ip + q] = sig[q]; }
      ip++;
    }
  }
  for (long i = 0; i < ndofe; i++) f_out[i] = nv[i];
}

void ref_hex8_heat_ip(const double *xc, const double *yc, const double *zc, double kk,
                      const double *tt, double *grad_out, double *flux_out, double *f_out)
{
  const long nne = 8;
  vector x(nne), y(nne), z(nne), t(nne), w(2), gp(2), dx(nne), dy(nne), dz(nne), grad(3),
      fl(3), contr(nne), nv(nne);
  matrix gm(3, nne), d(3, 3);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; t[i] = tt[i]; }
  fillm(0.0, d);
  d[0][0] = kk; d[1][1] = kk; d[2][2] = kk;
  gauss_points(gp.a, w.a, 2);
  nullv(nv);
  long ip = 0;
  for (long i = 0; i < 2; i++)
    for (long j = 0; j < 2; j++)
      for (long k = 0; k < 2; k++) {
        double xi = gp[i], eta = gp[j], zeta = gp[k], jac;
        dx_bf_lin_hex_3d(dx.a, eta, zeta);
        dy_bf_lin_hex_3d(dy.a, xi, zeta);
        dz_bf_lin_hex_3d(dz.a, xi, eta);
        derivatives_3d(dx, dy, dz, jac, x, y, z, xi, eta, zeta);
        nullm(gm);
        for (long a = 0; a < nne; a++) { gm[0][a] = dx[a]; gm[1][a] = dy[a]; gm[2][a] = dz[a]; }
        mxv(gm, t, grad);
        for (long jj = 0; jj < 3; jj++) {
          fl[jj] = 0.0;
          for (long q = 0; q < 3; q++) fl[jj] -= d[jj][q] * grad[q];
        }
        mtxv(gm, fl, contr);
        if (jac < 0.0) jac = fabs(jac);
        cmulv(jac * w[i] * w[j] * w[k], contr);
        addv(contr, nv, nv);
        for (long c = 0; c < 3; c++) { grad_out[3 * ip + c] = grad[c]; flux_out[3 * ip + c] = fl[c]; }
        ip++;
      }
  for (long i = 0; i < nne; i++) f_out[i] = nv[i];
}

// GEFEL/matrix.cpp:3500 locglob
void ref_locglob(double *gv, const double *lv, const long *cn, long n) { locglob(gv, lv, cn, n); }

// ------------------------------------------------------- nonstationary heat
// capacity matrix in the loop order of linhext::capacity_matrix
// (TRFEL/SRC/linhext.cpp:979-1029): GEFEL's gauss_points, jac_3d, bf_lin_hex_3d,
// scprd (linhext::approx) and nnj; c restates isotrmat::matcap
void ref_hex8_heat_cap(const double *xc, const double *yc, const double *zc, double c,
                       const double *dens_nodes, double *cm_out)
{
  const long nne = 8;
  vector x(nne), y(nne), z(nne), w(2), gp(2), dens(nne), bf(nne);
  matrix n(1, nne), cm(nne, nne);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; dens[i] = dens_nodes[i]; }
  gauss_points(gp.a, w.a, 2);
  nullm(cm);
  for (long i = 0; i < 2; i++) {
    double xi = gp[i], w1 = w[i];
    for (long j = 0; j < 2; j++) {
      double eta = gp[j], w2 = w[j];
      for (long k = 0; k < 2; k++) {
        double zeta = gp[k], w3 = w[k], jac, rho;
        jac_3d(jac, x, y, z, xi, eta, zeta);
        nullm(n);
        bf_lin_hex_3d(n.a, xi, eta, zeta);
        bf_lin_hex_3d(bf.a, xi, eta, zeta);
        scprd(bf, dens, rho);
        if (jac < 0.0) jac = fabs(jac);
        jac *= w1 * w2 * w3 * rho * c;
        nnj(cm.a, n.a, jac, n.m, n.n);
      }
    }
  }
  for (long i = 0; i < nne; i++)
    for (long j = 0; j < nne; j++) cm_out[i * nne + j] = cm[i][j];
}

// one_step_linear (TRFEL/SRC/npsolvert.cpp:367-482) on the reference's own
// symcomprow objects: K->mxv_scr, A->scalmat_scr, A->addmat_scr, A->cg_prec; the
// vector loops of the step are restated
void ref_np_step_linear(void *hK, void *hC, void *hA, long n, double alpha, double dt,
                        const double *rhs, double *lhs, double *tdlhs, long ni, double res,
                        double zero, long *ani, double *ares)
{
  symcomprow *K = (symcomprow *)hK, *C = (symcomprow *)hC, *A = (symcomprow *)hA;
  double *d = new double[n], *p = new double[n], *f = new double[n];
  for (long j = 0; j < n; j++) d[j] = lhs[j] + (1.0 - alpha) * dt * tdlhs[j];
  K->mxv_scr(d, p);
  memcpy(A->a, K->a, K->negm * sizeof(double));
  A->scalmat_scr(dt * alpha);
  A->addmat_scr(1.0, *C);
  for (long j = 0; j < n; j++) f[j] = rhs[j] - p[j];
  long a = *ani;
  double r = *ares;
  A->cg_prec(tdlhs, f, ni, res, a, r, zero, 1, 1, 0.0, NULL);
  *ani = a;
  *ares = r;
  for (long j = 0; j < n; j++) lhs[j] = d[j] + alpha * dt * tdlhs[j];
  delete[] d;
  delete[] p;
  delete[] f;
}

// one pass of linear_nonstat_solv_dform's time loop (TRFEL/SRC/npsolvert.cpp:
// 1152-1185) on the reference's own symcomprow objects: C->mxv_scr,
// A->scalmat_scr, A->addmat_scr, A->cg_prec; the vector loops are restated
void ref_np_step_dform(void *hK, void *hC, void *hA, long n, double alpha, double dt,
                       const double *rhs, double *lhs, double *tdlhs, long ni, double res,
                       double zero, long *ani, double *ares)
{
  symcomprow *K = (symcomprow *)hK, *C = (symcomprow *)hC, *A = (symcomprow *)hA;
  double *d = new double[n], *p = new double[n], *f = new double[n];
  for (long j = 0; j < n; j++) d[j] = lhs[j] + (1.0 - alpha) * dt * tdlhs[j];
  C->mxv_scr(d, p);
  memcpy(A->a, K->a, K->negm * sizeof(double));
  A->scalmat_scr(dt * alpha);
  A->addmat_scr(1.0, *C);
  for (long j = 0; j < n; j++) f[j] = rhs[j] * alpha * dt + p[j];
  long a = *ani;
  double r = *ares;
  A->cg_prec(lhs, f, ni, res, a, r, zero, 1, 1, 0.0, NULL);
  *ani = a;
  *ares = r;
  for (long j = 0; j < n; j++) tdlhs[j] = (lhs[j] - d[j]) / dt / alpha;
  delete[] d;
  delete[] p;
  delete[] f;
}

// ------------------------------------------------ element volume loads / sources
// kind 0: loadel::volumeload with linhex::load_matrix (MEFEL/SRC/loadel.cpp:438,
// MEFEL/SRC/linhex.cpp:969) and linhex::bf_matrix (:126); kind 1:
// linhext::quantity_source_vector (TRFEL/SRC/linhext.cpp:1039); GEFEL's
// gauss_points, jac_3d, bf_lin_hex_3d, nnj, mxv, addv
void ref_hex8_volume_load(int kind, const double *xc, const double *yc, const double *zc,
                          const double *nodval_in, double *nf_out)
{
  const long nne = 8, nd = kind == 0 ? 24 : 8, mm = kind == 0 ? 3 : 1;
  vector x(nne), y(nne), z(nne), w(2), gp(2), bf(nne), nodval(nd), v(nd), sv(nd);
  matrix n(mm, nd), lm(nd, nd);
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; }
  for (long i = 0; i < nd; i++) nodval[i] = nodval_in[i];
  gauss_points(gp.a, w.a, 2);
  nullm(lm);
  for (long i = 0; i < 2; i++)
    for (long j = 0; j < 2; j++)
      for (long k = 0; k < 2; k++) {
        double xi = gp[i], eta = gp[j], zeta = gp[k], jac;
        jac_3d(jac, x, y, z, xi, eta, zeta);
        nullm(n);
        if (kind == 0) {
          bf_lin_hex_3d(bf.a, xi, eta, zeta);
          long jj = 0, kk = 1, ll = 2;
          for (long a = 0; a < nne; a++) {
            n[0][jj] = bf[a]; jj += 3;
            n[1][kk] = bf[a]; kk += 3;
            n[2][ll] = bf[a]; ll += 3;
          }
        } else {
          bf_lin_hex_3d(n.a, xi, eta, zeta);
          if (jac < 0.0) jac = fabs(jac);
        }
        jac *= w[i] * w[j] * w[k];
        nnj(lm.a, n.a, jac, n.m, n.n);
      }
  if (kind == 0) {
    mxv(lm.a, nodval.a, v.a, nd, nd);
    for (long i = 0; i < nd; i++) nf_out[i] = v[i];
  } else {
    nullv(sv);
    mxv(lm, nodval, v);
    addv(sv, v, sv);
    for (long i = 0; i < nd; i++) nf_out[i] = sv[i];
  }
}

// linhex::node_forces_surf (MEFEL/SRC/linhex.cpp:2543-3047) composed from GEFEL's
// gauss_points, jac2d_3d (with the face index), bf_lin_hex_3d, nnj, mxv and addv in
// the member's loop order, with linhex::locglob_nodeval (:3049-3143) composed from
// scprd, cmulv and mxv; the per-face DOF positions are those the member writes
void ref_hex8_surface_load(const double *xc, const double *yc, const double *zc,
                           const long *is, const double *nv, double *nf_out)
{
  // for face f: the fixed coordinate (0 xi, 1 eta, 2 zeta) and value, the
  // coordinates of the outer / inner loop, the DOFs of the 4 nodal values, the
  // node pairs of locglob_nodeval's g1 and g2
  static const int fix[6] = {0, 1, 0, 1, 2, 2};
  static const double val[6] = {1.0, 1.0, -1.0, -1.0, 1.0, -1.0};
  static const int outer[6] = {1, 0, 1, 0, 0, 0}, inner[6] = {2, 2, 2, 2, 1, 1};
  static const int dof[6][4] = {{0, 9, 21, 12}, {3, 0, 12, 15}, {6, 3, 15, 18},
                                {9, 6, 18, 21}, {0, 3, 6, 9}, {12, 21, 18, 15}};
  static const int gn[6][4] = {{4, 7, 3, 7}, {5, 4, 0, 4}, {6, 5, 1, 5},
                               {7, 6, 2, 6}, {3, 2, 1, 2}, {5, 6, 7, 6}};
  const long nne = 8, ndofe = 24;
  vector x(nne), y(nne), z(nne), gp(2), w(2), av(ndofe), v(ndofe), nf(ndofe), bf(nne);
  matrix n(3, ndofe), am(ndofe, ndofe);
  double tnv[12];
  for (long i = 0; i < nne; i++) { x[i] = xc[i]; y[i] = yc[i]; z[i] = zc[i]; }
  gauss_points(gp.a, w.a, 2);
  for (int f = 0; f < 6; f++) {
    if (is[f] <= 0) continue;
    nullm(am);
    for (long i = 0; i < 2; i++)
      for (long j = 0; j < 2; j++) {
        double c[3], jac;
        c[fix[f]] = val[f];
        c[outer[f]] = gp[i];
        c[inner[f]] = gp[j];
        jac2d_3d(jac, x, y, z, c[outer[f]], c[inner[f]], f);
        nullm(n);
        bf_lin_hex_3d(bf.a, c[0], c[1], c[2]);
        long jj = 0, kk = 1, ll = 2;
        for (long a = 0; a < nne; a++) {
          n[0][jj] = bf[a]; jj += 3;
          n[1][kk] = bf[a]; kk += 3;
          n[2][ll] = bf[a]; ll += 3;
        }
        jac = jac * w[i] * w[j];
        nnj(am.a, n.a, jac, n.m, n.n);
      }
    const double *fv = nv + 12 * f;
    if (is[f] == 1) {
      for (int m = 0; m < 4; m++)
        for (int k = 0; k < 3; k++) av[dof[f][m] + k] = fv[3 * m + k];
    }
    if (is[f] == 2) {
      for (int q = 0; q < 12; q++) av[q] = fv[q];
      vector g1(3), g2(3), g3(3);
      matrix t(3, 3);
      double norm;
      g1[0] = x[gn[f][0]] - x[gn[f][1]];
      g1[1] = y[gn[f][0]] - y[gn[f][1]];
      g1[2] = z[gn[f][0]] - z[gn[f][1]];
      g2[0] = x[gn[f][2]] - x[gn[f][3]];
      g2[1] = y[gn[f][2]] - y[gn[f][3]];
      g2[2] = z[gn[f][2]] - z[gn[f][3]];
      scprd(g1, g1, norm);
      norm = 1.0 / sqrt(norm);
      cmulv(norm, g1, g1);
      scprd(g1, g2, norm);
      g2[0] = g2[0] - norm * g1[0];
      g2[1] = g2[1] - norm * g1[1];
      g2[2] = g2[2] - norm * g1[2];
      scprd(g2, g2, norm);
      norm = 1.0 / sqrt(norm);
      cmulv(norm, g2, g2);
      g3[0] = g1[1] * g2[2] - g1[2] * g2[1];
      g3[1] = g1[2] * g2[0] - g1[0] * g2[2];
      g3[2] = g1[0] * g2[1] - g1[1] * g2[0];
      for (int r = 0; r < 3; r++) {
        t[r][0] = g1[r];
        t[r][1] = g2[r];
        t[r][2] = g3[r];
      }
      for (int q = 0; q < 4; q++) mxv(t.a, av.a + 3 * q, tnv + 3 * q, 3, 3);
      nullv(av);
      for (int m = 0; m < 4; m++)
        for (int k = 0; k < 3; k++) av[dof[f][m] + k] = tnv[3 * m + k];
    }
    mxv(am, av, v);
    addv(v, nf, nf);
  }
  for (long i = 0; i < ndofe; i++) nf_out[i] = nf[i];
}

// gtopology::cuthill_mckee_renumb (GEFEL/gtopology.cpp:6242) itself, on a gtopology
// holding only the elements (linhexa for nne 8, linquad for nne 4); it builds its
// graph with adjacnodes_edge (:2632). ordering_out[k] = the node numbered k.
void ref_cuthill_mckee(long nn, long ne, long nne, const long *conn, int reverse,
                       long *ordering_out)
{
  gtopology *gt = new gtopology();
  gt->nn = nn;
  gt->ne = ne;
  gt->gelements = new gelement[ne];
  for (long e = 0; e < ne; e++) {
    gt->gelements[e].get = nne == 8 ? linhexa : linquad;
    gt->gelements[e].nne = nne;
    gt->gelements[e].nodes = new long[nne];
    for (long a = 0; a < nne; a++) gt->gelements[e].nodes[a] = conn[e * nne + a];
  }
  gt->nodren = reverse ? rev_cuthill_mckee : cuthill_mckee;
  gt->ordering = new long[nn];
  gt->cuthill_mckee_renumb(stdout);
  for (long i = 0; i < nn; i++) ordering_out[i] = gt->ordering[i];
  delete[] gt->ordering;
  gt->ordering = NULL;
  delete gt;
}

}  // extern "C"
