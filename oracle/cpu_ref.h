/*
 * oracle/cpu_ref.h -- CPU restatement of the SIFEL linear-system hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the timed
 * CPU baseline. The product (ssf_amd/) never calls it.
 *
 * Every routine restates the reference loop order with plain double arithmetic;
 * the library is compiled with -O2 -ffp-contract=off so that, like the
 * reference's x86-64 build (no FMA), each operation rounds on its own.
 * Integers are C `long` (64-bit), exactly as the reference's `long`.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   cr_codenum            gencodnum rules; tests vs closed-form numbering and flags
 *   cr_scr_pattern        tmatstore.tex known-answer arrays; bitwise vs oracle/_ref
 *   cr_cr_pattern         tmatstore.tex known-answer arrays; bitwise vs oracle/_ref
 *   cr_*_localized        bitwise vs oracle/_ref; dense numpy assembly
 *   cr_mxv_scr/cr         bitwise vs oracle/_ref; dense numpy product
 *   cr_scr_cg_prec        bitwise (x, ani, ares) vs oracle/_ref; scipy direct solve;
 *                         diagonal matrix converges in one iteration
 *   cr_scr_cg, cr_cr_cg_prec_noprec   bitwise vs oracle/_ref; scipy direct solve
 *   cr_hex8_elast_ke      bitwise vs oracle/_ref; symmetry, 6 rigid-body modes,
 *                         patch test (uniform strain energy closed form)
 *   cr_hex8_heat_ke       bitwise vs oracle/_ref; constants in kernel, closed form
 *                         unit-cube matrix, linear temperature patch test
 *   cr_quad4_pstress_ke   bitwise vs oracle/_ref; 3 rigid-body modes, patch test
 *   cr_eldispl            code-number rule on a hand-built table (all three branches)
 *   cr_*_ip, cr_elem_results_mesh
 *                         bitwise vs oracle/_ref (fixtures from the reference's
 *                         solutions); patch tests (exact constant strain / gradient
 *                         under linear fields), rigid-body motions give zero strain
 *                         and force, element forces = K_e r with the pinned K_e,
 *                         assembled forces = A x - lift, D / -k closed forms
 *   cr_hex8_heat_cap      bitwise vs oracle/_ref; unit-cube closed form (exact with
 *                         2x2x2 points); sum of entries = rho c V
 *   cr_scr_incomplete_ldl, cr_scr_prec_ildl, cr_scr_cg_prec_ildl
 *                         bitwise vs oracle/_ref (symcomprow::incomplete_ldl,
 *                         prec_ildl_scr, cg_prec with tprec = 10; gamma 0 and not);
 *                         a pattern without fill gives the exact LDL^T (dense and
 *                         tridiagonal: L D L^T = A, the preconditioned solve is
 *                         exact, cg_prec converges in one iteration); the product
 *                         L D L^T equals A on the pattern
 *   cr_mtxv_cr, cr_cr_bicg
 *                         bitwise vs oracle/_ref (comprow::mtxv_cr, comprow::bicg);
 *                         transpose identity b.(A^T c) = (A b).c; on a symmetric
 *                         matrix BiCG reaches the CG solution; on small dense
 *                         nonsymmetric systems the solution of a direct solve
 *   cr_hex8_surface_load, cr_surface_load_mesh
 *                         bitwise vs oracle/_ref (GEFEL jac2d_3d, bf_lin_hex_3d, nnj,
 *                         mxv, addv, scprd, cmulv in the loop order of
 *                         linhex::node_forces_surf / locglob_nodeval); a uniform
 *                         pressure on a rectangular face gives pA/4 per node; the
 *                         local-axes form equals the global form with the rotated
 *                         values; the loads integrate the traction (total = int t dA)
 *   cr_edge_adjacency, cr_cuthill_mckee
 *                         the ordering equals oracle/_ref's (the reference's own
 *                         gtopology::adjacnodes_edge + cuthill_mckee_renumb, both
 *                         forms); a path graph gives the path order; the ordering
 *                         is a permutation and a BFS (levels never decrease); the
 *                         profile of the renumbered matrix does not exceed a
 *                         shuffled numbering's on a mesh
 *   cr_newton_raphson     not run against the reference itself (gnewton_raphson
 *                         needs MEFEL's global problem objects, which the GEFEL-only
 *                         harness does not build); pinned by closed forms: the
 *                         attained displacements equal the direct solution at lambda
 *                         (linear elasticity, superposition of fc and fp), the
 *                         increment sequence (dlambda doubled after every second
 *                         accepted increment, capped at maxincrnr, cut at lambdar),
 *                         rejection and halving down to minincrnr when errnr is
 *                         below the solver's reach, the rel_load / absolute norms
 *   cr_np_step_linear, cr_np_step_dform
 *                         bitwise vs oracle/_ref (the reference's own symcomprow
 *                         mxv_scr / scalmat_scr / addmat_scr / cg_prec); amplification
 *                         factor of the trapezoidal rule on one unknown; steady state;
 *                         the two forms agree to rounding on the same problem
 *   cr_hex8_volume_load, cr_volume_load_mesh
 *                         bitwise vs oracle/_ref (GEFEL jac_3d, nnj, mxv in the loop
 *                         order of load_matrix / quantity_source_vector); constant
 *                         load on a parallelepiped: b V / 8 per node; totals = b V
 */
#ifndef SSF_ORACLE_CPU_REF_H
#define SSF_ORACLE_CPU_REF_H

#ifdef __cplusplus
extern "C" {
#endif

/* GEFEL/gtopology.cpp:1619-1680 gencodnum. dof[nn*ndofn] holds the flags on entry
 * (0 fixed, 1 free, k>1 coupled group, <0 prescribed) and the code numbers on exit.
 * ordering may be NULL (identity). Returns the number of unknowns. */
long cr_codenum(long nn, long ndofn, long *dof, const long *ordering);

/* GEFEL/gtopology.cpp:597-626 give_code_numbers for node-based code numbers. */
void cr_elem_codes(long ne, long nne, const long *conn, long ndofn, const long *dof,
                   long *cn);

/* GEFEL/scr.cpp:556-570 symcomprow::initiate (numcontr_elem :112, addresses :211,
 * fillarray_elem :163, sort_and_colindex_tko :317). adr[n+1] is always written;
 * ci is written when ci != NULL and cap >= result. Returns negm. */
long cr_scr_pattern(long n, long ne, long ndofe, const long *cn, long *adr, long *ci,
                    long cap);
/* GEFEL/cr.cpp:454-470 comprow::initiate (sort_and_colindex :247). */
long cr_cr_pattern(long n, long ne, long ndofe, const long *cn, long *adr, long *ci,
                   long cap);

/* GEFEL/scr.cpp:472-521 symcomprow::localized and GEFEL/cr.cpp:368-410
 * comprow::localized: a[] += element matrix lm[nc*nc] (row-major). */
void cr_scr_localized(long n, const long *adr, const long *ci, double *a,
                      const double *lm, const long *cn, long nc);
void cr_cr_localized(long n, const long *adr, const long *ci, double *a,
                     const double *lm, const long *cn, long nc);
/* element loop of MEFEL/SRC/globmat.cpp:41-142 / TRFEL/SRC/globmatt.cpp:19-67:
 * localize every element matrix ke[ne][nc*nc] in element order. */
void cr_scr_assemble(long n, const long *adr, const long *ci, double *a, long ne,
                     long nc, const long *cn, const double *ke);
void cr_cr_assemble(long n, const long *adr, const long *ci, double *a, long ne,
                    long nc, const long *cn, const double *ke);

/* GEFEL/scr.cpp:655-670 mxv_scr and GEFEL/cr.cpp:653-700 mxv_cr (serial branch). */
void cr_mxv_scr(long n, const long *adr, const long *ci, const double *a,
                const double *b, double *c);
void cr_mxv_cr(long n, const long *adr, const long *ci, const double *a,
               const double *b, double *c);

/* GEFEL/scr.cpp:1093-1203 cg_prec with tprec=1 (Jacobi, prec_diag_scr :897).
 * Returns 1 when the stopping test fired, 2 on the zero-denominator exit,
 * 0 when ni was exhausted (ani/ares then keep their input values, as in the
 * reference). */
int cr_scr_cg_prec(long n, const long *adr, const long *ci, const double *a, double *x,
                   const double *y, long ni, double res, long *ani, double *ares,
                   double zero, long iv);
void cr_scr_incomplete_ldl(long n, const long *adr, const long *ci, const double *a,
                           double gamma, double *incdec);
void cr_scr_prec_ildl(long n, const long *adr, const long *ci, const double *incdec,
                      double *x, const double *y);
int cr_scr_cg_prec_ildl(long n, const long *adr, const long *ci, const double *a,
                        const double *incdec, double *x, const double *y, long ni, double res,
                        long *ani, double *ares, double zero, long iv);
/* GEFEL/scr.cpp:786-880 symcomprow::cg (no preconditioner, sqrt test). */
int cr_scr_cg(long n, const long *adr, const long *ci, const double *a, double *x,
              const double *y, long ni, double res, long *ani, double *ares,
              double zero, long iv);
/* GEFEL/cr.cpp:1124-1232 comprow::cg_prec with a noprecond precond (h = r). */
int cr_cr_cg_prec_noprec(long n, const long *adr, const long *ci, const double *a,
                         double *x, const double *y, long ni, double res, long *ani,
                         double *ares, double zero, long iv);

/* MEFEL/SRC/linhex.cpp:583-640 gl_stiffness_matrix, 24x24 row-major. */
void cr_hex8_elast_ke(const double *x, const double *y, const double *z, double e,
                      double nu, double *ke);
/* TRFEL/SRC/linhext.cpp:774-840 conductivity_matrix, 8x8 row-major. */
void cr_hex8_heat_ke(const double *x, const double *y, const double *z, double k,
                     double *ke);
/* MEFEL/SRC/plelemlq.cpp:611-690 gl_stiffness_matrix, 8x8 row-major; t = nodal
 * thickness (4 values). */
void cr_quad4_pstress_ke(const double *x, const double *y, double e, double nu,
                         const double *t, double *ke);

/* Batched element matrices for a mesh: coordinates per node (xyz, stride 3 or 2),
 * connectivity conn[ne*nne] 0-based, per-element material. */
void cr_hex8_elast_ke_mesh(long ne, const long *conn, const double *xyz,
                           const double *e, const double *nu, double *ke);
void cr_hex8_heat_ke_mesh(long ne, const long *conn, const double *xyz,
                          const double *k, double *ke);
void cr_quad4_pstress_ke_mesh(long ne, const long *conn, const double *xy,
                              const double *e, const double *nu, const double *t,
                              double *ke);

/* TRFEL/SRC/globmatt.cpp:1635-1700 assemble_init for a stationary problem with
 * zero initial values: rhs += -(K_e r_e) with r_e the prescribed values pv[-cn-1]
 * of the element's prescribed DOFs (MATRIX.cpp:3500 locglob scatter). */
void cr_dirichlet_lift(long ne, long nc, const long *cn, const double *ke,
                       const double *pv, double *rhs);

/* ---- post-solve element pass (strains/stresses or gradients/fluxes at the
 * integration points, internal forces / fluxes) */

/* MEFEL/SRC/globmat.cpp:1305-1311 eldispl and TRFEL/SRC/globmatt.cpp:350-363
 * select_nodalval: r[i] = lhs[cn-1] (cn > 0), pv[-cn-1] (cn < 0), 0 (cn = 0). */
void cr_eldispl(long nc, const long *cn, const double *lhs, const double *pv, double *r);
/* MEFEL/SRC/linhex.cpp:1110-1145, elastisomat.cpp:627-651, linhex.cpp:2157-2191:
 * eps/sig [8][6] (x, y, z, yz, xz, xy; engineering shears), element forces f[24]. */
void cr_hex8_elast_ip(const double *x, const double *y, const double *z, double e,
                      double nu, const double *r, double *eps, double *sig, double *f);
/* MEFEL/SRC/plelemlq.cpp:1173-1205, elastisomat.cpp:627-651, plelemlq.cpp:2307-2359:
 * eps/sig [4][4] (x, y, xy, z), f[8]; t = nodal thickness. */
void cr_quad4_pstress_ip(const double *x, const double *y, double e, double nu,
                         const double *t, const double *r, double *eps, double *sig,
                         double *f);
/* TRFEL/SRC/linhext.cpp:432-498, transmat.cpp:1123-1148, linhext.cpp:1086-1146:
 * grad/flux [8][3], internal fluxes f[8]. */
void cr_hex8_heat_ip(const double *x, const double *y, const double *z, double k,
                     const double *r, double *grad, double *flux, double *f);
/* element loop of compute_ipstrains / compute_ipstresses / loc_internal_forces
 * (MEFEL/SRC/elemswitch.cpp:41,1215, globmat.cpp:506-650) and of compute_ipgrads /
 * compute_ipfluxes / internal_fluxes (TRFEL/SRC/elemswitcht.cpp:1260,1111,1845);
 * kind 0 hex8 elasticity, 1 hex8 heat, 2 quad4 plane stress. */
void cr_elem_results_mesh(int kind, long ne, const long *conn, const double *xyz,
                          const long *cn, const double *ma, const double *mb,
                          const double *th, long n, const double *lhs, const double *pv,
                          double *eps, double *sig, double *ifor);

/* ---- nonstationary heat */
/* TRFEL/SRC/linhext.cpp:979-1029 capacity_matrix, 8x8 row-major; dens = nodal
 * densities (8), c = specific heat (isotrmat::matcap). */
void cr_hex8_heat_cap(const double *x, const double *y, const double *z, double c,
                      const double *dens, double *cm);
void cr_hex8_heat_cap_mesh(long ne, const long *conn, const double *xyz, const double *c,
                           const double *rho, double *cm);
/* TRFEL/SRC/npsolvert.cpp:367-482 one_step_linear on symcomprow K and C (same
 * pattern); returns the cg_prec status (see the definition for the steps). */
void cr_mtxv_cr(long n, const long *adr, const long *ci, const double *a,
                const double *b, double *c);
int cr_cr_bicg(long n, const long *adr, const long *ci, const double *a, double *x,
               const double *y, long ni, double res, long *ani, double *ares, double zero,
               long iv);
double cr_newton_raphson(long n, const long *adr, const long *ci, const double *a, int kind,
                         long ne, const long *conn, const double *xyz, const long *cn,
                         const double *ma, const double *mb, const double *th,
                         const double *pv, const double *fc, const double *fp, double *fa,
                         double *ra, long li, double ilambda, long ninr, long niilnr,
                         double errnr, double incrnr, double minincrnr, double maxincrnr,
                         int normt, int check_lambdar, double lambdar, long sni, double sres,
                         double szero, double *lam, double *dlam, long *nsol, long *acc,
                         long *counts);
void cr_hex8_surface_load(const double *x, const double *y, const double *z, const long *ls,
                          const double *nodvals, double *nf);
void cr_surface_load_mesh(long nrec, const long *eid, const long *ls, const double *nodvals,
                          const long *conn, const double *xyz, const long *cn, double *rhs);
void cr_edge_adjacency(long nn, long ne, long nne, const long *conn, long *ptr, long *adj);
int cr_cuthill_mckee(long nn, const long *ptr, long *adj, int reverse, long *ordering);
int cr_np_step_linear(long n, const long *adr, const long *ci, const double *aK,
                      const double *aC, double *aA, double alpha, double dt,
                      const double *rhs, double *lhs, double *tdlhs, double *d, double *p,
                      double *f, long ni, double res, double zero, long *ani, double *ares);
int cr_np_step_dform(long n, const long *adr, const long *ci, const double *aK,
                     const double *aC, double *aA, double alpha, double dt,
                     const double *rhs, double *lhs, double *tdlhs, double *d, double *p,
                     double *f, long ni, double res, double zero, long *ani, double *ares);

/* ---- element volume loads (MEFEL loadel::volumeload with linhex::load_matrix,
 * kind 0) and heat sources (TRFEL linhext::quantity_source_vector, kind 1) */
void cr_hex8_volume_load(int kind, const double *x, const double *y, const double *z,
                         const double *nodval, double *nf);
void cr_volume_load_mesh(int kind, long ne, const long *conn, const double *xyz, const long *cn,
                         const double *nodval, double *rhs);

#ifdef __cplusplus
}
#endif
#endif
