/*
 * ssf_gpu.h -- C ABI of the MI355X (gfx950) linear-system path of SIFEL.
 *
 * The path: element stiffness / conductivity evaluation, assembly ("localize")
 * into GEFEL compressed-row storage (symcomprow = symmetric lower rows with the
 * diagonal last, comprow = full rows), and the conjugate-gradient solve behind
 * gmatrix::solve_system. Host callers keep the gmatrix API (GEFEL/gmatrix.h:29-153);
 * a gmatrix built on this library forwards each member below to one entry point.
 *
 * Conventions: plain pointers and sizes, 64-bit integers as the reference's
 * `long`, code numbers 1-based with entries <= 0 skipped (GEFEL/scr.cpp:427-433).
 * Every function returns SSF_OK (0) or a negative error code; ssf_last_error()
 * gives the message (per calling thread). Host pointers are read/written
 * synchronously; *_dev entry points take device pointers and are asynchronous
 * on the context's stream. A context is used by one host thread at a time.
 */
#ifndef SSF_GPU_H
#define SSF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SSF_OK 0
#define SSF_EINVAL -1     /* bad argument (size, pointer, storage type)   */
#define SSF_EHIP -2       /* HIP runtime error                           */
#define SSF_ENOMEM -3     /* device memory or HBM budget exhausted        */
#define SSF_ESTATE -4     /* call out of order (e.g. solve before values) */
#define SSF_EUNSUP -5     /* combination the reference does not support   */

/* identifiers kept from GEFEL/galias.h:59-62,74-77,103 */
#define SSF_STORAGE_COMPRESSED_ROWS 10 /* comprow            */
#define SSF_STORAGE_SYMM_COMP_ROWS 11  /* symcomprow        */
#define SSF_SOLVER_CG 20               /* linsolvertype cg   */
#define SSF_SOLVER_BICG 30             /* linsolvertype bicg */
#define SSF_PREC_NOPRECOND 0           /* precondtype        */
#define SSF_PREC_DIAG 1                /* diagprec (Jacobi)  */
#define SSF_PREC_ILDL 10               /* incomdec (incomplete LDL^T, symcomprow) */

/* element kinds of the batched element kernels */
#define SSF_ELEM_HEX8_ELAST 0  /* MEFEL linhex, 24x24, 3 DOF/node            */
#define SSF_ELEM_HEX8_HEAT 1   /* TRFEL linhext with isotrmat, 8x8            */
#define SSF_ELEM_QUAD4_PSTRESS 2 /* MEFEL planeelemlq, plane stress, 8x8      */
#define SSF_ELEM_HEX8_HEAT_CAP 3 /* TRFEL linhext capacity matrix, 8x8:
                                    mat_a = specific heat c, mat_b = density
                                    (may be NULL = 1)                          */

/* arithmetic of a solve */
#define SSF_PRECISION_FP64 0   /* plain FP64 PCG, recurrence of the reference   */
#define SSF_PRECISION_MIXED 1  /* FP32 inner PCG inside FP64 iterative refinement */

typedef struct ssf_ctx ssf_ctx;
typedef struct ssf_matrix ssf_matrix;

typedef struct ssf_solve_params {
  int32_t linsolver;   /* SSF_SOLVER_CG, SSF_SOLVER_BICG (compressed_rows)        */
  int32_t prec;        /* SSF_PREC_NOPRECOND / SSF_PREC_DIAG / SSF_PREC_ILDL       */
  int64_t ni;          /* maximum number of iterations (gmatrix::ni)              */
  double res;          /* required residual norm (gmatrix::res); squared test for
                          symcomprow + diagprec as GEFEL/scr.cpp:1188             */
  double zero;         /* computer zero (gmatrix::zero, 1e-30)                    */
  int32_t iv;          /* 0: start from x = 0, 1: start from lhs                  */
  int32_t precision;   /* SSF_PRECISION_*                                         */
  double inner_res;    /* mixed: delta, relative FP32 residual reduction between two
                          FP64 refreshes (0 = 1e-4)                              */
  int64_t inner_ni;    /* mixed: FP32 iterations between two FP64 refinements
                          (reliable updates); 0 = 128 (single systems), 32 (batch) */
  int64_t max_outer;   /* reserved (0)                                            */
} ssf_solve_params;

typedef struct ssf_solve_info {
  int64_t ani;         /* as gmatrix::ani (left unchanged when ni is exhausted)   */
  double ares;         /* as gmatrix::ares                                        */
  int32_t status;      /* 1 converged, 2 zero denominator, 3 zero rhs, 0 ni spent,
                          4 mixed: FP64 residual floor reached above res */
  int32_t pad;
  int64_t iterations;  /* PCG iterations actually run (FP32 ones for mixed)      */
  int64_t outer;       /* mixed: FP64 refinements (residual recomputations)       */
  double rel_residual; /* final ||b - A x|| / ||b|| in FP64                        */
} ssf_solve_info;

/* ------------------------------------------------------------------ context */
const char *ssf_last_error(void);
int ssf_ctx_create(int device, ssf_ctx **out);
int ssf_ctx_destroy(ssf_ctx *ctx);
int ssf_ctx_sync(ssf_ctx *ctx);
/* the context's HIP stream (hipStream_t), for callers that enqueue around it */
void *ssf_ctx_stream(ssf_ctx *ctx);
/* Kernel sampling inside solves (on = 1 enables and resets): the second SpMV of
 * every iteration chunk is bracketed by its own start / end events
 * (hipExtLaunchKernelGGL) on the solve stream.
 * ssf_ctx_kernel_time returns the mean duration (us) and sample count of the
 * FP64 (which = 0) or FP32 (which = 1) operator since the last reset. */
int ssf_ctx_set_profiling(ssf_ctx *ctx, int on);
int ssf_ctx_kernel_time(ssf_ctx *ctx, int which, double *avg_us, int64_t *count);

/* ------------------------------------------------------------ DOF numbering */
/* gtopology::gencodnum (GEFEL/gtopology.cpp:1619-1665), on the host: dof[nn*ndofn]
 * holds the node flags on entry (0 fixed, 1 free, k > 1 coupled group k, < 0
 * prescribed value index) and the code numbers on return: free DOFs get
 * 1, 2, ... in the node order `ordering` (NULL = identity; else a permutation,
 * e.g. from ssf_cuthill_mckee), each coupled group one number at its first DOF,
 * 0 and negative flags unchanged. *ndof = the number of unknowns. */
int ssf_gencodnum(int64_t nn, int64_t ndofn, int64_t *dof, const int64_t *ordering,
                  int64_t *ndof);
/* gtopology::give_code_numbers (GEFEL/gtopology.cpp:597-626, code numbers on the
 * nodes): cn[ne][nne*ndofn] = the codes of element e's nodes conn[e*nne + a]
 * (0-based) in node order, DOFs of a node in their order. */
int ssf_give_code_numbers(int64_t ne, int64_t nne, const int64_t *conn, int64_t nn,
                          int64_t ndofn, const int64_t *dof, int64_t *cn);

/* ------------------------------------------------------ node renumbering */
/* gtopology::cuthill_mckee_renumb (GEFEL/gtopology.cpp:6242-6355) on the edge graph
 * of adjacnodes_edge (:2632) of nn nodes and ne hex8 (nne 8) or quad4 (nne 4)
 * elements conn[ne*nne] (0-based): ordering[k] = the node numbered k (reverse =
 * 1: rev_cuthill_mckee, the order reversed), equal to the reference's ordering.
 * gencodnum then numbers the DOFs by visiting the nodes in this order. The graph
 * must be connected (SSF_EINVAL otherwise). Built on the device level by level;
 * synchronous. */
int ssf_cuthill_mckee(ssf_ctx *ctx, int64_t nn, int64_t ne, int64_t nne, const int64_t *conn,
                      int reverse, int64_t *ordering);

/* ------------------------------------------------------------ storage setup */
/* symcomprow::initiate / comprow::initiate (GEFEL/scr.cpp:556, GEFEL/cr.cpp:454,
 * reached through gmatrix::initiate GEFEL/gmatrix.cpp:268): symbolic pattern
 * from element code numbers cn[ne*ndofe]. storage = SSF_STORAGE_*. */
int ssf_matrix_create_elem(ssf_ctx *ctx, int storage, int64_t n, int64_t ne,
                           int64_t ndofe, const int64_t *cn, ssf_matrix **out);
/* storage from existing reference arrays adr[n+1], ci[adr[n]] (0-based columns;
 * symcomprow rows ascending with the diagonal last). */
int ssf_matrix_create_csr(ssf_ctx *ctx, int storage, int64_t n, const int64_t *adr,
                          const int64_t *ci, ssf_matrix **out);
/* symcomprow::initiate with the pattern built on the device: the same adr / ci,
 * from a sort of the elements' DOF pairs (no host pattern work). The matrix keeps
 * the SELL-64 layout only (16/32-bit columns, no row patterns): it assembles,
 * multiplies and solves in FP64 and in mixed precision; batches and the stream /
 * row-vector SpMV need a matrix from ssf_matrix_create_elem. */
int ssf_matrix_create_elem_device(ssf_ctx *ctx, int64_t n, int64_t ne, int64_t ndofe,
                                  const int64_t *cn, ssf_matrix **out);
int ssf_matrix_destroy(ssf_matrix *m);
/* number of stored entries in the reference layout (negm) and n */
int ssf_matrix_sizes(const ssf_matrix *m, int64_t *n, int64_t *negm, int64_t *nnz_full);
/* reference-layout arrays adr[n+1], ci[negm] (either may be NULL) */
int ssf_matrix_pattern(const ssf_matrix *m, int64_t *adr, int64_t *ci);

/* --------------------------------------------------------------- values */
/* symcomprow::nullmat / comprow::nullmat (GEFEL/scr.cpp:100) */
int ssf_matrix_zero(ssf_matrix *m);
/* a[negm] in the reference layout -> device (replaces the host-side array a) */
int ssf_matrix_set_values(ssf_matrix *m, const double *a);
/* device -> a[negm] in the reference layout */
int ssf_matrix_get_values(const ssf_matrix *m, double *a);
/* gmatrix::localized (GEFEL/gmatrix.cpp:581) for a batch of element matrices:
 * lm[ne][nc*nc] row-major added into the device values; every entry receives its
 * contributions in localize's order (element, row, column; symcomprow keeps
 * ii>=jj, comprow all; code numbers <= 0 skipped). cn[ne*nc] 1-based; code
 * numbers may repeat within and across elements (coupled DOFs). */
int ssf_matrix_localize_batch(ssf_matrix *m, int64_t ne, int64_t nc, const int64_t *cn,
                              const double *lm);
/* Batched element kernel + atomic-free assembly by entry into the device values,
 * each entry summed in localize's order (stiffness_matrix MEFEL/SRC/globmat.cpp:41 / conductivity_matrix
 * TRFEL/SRC/globmatt.cpp:19 with linhex / linhext / planeelemlq kernels).
 * xyz[nn*3] node coordinates (z ignored for quad4), conn[ne*nne] 0-based nodes,
 * cn[ne*ndofe] code numbers, mat_a = E (elasticity) or k (heat), mat_b = nu,
 * thick = thickness (quad4, may be NULL = 1). The pattern must have been built
 * from the same cn. rhs_lift (may be NULL, n doubles, host) receives
 * -K_e u_p for prescribed DOFs (cn<0 -> pv[-cn-1]) as assemble_init
 * (TRFEL/SRC/globmatt.cpp:1635) computes. Values are accumulated (call
 * ssf_matrix_zero first for a fresh assembly). Kind SSF_ELEM_HEX8_HEAT_CAP
 * assembles the capacity matrix of linhext (TRFEL/SRC/linhext.cpp:979). */
int ssf_matrix_assemble(ssf_matrix *m, int kind, int64_t nn, const double *xyz, int64_t ne,
                        const int64_t *conn, const int64_t *cn, const double *mat_a,
                        const double *mat_b, const double *thick, const double *pv,
                        int64_t npv, double *rhs_lift);

/* gmatrix::scalgm (GEFEL/gmatrix.cpp:1848; symcomprow::scalmat_scr
 * GEFEL/scr.cpp:760, comprow::scalmat_cr GEFEL/cr.cpp:996): a *= c */
int ssf_matrix_scale(ssf_matrix *m, double c);
/* gmatrix::addgm (GEFEL/gmatrix.cpp:1791; symcomprow::addmat_scr GEFEL/scr.cpp:682,
 * comprow::addmat_cr GEFEL/cr.cpp:981): a += c * other.a. Both matrices must have
 * the same storage and the same pattern (SSF_EINVAL otherwise). */
int ssf_matrix_add(ssf_matrix *m, double c, const ssf_matrix *other);

/* -------------------------------------------------------------- operations */
/* gmatrix::gmxv (GEFEL/gmatrix.cpp:1704): b = A a, host vectors of length n */
int ssf_gmxv(ssf_matrix *m, const double *a, double *b);
/* same on device pointers, asynchronous */
int ssf_gmxv_dev(ssf_matrix *m, const double *a_dev, double *b_dev);
/* gmatrix::solve_system (GEFEL/gmatrix.cpp:1049), dispatched on the solver,
 * storage and preconditioner as the reference does:
 *   cg (20): symcomprow + diagprec -> symcomprow::cg_prec GEFEL/scr.cpp:1093;
 *     symcomprow + noprecond -> symcomprow::cg :786; comprow ->
 *     comprow::cg_prec GEFEL/cr.cpp:1124 with the noprecond preconditioner.
 *     symcomprow + incomdec (10) -> symcomprow::incomplete_ldl GEFEL/scr.cpp:990
 *     with weight 0 (gmatrix.cpp:24,1234-1240), then cg_prec :1093 with
 *     prec_ildl_scr :1031; FP64. The factor is formed at the first such solve and
 *     again after every change of the values (the reference factors once); its
 *     rows run by dependency levels, each row in the reference's operation order.
 *   bicg (30): comprow only -> comprow::bicg GEFEL/cr.cpp:1353 (gmatrix.cpp:
 *     1273-1289), FP64, no preconditioner (prec is ignored, as there). Stops when
 *     r.r / y.y < res and rr.rr / y.y < res (:1400): status 1, ani = the
 *     iteration index, ares = r.r. y.y < zero: status 3, lhs untouched, ani 0,
 *     ares = y.y. ni spent: status 0, ani = ni, ares = the last r.r. Where the
 *     reference would divide by a zero dd.p or r.rr and carry NaN to ni, the
 *     device stops with status 2. The first bicg solve of a matrix lays out the
 *     union pattern of A and A^T on the device (kept with the matrix); its values
 *     are re-taken after every change of the matrix's values.
 * lhs/rhs host vectors of length n. */
int ssf_solve(ssf_matrix *m, const ssf_solve_params *p, const double *rhs, double *lhs,
              ssf_solve_info *info);
/* same with device vectors (rhs_dev read-only, lhs_dev in/out), synchronous */
int ssf_solve_dev(ssf_matrix *m, const ssf_solve_params *p, const double *rhs_dev,
                  double *lhs_dev, ssf_solve_info *info);
/* ssf_solve_dev without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it). The FP64 cg solver only
 * (symcomprow with diagprec or noprecond, comprow with noprecond), with the
 * results of ssf_solve_dev bit for bit, for n <= SSF_NOALLOC_MAX_N (SSF_EINVAL
 * beyond). Its working vectors, reduction partials and recurrence scalars live in
 * one device array of the library (4 SSF_NOALLOC_MAX_N + 32784 doubles, about
 * 80 MB), which each GPU the library is loaded on holds until the process ends.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU race.
 * Synchronous; no hipMalloc, hipFree or memset. */
#define SSF_NOALLOC_MAX_N 2500000
int ssf_solve_dev_noalloc(ssf_matrix *m, const ssf_solve_params *p, const double *rhs_dev,
                          double *lhs_dev, ssf_solve_info *info);
/* symcomprow::incomplete_ldl (GEFEL/scr.cpp:990-1022) with weight 0, as gmatrix
 * runs it (GEFEL/gmatrix.cpp:24,1237): the factor incdec[negm] of the current
 * values in the reference layout (L off the diagonal, D on it), copied to the
 * host. symm_comp_rows only. */
int ssf_matrix_ildl_factor(ssf_matrix *m, double *incdec);
/* symcomprow::prec_ildl_scr (GEFEL/scr.cpp:1031-1060): x_dev = (L D L^T)^-1 y_dev
 * with the factor of the current values (formed if needed); asynchronous. */
int ssf_ildl_apply_dev(ssf_matrix *m, const double *y_dev, double *x_dev);

/* one_step_linear (TRFEL/SRC/npsolvert.cpp:367-482): one step of the generalized
 * trapezoidal rule in the v-form for C dT/dt + K T = f, on device vectors:
 *   d = lhs + (1 - alpha) dt tdlhs;  p = K d;  A = K * (dt alpha) + 1.0 * C;
 *   f = rhs - p;  A tdlhs = f by ssf_solve_dev from the current tdlhs (the
 *   iv field of p is ignored: the previous rate is the start vector);
 *   lhs = d + alpha dt tdlhs.
 * K (conductivity), C (capacity) and A (the system matrix, overwritten) share
 * one context, storage and pattern. A is re-formed only when K's or C's values
 * or dt * alpha change since the last step. rhs_dev = f_{n+1} (prescribed
 * fluxes and the Dirichlet lift); lhs_dev = nodal values, tdlhs_dev = their time
 * derivatives, both updated. Synchronous. */
int ssf_np_step_linear(ssf_matrix *K, ssf_matrix *C, ssf_matrix *A, double alpha, double dt,
                       const ssf_solve_params *p, const double *rhs_dev, double *lhs_dev,
                       double *tdlhs_dev, ssf_solve_info *info);

/* One step of linear_nonstat_solv_dform (TRFEL/SRC/npsolvert.cpp:1071-1227, the
 * loop body :1152-1185): the generalized trapezoidal rule in the d-form, the
 * scheme the reference selects for tnpsolver = trapezoidd and for discontinuous
 * problems (:19-31):
 *   d = lhs + (1 - alpha) dt tdlhs;  p = C d;  A = K * (dt alpha) + 1.0 * C;
 *   f = rhs alpha dt + p;  A lhs = f by ssf_solve_dev from the current lhs (the
 *   iv field of p is ignored);  tdlhs = (lhs - d) / dt / alpha.
 * Arguments, A's caching and the errors are those of ssf_np_step_linear; alpha
 * and dt must be nonzero. Synchronous. */
int ssf_np_step_dform(ssf_matrix *K, ssf_matrix *C, ssf_matrix *A, double alpha, double dt,
                      const ssf_solve_params *p, const double *rhs_dev, double *lhs_dev,
                      double *tdlhs_dev, ssf_solve_info *info);

/* ------------------------------------------------------------------ batches */
/* Many assembled systems solved together, each with nrhs right-hand sides
 * (1, 2, 4 or 8). Replaces the loop of gmatrix::solve_system calls over the
 * load cases of one matrix (MEFEL/SRC/lssolver.cpp:78-79) and over independent
 * meshes: one launch sequence serves every (system, rhs) pair, each pair runs
 * the symcomprow::cg_prec recurrence (GEFEL/scr.cpp:1093-1203, Jacobi) on its
 * own and stops on its own test. The matrices' values are copied at creation
 * (later changes to them are not seen); the matrices may be destroyed after.
 * Vectors are interleaved: entry (system s, row i, rhs r) is at
 * (row0[s] + i) * nrhs + r, row0 from ssf_batch_sizes. */
typedef struct ssf_batch ssf_batch;
int ssf_batch_create(ssf_ctx *ctx, int64_t nsys, ssf_matrix *const *mats, int32_t nrhs,
                     ssf_batch **out);
int ssf_batch_destroy(ssf_batch *b);
/* total rows, stored entries, and row0[nsys + 1] (may be NULL) */
int ssf_batch_sizes(const ssf_batch *b, int64_t *nrows, int64_t *nnz, int64_t *row0);
/* solve every pair: p as for ssf_solve_dev (cg + diagprec; FP64 or mixed), infos
 * receives nsys * nrhs results in pair order s * nrhs + r; the result of a pair
 * does not depend on the other pairs of the batch. Synchronous. */
int ssf_batch_solve_dev(ssf_batch *b, const ssf_solve_params *p, const double *rhs_dev,
                        double *lhs_dev, ssf_solve_info *infos);
/* same with host vectors */
int ssf_batch_solve(ssf_batch *b, const ssf_solve_params *p, const double *rhs, double *lhs,
                    ssf_solve_info *infos);
/* time the batched SpMM formats for precision 0 (FP64) / 1 (FP32): SELL-64
 * slices (format 2, the default) and the CSR row-vector kernel with R..64
 * lanes per row (format 0); keep the fastest. tune_out (may be NULL) receives
 * {format, lanes}. */
int ssf_batch_autotune(ssf_batch *b, int precision, int reps, int32_t *tune_out);

/* ------------------------------------------------------------------ serving */
/* A request loop for element-mesh solves on one GPU: requests are admitted
 * under an HBM budget and solved with continuous batching (requests join and
 * leave the resident set between iteration chunks). Each request runs what a
 * SIFEL problem does for one load case: gmatrix::initiate (pattern,
 * GEFEL/gmatrix.cpp:268), stiffness_matrix / conductivity_matrix assembly
 * (MEFEL/SRC/globmat.cpp:41, TRFEL/SRC/globmatt.cpp:19) with the Dirichlet lift,
 * and gmatrix::solve_system with symcomprow::cg_prec (GEFEL/scr.cpp:1093, Jacobi,
 * FP64). Its result does not depend on the other requests in flight. Matrices
 * and solver state stay on the device until the request is finished. */
typedef struct ssf_server ssf_server;
/* p: cg + diagprec, precision FP64, iv 0 (ni, res, zero used); chunk: iterations
 * between admissions (0 = 32); max_slots: resident requests at most (0 = 1024);
 * prep_threads: host threads building and assembling requests (0 = 2). */
int ssf_server_create(int device, int64_t budget_bytes, const ssf_solve_params *p,
                      int32_t chunk, int32_t max_slots, int32_t prep_threads,
                      ssf_server **out);
/* stops accepting, finishes every request in flight, releases everything */
int ssf_server_destroy(ssf_server *s);
/* one request, arguments as ssf_matrix_assemble plus the number of unknowns n
 * and the load vector rhs[n] (the lift of prescribed values is added); the
 * arrays are copied. ticket receives the request's id. */
int ssf_server_submit(ssf_server *s, int kind, int64_t n, int64_t nn, const double *xyz,
                      int64_t ne, const int64_t *conn, const int64_t *cn, const double *mat_a,
                      const double *mat_b, const double *thick, const double *pv, int64_t npv,
                      const double *rhs, int64_t *ticket);
/* blocks until the request is finished; x[n] and info (either may be NULL) get
 * its result; times[4] (may be NULL): submitted, admitted under the budget,
 * assembled, finished, in seconds since the server was created */
int ssf_server_wait(ssf_server *s, int64_t ticket, double *x, ssf_solve_info *info,
                    double *times);
/* stats[8]: finished requests, loop steps, peak reserved bytes, reserved bytes,
 * the microseconds spent by the prep threads on patterns, on matrix set-up and
 * on device assembly, summed over requests, and the requests whose solve
 * streams the symmetric half of the stencil form */
int ssf_server_stats(ssf_server *s, int64_t *stats);

/* ------------------------------------------------------- post-solve element pass */
/* Results at the integration points and the internal forces, from a solution.
 * Replaces the element loops run after the solve: compute_ipstrains
 * (MEFEL/SRC/elemswitch.cpp:41), compute_ipstresses (:1215) and
 * loc_internal_forces (MEFEL/SRC/globmat.cpp:506) for the elastic kinds, with
 * linhex::gl_ip_strains / elem_integration (MEFEL/SRC/linhex.cpp:1110,2157),
 * planeelemlq::gl_ip_strains / elem_integration (MEFEL/SRC/plelemlq.cpp:1173,2307)
 * and elastisomat::nlstresses (MEFEL/SRC/elastisomat.cpp:627); and
 * compute_ipgrads / compute_ipfluxes / internal_fluxes
 * (TRFEL/SRC/elemswitcht.cpp:1260,1111,1845) with linhext::intpointgrad /
 * internal_fluxes (TRFEL/SRC/linhext.cpp:432,1086) for heat. The element set
 * (mesh, code numbers, material) stays on the device across calls. */
typedef struct ssf_elements ssf_elements;
/* arguments as ssf_matrix_assemble; n = number of unknowns (code numbers <= n) */
int ssf_elements_create(ssf_ctx *ctx, int kind, int64_t n, int64_t nn, const double *xyz,
                        int64_t ne, const int64_t *conn, const int64_t *cn, const double *mat_a,
                        const double *mat_b, const double *thick, ssf_elements **out);
int ssf_elements_destroy(ssf_elements *el);
/* integration points per element (8 hex8, 4 quad4), components per point
 * (hex8 elasticity 6: eps x, y, z, gamma yz, xz, xy; hex8 heat 3: x, y, z;
 * quad4 plane stress 4: eps x, y, gamma xy, eps z with stress x, y, xy, 0) and
 * DOFs per element */
int ssf_elements_sizes(const ssf_elements *el, int32_t *ngp, int32_t *ncomp, int32_t *ndofe);
/* For the solution lhs[n] and prescribed values pv[npv] (eldispl,
 * MEFEL/SRC/globmat.cpp:1207: cn > 0 -> lhs[cn-1], cn < 0 -> pv[-cn-1], 0 -> 0):
 * strain[ne][ngp][ncomp] (heat: temperature gradients), stress[ne][ngp][ncomp]
 * (heat: fluxes q = -k grad T) and ifor[n], the element forces int B^T sigma dV
 * (heat: internal fluxes) added per unknown in element order (locglob,
 * GEFEL/matrix.cpp:3500). Any output may be NULL. Host arrays, synchronous. */
int ssf_elements_results(ssf_elements *el, const double *lhs, const double *pv, int64_t npv,
                         double *strain, double *stress, double *ifor);
/* same on device pointers, asynchronous on the context's stream */
int ssf_elements_results_dev(ssf_elements *el, const double *lhs_dev, const double *pv_dev,
                             int64_t npv, double *strain_dev, double *stress_dev,
                             double *ifor_dev);

/* Element volume loads of the set added into rhs[n] (on top of its contents,
 * e.g. the nodal loads), per unknown in element order as loadcase::forcecontrib
 * (MEFEL/SRC/loadcase.cpp:571-597) localizes them. Hex8 elasticity: the nodal
 * force vector of a volume load, loadel::volumeload (MEFEL/SRC/loadel.cpp:438)
 * with linhex::load_matrix (MEFEL/SRC/linhex.cpp:969); hex8 heat: the source
 * vector of linhext::quantity_source_vector (TRFEL/SRC/linhext.cpp:1039).
 * nodval[ne][ndofe] are the load's nodal values on each element. Host arrays,
 * synchronous (SSF_EUNSUP for quad4). */
int ssf_elements_volume_load(ssf_elements *el, const double *nodval, double *rhs);
/* same on device pointers, asynchronous on the context's stream */
int ssf_elements_volume_load_dev(ssf_elements *el, const double *nodval_dev, double *rhs_dev);
/* Surface loads of hex8 elasticity element records (loadel::surfaceload ->
 * linhex::node_forces_surf, MEFEL/SRC/loadel.cpp:549, MEFEL/SRC/linhex.cpp:
 * 2543-3047): record r loads element eid[r] (0-based) on the faces f with
 * ls[6r + f] > 0 (1: nodal values in global components, 2: in the face's local
 * axes of linhex::locglob_nodeval), nodvals[72r + 12f + 3m + c] = component c at
 * the face's m-th node. Faces in the reference's order: xi = 1, eta = 1, xi = -1,
 * eta = -1, zeta = 1, zeta = -1. The nodal forces are added to rhs by locglob
 * record after record (loadcase::assemble, MEFEL/SRC/loadcase.cpp:572-597), bit
 * for bit the reference's arithmetic. Host arrays, rhs in / out, synchronous; the
 * _dev form takes device arrays (int32 ids and flags). */
int ssf_elements_surface_load(ssf_elements *el, int64_t nrec, const int64_t *eid,
                              const int64_t *ls, const double *nodvals, double *rhs);
int ssf_elements_surface_load_dev(ssf_elements *el, int64_t nrec, const int32_t *eid_dev,
                                  const int32_t *ls_dev, const double *nodvals_dev,
                                  double *rhs_dev);

/* ----------------------------------------------------------- Newton-Raphson */
/* nonlinman fields of the Newton-Raphson load control (MEFEL/SRC/nonlinman.h) */
typedef struct ssf_nr_params {
  int64_t ninr;          /* maximum number of increments                         */
  int64_t niilnr;        /* maximum number of inner iterations per increment     */
  int64_t li;            /* index of the first increment (restart), 0 normally   */
  double errnr;          /* required norm of the unbalanced forces               */
  double incrnr;         /* initial load increment dlambda                       */
  double minincrnr;      /* minimum dlambda                                      */
  double maxincrnr;      /* maximum dlambda                                      */
  double ilambda;        /* initial load coefficient                             */
  double lambdar;        /* required load coefficient when check_lambdar = 1     */
  int32_t rnormtnr;      /* resnormt: 1 rel_load_norm, 4 absol_norm               */
  int32_t check_lambdar; /* flagsw: 1 stop at lambdar                            */
  int32_t check_div;     /* must be 0: the divergence check is not on this path  */
  int32_t pad;
} ssf_nr_params;

typedef struct ssf_nr_info {
  double lambda;         /* attained load coefficient (gnewton_raphson's result) */
  double norf;           /* last norm of the unbalanced forces                   */
  int64_t increments;    /* increments run                                       */
  int64_t solves;        /* linear solves                                        */
  int64_t rejected;      /* increments rejected (dlambda halved)                 */
  int64_t inner;         /* inner iterations                                     */
  int32_t status;        /* 1 ended (ninr or lambdar), 2 stopped: no equilibrium
                            at the minimum increment                            */
  int32_t pad;
} ssf_nr_info;

/* gnewton_raphson (MEFEL/SRC/newtonraph.cpp:51-268) for linear elastic element
 * sets (hex8 elasticity, quad4 plane stress) on device vectors. Per increment:
 * fa += dlambda fp; fi = the internal forces of ra (ssf_elements, as
 * loc_internal_forces); fb = fa - fi; K dr = fb by ssf_solve_dev with sp;
 * ra += dr; the residual norm (1: |fb| / |fa| when |fa| > 1e-9, 4: |fb|);
 * accepted below errnr, else up to niilnr inner corrections; a rejected
 * increment restores fa and ra and halves dlambda (not below minincrnr; at
 * minincrnr the loop stops); dlambda doubles after every second increment
 * accepted without inner iterations, capped at maxincrnr; with check_lambdar
 * the last increment is cut to reach lambdar and the loop ends there. With
 * ilambda = 0, fa = fc + 0 fp first (assemble_attained_load_vector).
 * K is the stiffness matrix of el's mesh (the linear elastic tangent at every
 * state); pv_dev[npv] are the prescribed values as in ssf_elements_results_dev.
 * fa_dev and ra_dev are updated. Histories (may be NULL, ninr - li entries):
 * lambda after each increment, its dlambda, its linear solves, accepted (1) or
 * rejected (0). Synchronous. */
int ssf_newton_raphson(ssf_matrix *K, ssf_elements *el, const ssf_nr_params *nr,
                       const ssf_solve_params *sp, const double *pv_dev, int64_t npv,
                       const double *fc_dev, const double *fp_dev, double *fa_dev,
                       double *ra_dev, double *lam_hist, double *dlam_hist, int64_t *nsol_hist,
                       int32_t *acc_hist, ssf_nr_info *info);

/* ------------------------------------------------------------ diagnostics */
/* Time `reps` calls of ssf_elements_results_dev: usec[0] = mean duration of the
 * element kernel (its own start and end events), usec[1] = mean duration of the
 * whole call (element kernel and the per-unknown force sums). */
int ssf_elements_bench(ssf_elements *el, const double *lhs_dev, const double *pv_dev,
                       int64_t npv, double *strain_dev, double *stress_dev, double *ifor_dev,
                       int reps, double *usec);
/* Time `reps` launches of the FP64 (precision 0) or FP32 (1) SpMV on device
 * vectors; returns the mean kernel time in microseconds measured with HIP
 * events on the context's stream. */
int ssf_spmv_bench(ssf_matrix *m, int precision, int reps, double *usec);
/* SpMV autotuner: for every tiling of the FP64 (precision 0) or FP32 (1) SpMV on
 * this matrix, time one iteration of the recurrence a solve runs with it (the
 * product and the element-wise kernels), and keep the fastest for later solves.
 * tune_out (may be NULL)
 * receives 5 values {format (0 row-vector, 1 stream tiles, 2 SELL-64 slices),
 * lanes per row (formats 0, 1) or value loads in flight per lane (format 2 with
 * row patterns: 2, 4 or 8; stencil form: 3 or 4 takes that many groups of 4
 * slots per round, other values all of a row's groups (7 per round for blocks
 * of 3 rows, 5 for blocks of 2)), blocks (at most 8192), idx16 (column form: 0 int32,
 * 1 16-bit tile or slice offsets, 2 SELL row patterns, 3 SELL stencil form: one
 * slot per column offset in every row, offered when the matrix has at most 32
 * offsets from the row itself, or at most 96 from the first row of the row's
 * block of 2 or 3 rows (the unknowns of a node numbered together: 81 for hex8
 * elasticity), costs at most 5 % more entries than the compact layout, and its
 * slices fit 32-bit byte offsets; a slot the row lacks holds 0
 * and multiplies x at the row plus that offset, so its products equal the other
 * forms' for finite x only: an Inf or NaN in x makes NaN in rows that do not
 * couple to it; 4 the symmetric half of the stencil form, for symm_comp_rows
 * matrices whose stencil is symmetric: the lower slots of a row (or of a node's
 * block of 2 or 3 rows) and its diagonal stored once, the upper entries read from
 * the mirrored rows, lanes 1 or 2 = blocks of 256 or 512 threads; its products
 * equal the other forms' bit for bit), map (SELL slices to wavefronts: 0 a contiguous range per block, 1
 * each of the 8 XCDs a contiguous eighth whose slices its wavefronts take in
 * turn; the default)}. */
int ssf_matrix_autotune(ssf_matrix *m, int precision, int reps, int32_t *tune_out);
/* pin the SpMV tiling {format, lanes, blocks, idx16, map} (5 values); the result
 * of a solve depends only on the matrix, the right-hand side and the two tilings */
int ssf_matrix_set_tune(ssf_matrix *m, int precision, const int32_t *tune);
/* Time `reps` launches of a plain device copy of `bytes` bytes (read + write
 * 2*bytes of HBM per launch): the achievable-bandwidth reference of the
 * roofline and the size-known kernel of the traffic-counter calibration. */
int ssf_copy_bench(ssf_ctx *ctx, int64_t bytes, int reps, double *usec);
/* bytes the SpMV of the given precision streams by the canonical count
 * nnz*(s_v+4) + 4(n+1) + 2 s_v n, and the bytes of the format actually used */
int ssf_spmv_bytes(const ssf_matrix *m, int precision, double *canonical, double *actual);
/* incomplete LDL^T of a symcomprow matrix (forms its plan and factor if needed):
 * the number of dependency levels of the forward and backward sweeps, each one
 * launch per application, and usec = mean time of `reps` applications
 * (prec_ildl_scr) on device vectors, HIP events on the context's stream;
 * graph = 1 replays the captured sweeps, 0 launches them on the stream */
int ssf_ildl_bench(ssf_matrix *m, int reps, int graph, int64_t *nfwd, int64_t *nbwd,
                   double *usec);

#ifdef __cplusplus
}
#endif
#endif
