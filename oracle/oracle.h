/* CPU oracle for the matrix-free layer-potential apply -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or as the timed CPU baseline.
 * The product (easybem_amd/) never links, loads or calls it.
 *
 * It restates, in FP64 plain C, the reference easybem CPU operator:
 *   Laplace single/double layer integrators  src/BEM3dLaplace.cpp:22-437
 *   pair classification and local rotations  src/BEM3dLaplace.cpp:692-765
 *   dense assembly driver getLaplace3d       src/BEM3dLaplace.cpp:561-781
 *   mass matrix getLaplace3dMass             src/BEM3dLaplace.cpp:448-486
 *   Kelvin single-layer integrators          src/BEM3dElasticity.cpp:30-348
 *   Kelvin block assembly                    src/BEM3dElasticity.cpp:612,838-846
 * and adds the operators the reference lacks (P1 test/trial single layer,
 * P1-test double layer, P1 Kelvin), built from Sauter-Schwab maps by exact
 * inner integration.  Which checks pin which function is listed in the
 * header of oracle/oracle.c.
 */
#ifndef EASYBEM_ORACLE_H
#define EASYBEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Quadrature rules indexed like the reference tables (src/quadrature.cpp). */
int orc_line_rule(int idx, double *pts, double *wts);
int orc_tri_rule(int idx, double *pts /* 2n */, double *wts);
int orc_gauss01(int n, double *pts, double *wts);

/* Dense P0 Laplace assembly, restating getLaplace3d with separate near
 * (line_idx) and far (tri_idx) rule indices; line_idx == tri_idx == order
 * reproduces the reference.  V row-major [test*ne + trial]; K, M row-major
 * [test*np + node].  Any output may be NULL. */
int orc_laplace_dense(int np, const double *nodes, int ne, const int *elems,
                      int line_idx, int tri_idx, double *V, double *K,
                      double *M);

/* Dense Kelvin single layer, P0, 3 components per element, restating
 * getElasticity3d's V; row-major [(a*ne+i)*3ne + b*ne+j]. */
int orc_kelvin_dense(int np, const double *nodes, int ne, const int *elems,
                     double nu, int line_idx, int tri_idx, double *V);

/* Row-sampled matrix-free applies.  Operators:
 *   op 0: Laplace single layer V    (test_basis, trial_basis in {0,1})
 *   op 1: Laplace double layer K    (trial basis must be 1)
 *   op 2: Kelvin single layer       (test == trial basis; x, y component-blocked)
 * Rows index test DOFs (triangles for P0, nodes for P1); for op 2 a row r
 * yields the 3 components y[a*nrows_total + r] written to y[a*nrows + k].
 * Uses OpenMP over rows; nthreads <= 0 means the OpenMP default. */
int orc_apply(int op, int test_basis, int trial_basis, int np,
              const double *nodes, int ne, const int *elems, int line_idx,
              int tri_idx, double nu, const double *x, int nrows,
              const int *rows, double *y, int nthreads);

/* Surface curls of the P1 hat functions restating getLaplace3dCurl
 * (src/BEM3dLaplace.cpp:489-558): curl[(3i + c)*3 + k] is component c of
 * the curl of triangle i's local hat function k.  elems are 1-based. */
int orc_surface_curls(int np, const double *nodes, int ne, const int *elems,
                      double *curl);

/* Representation formula at off-surface points restating evaluateLaplace3d
 * (src/BEM3dLaplace.cpp:784-843): t P0 (ne), u P1 (np), result npts. */
int orc_evaluate(int np, const double *nodes, int ne, const int *elems,
                 int npts, const double *points, const double *t,
                 const double *u, int tri_idx, double *result, int nthreads);

/* One pair's local block in the pair's local vertex orders, for tests:
 * out[9] for op 0/1 (k test, l trial, out[3k+l]); out[63] for op 2
 * (out[7*(3k+l) + c], c = I, I11, I22, I33, I12, I13, I23).
 * order_i/order_j return the global 0-based node numbers of the local
 * vertices.  use_tables=0 uses the reference reduced formulas (P0 block
 * sum in out[0] only); 1 uses the Sauter-Schwab tables. */
int orc_pair_block(int op, int np, const double *nodes, int ne,
                   const int *elems, int i, int j, int line_idx, int tri_idx,
                   int use_tables, double *out, int *order_i, int *order_j);

/* Full 4-D Sauter-Schwab integration of a test integrand over the pair's
 * reference triangles, for pinning the maps:
 *   kind 0: polynomial p(s,t) = s1^a s2^b t1^c t2^d (exps[4])
 *   kind 1: phi_k(s) phi_l(t) / |F_i(s)-F_j(t)| with geometry geom[12] =
 *           (u,v,w,z) as in the case's parameterisation; exps[0]=k, [1]=l
 * cls: 0 coincident, 1 edge, 2 vertex.  n: Gauss points per dimension. */
double orc_ss_full(int cls, int kind, const int *exps, const double *geom,
                   int n);

/* Weight tables of the reduced P1 Sauter-Schwab rules (see oracle.c). */
int orc_ss_table(int cls, int kernel_deg, int nq, double *W);

#ifdef __cplusplus
}
#endif
#endif
