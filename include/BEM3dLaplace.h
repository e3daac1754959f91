/* Drop-in C++ surface of the reference's dense Laplace assembly, same names
 * and signatures as the reference's include/BEM3dLaplace.h:8-26, computed on
 * the device through the C-ABI in easybem_hip.h
 * (easybem_amd/csrc/legacy_dense.cpp).
 *
 * Conventions as the reference's src/BEM3dLaplace.cpp:
 *   - elements are 1-based; the mesh must be a closed, consistently oriented
 *     surface (outward normal (P2-P1)x(P3-P1));
 *   - outputs are allocated with new[] by the callee and owned (delete[]) by
 *     the caller;
 *   - V is ne x ne with V[i*ne + j] the entry of test triangle i and trial
 *     triangle j (:602-726); K (P0 test x P1 trial) and M are ne x np with
 *     K[i + node*ne] (:730, :616); D is np x np with D[m + n*np] (:652);
 *     the curls C1..C3 are ne x np with C[node*ne + i] (:541-543);
 *   - `order` indexes both quadrature tables (:581-586): the Gauss rule of
 *     the near field and the triangle rule of the far field, 0..18
 *     (triangle rules stop at index 19).
 * Differences: nothing is printed; invalid input (open mesh, bad order)
 * prints the library's message to stderr and aborts. */
#ifndef BEM3DLAPLACE_H
#define BEM3DLAPLACE_H

void getLaplace3d(int nNodes, const double *nodes, int nElements, const int *elements,
                  double *&V, double *&K, double *&D, double *&M, int order);

void getLaplace3dMass(int nNodes, const double *nodes, int nElements, const int *elements,
                      double *&M);

void getLaplace3dCurl(int nNodes, const double *nodes, int nElements, const int *elements,
                      double *&C1, double *&C2, double *&C3);

/* result: nPoints values of the representation formula (:784-843) with the
 * P0 Neumann data `neumann` (nElements) and the P1 Dirichlet data
 * `dirichlet` (one value per node referenced by elements). */
void evaluateLaplace3d(const double *nodes, int nElements, const int *elements,
                       int nPoints, const double *points, const double *neumann,
                       const double *dirichlet, double *&result, int order);

#endif
