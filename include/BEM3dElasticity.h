/* Drop-in C++ surface of the reference's dense elasticity assembly, same
 * name and signature as the reference's include/BEM3dElasticity.h:7-11,
 * computed on the device through the C-ABI in easybem_hip.h
 * (easybem_amd/csrc/legacy_dense.cpp).
 *
 * E = 1, Poisson ratio nu.  Outputs are allocated with new[] by the callee
 * and owned by the caller, component-blocked and column-major as the
 * reference writes them (src/BEM3dElasticity.cpp:604-952): V (3ne x 3ne),
 * K (3ne x 3np), D (3np x 3np), M (3ne x 3np), entry (row r, column c) at
 * [r + c*rows], row a*ne + i for component a of triangle i (a*np + n for
 * node n).  Nothing is printed; invalid input aborts with the library's
 * message. */
#ifndef BEM3DELASTICITY_H
#define BEM3DELASTICITY_H

void getElasticity3d(int nNodes, const double *nodes, int nElements, const int *elements,
                     double nu, double *&V, double *&K, double *&D, double *&M, int order);

#endif
