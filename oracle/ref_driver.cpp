// Test-infrastructure only: a C-ABI driver over the reference easybem CPU
// operator, compiled together with the reference's own sources (see
// oracle/Makefile) into oracle/_ref/libeasybem_ref.so.  Nothing in the
// product links or loads this file; tests/ use it to pin the C oracle.
//
// Entry points wrap (reference file:line):
//   getLaplace3d      src/BEM3dLaplace.cpp:561-781   (include/BEM3dLaplace.h:8-11)
//   getLaplace3dMass  src/BEM3dLaplace.cpp:448-486
//   getElasticity3d   src/BEM3dElasticity.cpp:604-952
//   quadrature tables src/quadrature.cpp:286-287,1047-1048
#include <cstring>
#include "BEM3dLaplace.h"
#include "BEM3dElasticity.h"
#include "quadrature.h"

extern "C" {

// V (ne*ne), K (ne*np), D (np*np), M (ne*np) copied out in the reference's
// own storage order: V row-major [test*ne + trial]; K, D, M column-major.
int ebref_laplace3d(int np, const double *nodes, int ne, const int *elems,
                    int order, double *V, double *K, double *D, double *M)
{
  double *v = nullptr, *k = nullptr, *d = nullptr, *m = nullptr;
  getLaplace3d(np, nodes, ne, elems, v, k, d, m, order);
  std::memcpy(V, v, sizeof(double) * (size_t)ne * ne);
  std::memcpy(K, k, sizeof(double) * (size_t)ne * np);
  std::memcpy(D, d, sizeof(double) * (size_t)np * np);
  std::memcpy(M, m, sizeof(double) * (size_t)ne * np);
  delete[] v; delete[] k; delete[] d; delete[] m;
  return 0;
}

// C1, C2, C3 (ne x np, column-major as the reference writes them).
int ebref_laplace3d_curl(int np, const double *nodes, int ne, const int *elems,
                         double *C1, double *C2, double *C3)
{
  double *c1 = nullptr, *c2 = nullptr, *c3 = nullptr;
  getLaplace3dCurl(np, nodes, ne, elems, c1, c2, c3);
  std::memcpy(C1, c1, sizeof(double) * (size_t)ne * np);
  std::memcpy(C2, c2, sizeof(double) * (size_t)ne * np);
  std::memcpy(C3, c3, sizeof(double) * (size_t)ne * np);
  delete[] c1; delete[] c2; delete[] c3;
  return 0;
}

// evaluateLaplace3d at npts points (result npts).
int ebref_evaluate_laplace3d(int np, const double *nodes, int ne, const int *elems,
                             int npts, const double *points, const double *t,
                             const double *u, int order, double *result)
{
  (void)np;
  double *r = nullptr;
  evaluateLaplace3d(nodes, ne, elems, npts, points, t, u, r, order);
  std::memcpy(result, r, sizeof(double) * (size_t)npts);
  delete[] r;
  return 0;
}

int ebref_laplace3d_mass(int np, const double *nodes, int ne,
                         const int *elems, double *M)
{
  double *m = nullptr;
  getLaplace3dMass(np, nodes, ne, elems, m);
  std::memcpy(M, m, sizeof(double) * (size_t)ne * np);
  delete[] m;
  return 0;
}

// V (3ne x 3ne), K (3ne x 3np), D (3np x 3np), M (3ne x 3np), all
// column-major as the reference writes them; K, D, M may be NULL.
int ebref_elasticity3d(int np, const double *nodes, int ne, const int *elems,
                       double nu, int order, double *V, double *K, double *D,
                       double *M)
{
  double *v = nullptr, *k = nullptr, *d = nullptr, *m = nullptr;
  getElasticity3d(np, nodes, ne, elems, nu, v, k, d, m, order);
  std::memcpy(V, v, sizeof(double) * (size_t)9 * ne * ne);
  if (K) std::memcpy(K, k, sizeof(double) * (size_t)9 * ne * np);
  if (D) std::memcpy(D, d, sizeof(double) * (size_t)9 * np * np);
  if (M) std::memcpy(M, m, sizeof(double) * (size_t)9 * ne * np);
  delete[] v; delete[] k; delete[] d; delete[] m;
  return 0;
}

int ebref_line_rule(int idx, double *pts, double *wts)
{
  int n = LineQuadratureSizes[idx];
  for (int i = 0; i < n; i++) {
    pts[i] = LineQuadraturePoints[idx][i];
    wts[i] = LineQuadratureWeights[idx][i];
  }
  return n;
}

int ebref_tri_rule(int idx, double *pts, double *wts)
{
  int n = TriangleQuadratureSizes[idx];
  for (int i = 0; i < n; i++) {
    pts[2 * i] = TriangleQuadraturePoints[idx][i].x;
    pts[2 * i + 1] = TriangleQuadraturePoints[idx][i].y;
    wts[i] = TriangleQuadratureWeights[idx][i];
  }
  return n;
}

}  // extern "C"
