// The reference's dense C++ surface (include/BEM3dLaplace.h,
// include/BEM3dElasticity.h) over the C-ABI: each call creates a temporary
// device mesh, assembles on the device (ebh_assemble_laplace,
// ebh_assemble_elasticity, ebh_evaluate_laplace_host, ebh_surface_curls)
// and returns new[] arrays in the reference's layouts.
#include "BEM3dElasticity.h"
#include "BEM3dLaplace.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "easybem_hip.h"

namespace {

[[noreturn]] void fail(const char *where)
{
  std::fprintf(stderr, "easybem-hip: %s: %s\n", where, ebh_last_error());
  std::abort();
}

struct TempMesh {
  ebh_mesh *m = nullptr;
  TempMesh(int np, const double *nodes, int ne, const int *elems, const char *where)
  {
    if (ebh_mesh_create(np, nodes, ne, elems, &m) != EBH_OK) fail(where);
  }
  ~TempMesh() { ebh_mesh_destroy(m); }
};

// a new[] copy of the row-major rows x cols matrix A, column-major
double *column_major(const std::vector<double> &A, size_t rows, size_t cols)
{
  double *out = new double[rows * cols];
  for (size_t r = 0; r < rows; r++)
    for (size_t c = 0; c < cols; c++) out[r + c * rows] = A[r * cols + c];
  return out;
}

}  // namespace

void getLaplace3d(int nNodes, const double *nodes, int nElements, const int *elements,
                  double *&V, double *&K, double *&D, double *&M, int order)
{
  TempMesh tm(nNodes, nodes, nElements, elements, "getLaplace3d");
  const size_t ne = nElements, np = nNodes;
  std::vector<double> v(ne * ne), k(ne * np), d(np * np), m(ne * np);
  if (ebh_assemble_laplace(tm.m, order, v.data(), k.data(), d.data(), m.data()) != EBH_OK)
    fail("getLaplace3d");
  // V[i*ne + j] (test i, trial j) is the row-major matrix itself
  V = new double[ne * ne];
  std::copy(v.begin(), v.end(), V);
  K = column_major(k, ne, np);
  D = column_major(d, np, np);
  M = column_major(m, ne, np);
}

void getLaplace3dMass(int nNodes, const double *nodes, int nElements, const int *elements,
                      double *&M)
{
  TempMesh tm(nNodes, nodes, nElements, elements, "getLaplace3dMass");
  const size_t ne = nElements, np = nNodes;
  std::vector<double> m(ne * np);
  if (ebh_assemble_laplace(tm.m, 0, nullptr, nullptr, nullptr, m.data()) != EBH_OK)
    fail("getLaplace3dMass");
  M = column_major(m, ne, np);
}

void getLaplace3dCurl(int nNodes, const double *nodes, int nElements, const int *elements,
                      double *&C1, double *&C2, double *&C3)
{
  const size_t ne = nElements, np = nNodes;
  std::vector<double> cu(9 * ne);
  if (ebh_surface_curls(nNodes, nodes, nElements, elements, cu.data()) != EBH_OK)
    fail("getLaplace3dCurl");
  double *C[3];
  for (int c = 0; c < 3; c++) {
    C[c] = new double[ne * np];
    std::fill(C[c], C[c] + ne * np, 0.0);
    for (size_t i = 0; i < ne; i++)
      for (int k = 0; k < 3; k++)
        C[c][(size_t)(elements[3 * i + k] - 1) * ne + i] = cu[(3 * i + c) * 3 + k];
  }
  C1 = C[0];
  C2 = C[1];
  C3 = C[2];
}

void evaluateLaplace3d(const double *nodes, int nElements, const int *elements,
                       int nPoints, const double *points, const double *neumann,
                       const double *dirichlet, double *&result, int order)
{
  // the reference takes no node count: the nodes are those the elements use
  int np = 0;
  for (size_t k = 0; k < 3 * (size_t)nElements; k++) np = std::max(np, elements[k]);
  TempMesh tm(np, nodes, nElements, elements, "evaluateLaplace3d");
  result = new double[nPoints];
  if (ebh_evaluate_laplace_host(tm.m, order, EBH_FP64, nPoints, points, neumann, dirichlet,
                                result) != EBH_OK)
    fail("evaluateLaplace3d");
}

void getElasticity3d(int nNodes, const double *nodes, int nElements, const int *elements,
                     double nu, double *&V, double *&K, double *&D, double *&M, int order)
{
  TempMesh tm(nNodes, nodes, nElements, elements, "getElasticity3d");
  const size_t ne = nElements, np = nNodes;
  std::vector<double> v(9 * ne * ne), k(9 * ne * np), d(9 * np * np), m(9 * ne * np);
  if (ebh_assemble_elasticity(tm.m, nu, order, v.data(), k.data(), d.data(), m.data()) !=
      EBH_OK)
    fail("getElasticity3d");
  V = column_major(v, 3 * ne, 3 * ne);
  K = column_major(k, 3 * ne, 3 * np);
  D = column_major(d, 3 * np, 3 * np);
  M = column_major(m, 3 * ne, 3 * np);
}
