// pybind11 module `easybem`: the reference module's GPU entry points
// (python/bindings.cpp:186-261) and its mass matrix (:42-60), with the same
// names, arguments and return shapes, over the HIP library.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cmath>
#include <stdexcept>
#include <vector>

#include "BEM3dLaplaceGPU.h"
#include "easybem_hip.h"

namespace py = pybind11;

namespace {

using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using iarr = py::array_t<int, py::array::c_style | py::array::forcecast>;

void check_mesh(const darr &nodes, const iarr &elements)
{
  if (nodes.ndim() != 2 || nodes.shape(1) != 3)
    throw std::invalid_argument("nodes must have shape (nNodes, 3)");
  if (elements.ndim() != 2 || elements.shape(1) != 3)
    throw std::invalid_argument("elements must have shape (nElements, 3)");
}

void need_live()
{
  if (gpuNumElements() == 0) throw std::runtime_error("call init_gpu first");
}

py::array_t<double> apply(const darr &x, size_t n_in, size_t n_out,
                          void (*fn)(const double *, double *, int), int ansatz)
{
  if ((size_t)x.size() != n_in)
    throw std::invalid_argument("vector has " + std::to_string(x.size()) +
                                " entries, expected " + std::to_string(n_in));
  py::array_t<double> out(n_out);
  const double *xp = x.data();
  double *yp = out.mutable_data();
  {
    py::gil_scoped_release nogil;
    fn(xp, yp, ansatz);
  }
  return out;
}

// A device mesh for the calls that take nodes and elements directly.
struct TempMesh {
  ebh_mesh *m = nullptr;
  TempMesh(const darr &nodes, const iarr &elements)
  {
    check_mesh(nodes, elements);
    if (ebh_mesh_create((int)nodes.shape(0), nodes.data(), (int)elements.shape(0),
                        elements.data(), &m) != EBH_OK)
      throw std::runtime_error(ebh_last_error());
  }
  ~TempMesh() { ebh_mesh_destroy(m); }
};

void v_apply(const double *x, double *y, int ansatz) { applyGPULaplace3d1Layer(x, y, ansatz); }
void k_apply(const double *x, double *y, int) { applyGPULaplace3d2Layer(x, y); }
void d_apply(const double *x, double *y, int) { applyGPULaplace3dHyperSing(x, y); }

}  // namespace

PYBIND11_MODULE(easybem, m)
{
  m.doc() = "easybem drop-in module over the MI355X HIP layer-potential apply";

  // P0 x P1 mass matrix, (nElements, nNodes), like getLaplace3dMass
  // (src/BEM3dLaplace.cpp:448-486): each triangle adds J/6 to its nodes.
  m.def("get_laplace3d_mass", [](darr nodes, iarr elements) {
    check_mesh(nodes, elements);
    const ssize_t np = nodes.shape(0), ne = elements.shape(0);
    py::array_t<double> M({ne, np});
    auto o = M.mutable_unchecked<2>();
    for (ssize_t i = 0; i < ne; i++)
      for (ssize_t j = 0; j < np; j++) o(i, j) = 0.0;
    auto P = nodes.unchecked<2>();
    auto T = elements.unchecked<2>();
    for (ssize_t i = 0; i < ne; i++) {
      int a = T(i, 0) - 1, b = T(i, 1) - 1, c = T(i, 2) - 1;
      if (a < 0 || b < 0 || c < 0 || a >= np || b >= np || c >= np)
        throw std::invalid_argument("element index out of range");
      double u[3], v[3];
      for (int k = 0; k < 3; k++) { u[k] = P(b, k) - P(a, k); v[k] = P(c, k) - P(a, k); }
      double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
             n2 = u[0] * v[1] - u[1] * v[0];
      double J = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      o(i, a) += J / 6.0; o(i, b) += J / 6.0; o(i, c) += J / 6.0;
    }
    return M;
  });

  // (C1, C2, C3), each (nElements, nNodes), like getLaplace3dCurl
  // (src/BEM3dLaplace.cpp:489-558; python/bindings.cpp:62-86).
  m.def("get_laplace3d_curl", [](darr nodes, iarr elements) {
    check_mesh(nodes, elements);
    const ssize_t np = nodes.shape(0), ne = elements.shape(0);
    std::vector<double> cu((size_t)9 * ne);
    if (ebh_surface_curls((int)np, nodes.data(), (int)ne, elements.data(), cu.data()) !=
        EBH_OK)
      throw std::invalid_argument(ebh_last_error());
    auto T = elements.unchecked<2>();
    py::array_t<double> C[3] = {py::array_t<double>({ne, np}), py::array_t<double>({ne, np}),
                                py::array_t<double>({ne, np})};
    for (int c = 0; c < 3; c++) {
      auto o = C[c].mutable_unchecked<2>();
      for (ssize_t i = 0; i < ne; i++)
        for (ssize_t j = 0; j < np; j++) o(i, j) = 0.0;
      for (ssize_t i = 0; i < ne; i++)
        for (int k = 0; k < 3; k++) o(i, T(i, k) - 1) = cu[(3 * (size_t)i + c) * 3 + k];
    }
    return py::make_tuple(C[0], C[1], C[2]);
  });

  // (V, K, D, M) like getLaplace3d (src/BEM3dLaplace.cpp:561-781;
  // python/bindings.cpp:88-116), assembled on the device.  V is returned as
  // [test, trial]; the reference module returns its transpose
  // (python/bindings.cpp:28-38 reads the row-major V column-major).
  m.def("get_laplace3d", [](darr nodes, iarr elements, int order) {
    TempMesh tm(nodes, elements);
    const ssize_t np = nodes.shape(0), ne = elements.shape(0);
    py::array_t<double> V({ne, ne}), K({ne, np}), D({np, np}), M({ne, np});
    double *pv = V.mutable_data(), *pk = K.mutable_data(), *pd = D.mutable_data(),
           *pm = M.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = ebh_assemble_laplace(tm.m, order, pv, pk, pd, pm);
    }
    if (rc != EBH_OK) throw std::runtime_error(ebh_last_error());
    return py::make_tuple(V, K, D, M);
  });

  // (V, K, D, M) like getElasticity3d (src/BEM3dElasticity.cpp:604-952;
  // python/bindings.cpp:151-182), E = 1, component-blocked.
  m.def("get_elasticity3d", [](darr nodes, iarr elements, double nu, int order) {
    TempMesh tm(nodes, elements);
    const ssize_t np = nodes.shape(0), ne = elements.shape(0);
    py::array_t<double> V({3 * ne, 3 * ne}), K({3 * ne, 3 * np}), D({3 * np, 3 * np}),
        M({3 * ne, 3 * np});
    double *pv = V.mutable_data(), *pk = K.mutable_data(), *pd = D.mutable_data(),
           *pm = M.mutable_data();
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = ebh_assemble_elasticity(tm.m, nu, order, pv, pk, pd, pm);
    }
    if (rc != EBH_OK) throw std::runtime_error(ebh_last_error());
    return py::make_tuple(V, K, D, M);
  });

  // Representation formula at points, evaluateLaplace3d
  // (src/BEM3dLaplace.cpp:784-843; python/bindings.cpp:118-149): returns a
  // 1-tuple holding an (nPoints, 1) array.
  m.def("eval_laplace3d", [](darr nodes, iarr elements, darr points, darr neumann,
                             darr dirichlet, int order) {
    TempMesh tm(nodes, elements);
    if (points.ndim() != 2 || points.shape(1) != 3)
      throw std::invalid_argument("points must have shape (nPoints, 3)");
    if (neumann.size() != elements.shape(0) || dirichlet.size() != nodes.shape(0))
      throw std::invalid_argument("neumann needs nElements, dirichlet nNodes entries");
    const ssize_t npts = points.shape(0);
    py::array_t<double> out({npts, (ssize_t)1});
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = ebh_evaluate_laplace_host(tm.m, order, EBH_FP64, (int)npts, points.data(),
                                     neumann.data(), dirichlet.data(),
                                     out.mutable_data());
    }
    if (rc != EBH_OK) throw std::runtime_error(ebh_last_error());
    return py::make_tuple(out);
  });

  m.def("init_gpu", [](darr nodes, iarr elements, int order, int nblocks, int nthreads) {
    check_mesh(nodes, elements);
    const double *P = nodes.data();
    const int *T = elements.data();
    const int np = (int)nodes.shape(0), ne = (int)elements.shape(0);
    py::gil_scoped_release nogil;
    initGPU(np, P, ne, T, order, nblocks, nthreads);
  }, py::arg("nodes"), py::arg("elements"), py::arg("order"),
     py::arg("nblocks") = 0, py::arg("nthreads") = 0);

  m.def("close_gpu", []() { closeGPU(); });

  m.def("set_far_precision", [](int fp32) { setGPUFarPrecision(fp32); },
        "0: FP64 far field (default); 1: FP32 far field");

  m.def("apply_gpu_laplace3d_1layer", [](darr x) {
    need_live();
    return apply(x, gpuNumElements(), gpuNumElements(), v_apply, 0);
  });
  m.def("apply_gpu_laplace3d_1layer_linears", [](darr x) {
    need_live();
    return apply(x, gpuNumNodes(), gpuNumNodes(), v_apply, 1);
  });
  m.def("apply_gpu_laplace3d_2layer", [](darr u) {
    need_live();
    return apply(u, gpuNumNodes(), gpuNumElements(), k_apply, 0);
  });
  m.def("apply_gpu_laplace3d_hypersing", [](darr u) {
    need_live();
    return apply(u, gpuNumNodes(), gpuNumNodes(), d_apply, 0);
  });

  m.def("version", []() { return "1.0"; });
}
