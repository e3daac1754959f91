// The reference's C++ GPU surface (include/BEM3dLaplaceGPU.h) over the C-ABI.
#include "BEM3dLaplaceGPU.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "easybem_hip.h"

namespace {

struct LegacyState {
  ebh_mesh *mesh = nullptr;
  ebh_op *ops[4] = {nullptr, nullptr, nullptr, nullptr};  // V P0, V P1, K, V P0 (D)
  int order = 4;
  int fp32 = 0;
  int np = 0, ne = 0;
  std::vector<int> tri;           // 0-based
  std::vector<double> curl;       // ne x 3 comps x 3 local nodes
} g;

[[noreturn]] void fail(const char *where)
{
  std::fprintf(stderr, "easybem-hip: %s: %s\n", where, ebh_last_error());
  std::abort();
}

void need_mesh(const char *where)
{
  if (!g.mesh) {
    std::fprintf(stderr, "easybem-hip: %s called before initGPU\n", where);
    std::abort();
  }
}

ebh_op *get_op(int which)
{
  if (g.ops[which]) return g.ops[which];
  ebh_op_desc d{};
  d.kernel = which == 2 ? EBH_LAPLACE_DL : EBH_LAPLACE_SL;
  d.test_basis = which == 1 ? EBH_P1 : EBH_P0;
  d.trial_basis = which == 1 || which == 2 ? EBH_P1 : EBH_P0;
  d.far_precision = g.fp32 ? EBH_FP32 : EBH_FP64;
  d.line_rule = g.order;
  d.tri_rule = g.order;
  d.nu = 0.0;
  d.max_rhs = which == 3 ? 3 : 1;
  if (ebh_op_create(g.mesh, &d, &g.ops[which]) != EBH_OK) fail("operator setup");
  return g.ops[which];
}

void drop_ops()
{
  for (auto &o : g.ops) {
    ebh_op_destroy(o);
    o = nullptr;
  }
}

// Surface curls of the P1 hat functions on each triangle, restating
// getLaplace3dCurl (src/BEM3dLaplace.cpp:489-558): curl[(3i + c)*3 + k].
void build_curls(const double *nodes)
{
  g.curl.assign((size_t)9 * g.ne, 0.0);
  for (int i = 0; i < g.ne; i++) {
    const double *P[3];
    for (int k = 0; k < 3; k++) P[k] = nodes + 3 * g.tri[3 * i + k];
    double e1[3], e2[3];
    for (int k = 0; k < 3; k++) { e1[k] = P[1][k] - P[0][k]; e2[k] = P[2][k] - P[0][k]; }
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    const double J = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (double &c : n) c /= J;
    const double a11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double a22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    const double a12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    const double det = a11 * a22 - a12 * a12;
    // b_c = (e x n)_c for the two edges; y solves the 2x2 metric system
    const double b1[3] = {-e1[1] * n[2] + e1[2] * n[1], e1[0] * n[2] - e1[2] * n[0],
                          -e1[0] * n[1] + e1[1] * n[0]};
    const double b2[3] = {-e2[1] * n[2] + e2[2] * n[1], e2[0] * n[2] - e2[2] * n[0],
                          -e2[0] * n[1] + e2[1] * n[0]};
    for (int c = 0; c < 3; c++) {
      const double y1 = (b1[c] * a22 - b2[c] * a12) / det;
      const double y2 = (a11 * b2[c] - a12 * b1[c]) / det;
      double *o = &g.curl[(3 * (size_t)i + c) * 3];
      o[0] = -y1 - y2;
      o[1] = y1;
      o[2] = y2;
    }
  }
}

}  // namespace

void initGPU(int nNodes, const double *nodes, int nElements, const int *elements,
             int order, int nblocks, int nthreads)
{
  (void)nblocks;
  (void)nthreads;
  closeGPU();
  if (ebh_mesh_create(nNodes, nodes, nElements, elements, &g.mesh) != EBH_OK)
    fail("initGPU");
  g.order = order;
  g.np = nNodes;
  g.ne = nElements;
  g.tri.resize((size_t)3 * nElements);
  for (size_t k = 0; k < g.tri.size(); k++) g.tri[k] = elements[k] - 1;
  build_curls(nodes);
  get_op(0);  // validates the order now, as initGPU uploads the rules
}

void closeGPU()
{
  drop_ops();
  ebh_mesh_destroy(g.mesh);
  g.mesh = nullptr;
}

void setGPUFarPrecision(int fp32)
{
  if ((fp32 != 0) != (g.fp32 != 0)) drop_ops();
  g.fp32 = fp32 != 0;
}

int gpuNumElements() { return g.mesh ? g.ne : 0; }
int gpuNumNodes() { return g.mesh ? g.np : 0; }

void applyGPULaplace3d1Layer(const double *t, double *v, int ansatz)
{
  need_mesh("applyGPULaplace3d1Layer");
  if (ebh_op_apply_host(get_op(ansatz ? 1 : 0), t, v, 1) != EBH_OK)
    fail("applyGPULaplace3d1Layer");
}

void applyGPULaplace3d2Layer(const double *u, double *q)
{
  need_mesh("applyGPULaplace3d2Layer");
  if (ebh_op_apply_host(get_op(2), u, q, 1) != EBH_OK) fail("applyGPULaplace3d2Layer");
}

void applyGPULaplace3dHyperSing(const double *u, double *q)
{
  need_mesh("applyGPULaplace3dHyperSing");
  const int ne = g.ne, np = g.np;
  std::vector<double> w((size_t)3 * ne), y((size_t)3 * ne);
  for (int i = 0; i < ne; i++)
    for (int c = 0; c < 3; c++) {
      const double *o = &g.curl[(3 * (size_t)i + c) * 3];
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += o[k] * u[g.tri[3 * i + k]];
      w[(size_t)c * ne + i] = s;
    }
  if (ebh_op_apply_host(get_op(3), w.data(), y.data(), 3) != EBH_OK)
    fail("applyGPULaplace3dHyperSing");
  for (int n = 0; n < np; n++) q[n] = 0.0;
  for (int i = 0; i < ne; i++)
    for (int c = 0; c < 3; c++) {
      const double *o = &g.curl[(3 * (size_t)i + c) * 3];
      const double yc = y[(size_t)c * ne + i];
      for (int k = 0; k < 3; k++) q[g.tri[3 * i + k]] += o[k] * yc;
    }
}
