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
  ebh_op *ops[4] = {nullptr, nullptr, nullptr, nullptr};  // V P0, V P1, K, D
  int order = 4;
  int fp32 = 0;
  int np = 0, ne = 0;
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
  d.kernel = which == 2 ? EBH_LAPLACE_DL : which == 3 ? EBH_LAPLACE_HS : EBH_LAPLACE_SL;
  d.test_basis = which == 1 || which == 3 ? EBH_P1 : EBH_P0;
  d.trial_basis = which >= 1 ? EBH_P1 : EBH_P0;
  d.far_precision = g.fp32 ? EBH_FP32 : EBH_FP64;
  d.line_rule = g.order;
  d.tri_rule = g.order;
  d.nu = 0.0;
  d.max_rhs = 1;
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

// D = sum_c C_c^T V C_c through the device hypersingular operator (the
// reference's stub, src/BEM3dLaplaceGPU.cu:1554-1555, is empty).
void applyGPULaplace3dHyperSing(const double *u, double *q)
{
  need_mesh("applyGPULaplace3dHyperSing");
  if (ebh_op_apply_host(get_op(3), u, q, 1) != EBH_OK) fail("applyGPULaplace3dHyperSing");
}
