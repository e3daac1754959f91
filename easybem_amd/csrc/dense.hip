// Dense Galerkin matrices on the device, getLaplace3d
// (src/BEM3dLaplace.cpp:561-781): V (P0 x P0), K (P0 test x P1 trial),
// D = sum_c C_c^T V C_c (P1 x P1) and M (P0 x P1).
//
// Far entries come from the same FP64 quadrature data as the matrix-free
// operators (target points and weights, source points and weights,
// normals); touching pairs take the near-field CSR values of those
// operators, so every entry equals the corresponding apply's.  One thread
// per entry:
//   V[i][j]  sum_a sum_b W_a(i) W_b(j) / (4 pi r)            (j not touching i)
//   K[i][n]  sum over the triangles j at node n, not touching i, of
//            sum_a sum_b W_a(i) W_b(j) phi_l(b) (x_a - y_b).n_j / (4 pi r^3)
//   D        two gathers through the node -> triangle lists per component c:
//            W_c = V C_c (ne x np), then D += C_c^T W_c.
#include <algorithm>
#include <cmath>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr double kInv4PiD = 0.07957747154594766788;

__device__ __forceinline__ bool touching(const int *ptr, const int *list, int i, int j)
{
  int lo = ptr[i], hi = ptr[i + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int v = list[mid];
    if (v == j) return true;
    if (v < j) lo = mid + 1; else hi = mid;
  }
  return false;
}

struct DenseArgs {
  int ne, np, nt;
  const double *tq;      // [nt][3][ne] target points
  const double *tw;      // [nt][ne]    w_a J
  const double *sg;      // [ne*nt][3]  source points
  const double *swt;     // [ne*nt]     w_b J
  const double *normals; // [ne][3]
  const int *touch_ptr, *touch;
  const int *ntp, *ntri, *ntk;  // node -> (triangle, local index)
  double phi[kMaxTriPoints][3];  // trial shape functions at the rule points
};

__global__ __launch_bounds__(256) void k_dense_v(DenseArgs a, double *__restrict__ V)
{
  const int j = blockIdx.x * 16 + threadIdx.x % 16;
  const int i = blockIdx.y * 16 + threadIdx.x / 16;
  if (i >= a.ne || j >= a.ne || touching(a.touch_ptr, a.touch, i, j)) return;
  double s = 0.0;
  for (int q = 0; q < a.nt; q++) {
    const double x0 = a.tq[(3 * q) * (size_t)a.ne + i];
    const double x1 = a.tq[(3 * q + 1) * (size_t)a.ne + i];
    const double x2 = a.tq[(3 * q + 2) * (size_t)a.ne + i];
    double t = 0.0;
    for (int b = 0; b < a.nt; b++) {
      const double *y = a.sg + 3 * ((size_t)j * a.nt + b);
      const double d0 = x0 - y[0], d1 = x1 - y[1], d2 = x2 - y[2];
      t += a.swt[(size_t)j * a.nt + b] / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
    s += a.tw[(size_t)q * a.ne + i] * t;
  }
  V[(size_t)i * a.ne + j] = s * kInv4PiD;
}

__global__ __launch_bounds__(256) void k_dense_k(DenseArgs a, double *__restrict__ K)
{
  const int n = blockIdx.x * 16 + threadIdx.x % 16;
  const int i = blockIdx.y * 16 + threadIdx.x / 16;
  if (i >= a.ne || n >= a.np) return;
  double s = 0.0;
  for (int e = a.ntp[n]; e < a.ntp[n + 1]; e++) {
    const int j = a.ntri[e], l = a.ntk[e];
    if (touching(a.touch_ptr, a.touch, i, j)) continue;
    const double n0 = a.normals[3 * j], n1 = a.normals[3 * j + 1], n2 = a.normals[3 * j + 2];
    for (int q = 0; q < a.nt; q++) {
      const double x0 = a.tq[(3 * q) * (size_t)a.ne + i];
      const double x1 = a.tq[(3 * q + 1) * (size_t)a.ne + i];
      const double x2 = a.tq[(3 * q + 2) * (size_t)a.ne + i];
      double t = 0.0;
      for (int b = 0; b < a.nt; b++) {
        const double *y = a.sg + 3 * ((size_t)j * a.nt + b);
        const double d0 = x0 - y[0], d1 = x1 - y[1], d2 = x2 - y[2];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        const double ri = 1.0 / sqrt(r2);
        t += a.swt[(size_t)j * a.nt + b] * a.phi[b][l] * (d0 * n0 + d1 * n1 + d2 * n2) *
             ri * ri * ri;
      }
      s += a.tw[(size_t)q * a.ne + i] * t;
    }
  }
  K[(size_t)i * a.np + n] = s * kInv4PiD;
}

// adds a CSR (rows x cols) into a dense row-major matrix
__global__ void k_dense_add_csr(int rows, int cols, const int *__restrict__ ptr,
                                const int *__restrict__ col, const double *__restrict__ val,
                                double *__restrict__ A)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  for (int e = ptr[r]; e < ptr[r + 1]; e++) A[(size_t)r * cols + col[e]] += val[e];
}

// W[i][n] = sum_{(j,k) at n} V[i][j] curl[(3j + c) 3 + k]
__global__ __launch_bounds__(256) void k_dense_vc(DenseArgs a, int c,
                                                  const double *__restrict__ curl,
                                                  const double *__restrict__ V,
                                                  double *__restrict__ W)
{
  const int n = blockIdx.x * 16 + threadIdx.x % 16;
  const int i = blockIdx.y * 16 + threadIdx.x / 16;
  if (i >= a.ne || n >= a.np) return;
  double s = 0.0;
  for (int e = a.ntp[n]; e < a.ntp[n + 1]; e++)
    s += V[(size_t)i * a.ne + a.ntri[e]] * curl[(3 * (size_t)a.ntri[e] + c) * 3 + a.ntk[e]];
  W[(size_t)i * a.np + n] = s;
}

// D[m][n] += sum_{(i,k) at m} curl[(3i + c) 3 + k] W[i][n]
__global__ __launch_bounds__(256) void k_dense_ctw(DenseArgs a, int c,
                                                   const double *__restrict__ curl,
                                                   const double *__restrict__ W,
                                                   double *__restrict__ D)
{
  const int n = blockIdx.x * 16 + threadIdx.x % 16;
  const int m = blockIdx.y * 16 + threadIdx.x / 16;
  if (m >= a.np || n >= a.np) return;
  double s = 0.0;
  for (int e = a.ntp[m]; e < a.ntp[m + 1]; e++)
    s += curl[(3 * (size_t)a.ntri[e] + c) * 3 + a.ntk[e]] * W[(size_t)a.ntri[e] * a.np + n];
  D[(size_t)m * a.np + n] += s;
}

__global__ void k_dense_zero(size_t n, double *p)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_set_entry(double *x, size_t i, double v) { x[i] = v; }

// column c of a row-major rows x cols matrix
__global__ void k_put_column(int rows, int cols, int c, const double *__restrict__ y,
                             double *__restrict__ A)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) A[(size_t)r * cols + c] = y[r];
}

// Dense matrix of an operator by applying it to the unit vectors (one
// apply per column, all on the op's stream), copied to host row-major.
int dense_by_columns(ebh_op *op, double *A)
{
  int rows = 0, cols = 0;
  EBH_TRY(ebh_op_dims(op, &rows, &cols));
  DevBuf<double> x, y, dA;
  EBH_TRY(x.alloc(cols));
  EBH_TRY(y.alloc(rows));
  EBH_TRY(dA.alloc((size_t)rows * cols));
  const hipStream_t s = op->stream;
  hipLaunchKernelGGL(k_dense_zero, dim3((cols + 255) / 256), dim3(256), 0, s,
                     (size_t)cols, x.p);
  for (int c = 0; c < cols; c++) {
    hipLaunchKernelGGL(k_set_entry, dim3(1), dim3(1), 0, s, x.p, (size_t)c, 1.0);
    EBH_TRY(op_apply_one(op, x.p, y.p, s, false));
    hipLaunchKernelGGL(k_put_column, dim3((rows + 255) / 256), dim3(256), 0, s, rows, cols,
                       c, y.p, dA.p);
    hipLaunchKernelGGL(k_set_entry, dim3(1), dim3(1), 0, s, x.p, (size_t)c, 0.0);
  }
  EBH_HIP(hipGetLastError());
  EBH_HIP(hipStreamSynchronize(s));
  EBH_HIP(gated([&] { return hipMemcpy(A, dA.p, dA.bytes(), hipMemcpyDeviceToHost); }));
  return EBH_OK;
}

DenseArgs dense_args(const ebh_op *v)
{
  DenseArgs a;
  const Topology &T = v->mesh->T;
  a.ne = T.ne;
  a.np = T.np;
  a.nt = v->rule.nt;
  a.tq = v->tq64.p;
  a.tw = v->tw.p;
  a.sg = v->sg64.p;
  a.swt = v->swt.p;
  a.normals = v->normals.p;
  a.touch_ptr = v->mesh->touch_ptr.p;
  a.touch = v->mesh->touch.p;
  a.ntp = v->mesh->node_tri_ptr.p;
  a.ntri = v->mesh->node_tri.p;
  a.ntk = v->mesh->node_tri_k.p;
  for (int q = 0; q < kMaxTriPoints; q++) {
    const bool in = q < a.nt;
    a.phi[q][0] = in ? 1.0 - v->rule.x[q] - v->rule.y[q] : 0.0;
    a.phi[q][1] = in ? v->rule.x[q] : 0.0;
    a.phi[q][2] = in ? v->rule.y[q] : 0.0;
  }
  return a;
}

}  // namespace
}  // namespace ebh

using namespace ebh;

extern "C" {

int ebh_assemble_laplace(ebh_mesh *mesh, int order, double *V, double *K, double *D,
                         double *M)
{
  if (!mesh) { set_error("null mesh"); return EBH_ERR_ARG; }
  const Topology &T = mesh->T;
  const int ne = T.ne, np = T.np;
  EBH_HIP(hipSetDevice(mesh->device));
  if (M) {
    std::fill(M, M + (size_t)ne * np, 0.0);
    for (int i = 0; i < ne; i++) {
      const int *f = &T.tri[3 * i];
      const double *a = &T.nodes[3 * f[0]], *b = &T.nodes[3 * f[1]], *c = &T.nodes[3 * f[2]];
      const double p[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double q[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      const double n0 = p[1] * q[2] - p[2] * q[1], n1 = p[2] * q[0] - p[0] * q[2],
                   n2 = p[0] * q[1] - p[1] * q[0];
      const double J = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      for (int k = 0; k < 3; k++) M[(size_t)i * np + f[k]] += J / 6.0;
    }
  }
  if (!V && !K && !D) return EBH_OK;
  ebh_op_desc d{};
  d.kernel = EBH_LAPLACE_SL;
  d.test_basis = d.trial_basis = EBH_P0;
  d.far_precision = EBH_FP64;
  d.line_rule = d.tri_rule = order;
  d.max_rhs = 1;
  ebh_op *vop = nullptr, *kop = nullptr;
  EBH_TRY(ebh_op_create(mesh, &d, &vop));
  int rc = EBH_OK;
  DevBuf<double> dV, dK, dD, dW, dcurl;
  const DenseArgs a = dense_args(vop);
  const hipStream_t s = vop->stream;
  auto run = [&]() -> int {
    if (V || D) {
      EBH_TRY(dV.alloc((size_t)ne * ne));
      hipLaunchKernelGGL(k_dense_zero, dim3((unsigned)(((size_t)ne * ne + 255) / 256)),
                         dim3(256), 0, s, (size_t)ne * ne, dV.p);
      hipLaunchKernelGGL(k_dense_v, dim3((ne + 15) / 16, (ne + 15) / 16), dim3(256), 0, s,
                         a, dV.p);
      hipLaunchKernelGGL(k_dense_add_csr, dim3((ne + 255) / 256), dim3(256), 0, s, ne, ne,
                         vop->near_ptr.p, vop->near_col.p, vop->near_val.p, dV.p);
      EBH_HIP(hipGetLastError());
    }
    if (K) {
      ebh_op_desc dk = d;
      dk.kernel = EBH_LAPLACE_DL;
      dk.trial_basis = EBH_P1;
      EBH_TRY(ebh_op_create(mesh, &dk, &kop));
      EBH_TRY(dK.alloc((size_t)ne * np));
      hipLaunchKernelGGL(k_dense_k, dim3((np + 15) / 16, (ne + 15) / 16), dim3(256), 0, s,
                         a, dK.p);
      hipLaunchKernelGGL(k_dense_add_csr, dim3((ne + 255) / 256), dim3(256), 0, s, ne, np,
                         kop->near_ptr.p, kop->near_col.p, kop->near_val.p, dK.p);
      EBH_HIP(hipGetLastError());
    }
    if (D) {
      std::vector<double> curl;
      surface_curls(T, curl);
      EBH_TRY(dcurl.upload(curl));
      EBH_TRY(dW.alloc((size_t)ne * np));
      EBH_TRY(dD.alloc((size_t)np * np));
      hipLaunchKernelGGL(k_dense_zero, dim3((unsigned)(((size_t)np * np + 255) / 256)),
                         dim3(256), 0, s, (size_t)np * np, dD.p);
      for (int c = 0; c < 3; c++) {
        hipLaunchKernelGGL(k_dense_vc, dim3((np + 15) / 16, (ne + 15) / 16), dim3(256), 0,
                           s, a, c, dcurl.p, dV.p, dW.p);
        hipLaunchKernelGGL(k_dense_ctw, dim3((np + 15) / 16, (np + 15) / 16), dim3(256), 0,
                           s, a, c, dcurl.p, dW.p, dD.p);
      }
      EBH_HIP(hipGetLastError());
    }
    EBH_HIP(hipStreamSynchronize(s));
    if (V) EBH_HIP(gated([&] { return hipMemcpy(V, dV.p, dV.bytes(), hipMemcpyDeviceToHost); }));
    if (K) EBH_HIP(gated([&] { return hipMemcpy(K, dK.p, dK.bytes(), hipMemcpyDeviceToHost); }));
    if (D) EBH_HIP(gated([&] { return hipMemcpy(D, dD.p, dD.bytes(), hipMemcpyDeviceToHost); }));
    return EBH_OK;
  };
  rc = run();
  ebh_op_destroy(kop);
  ebh_op_destroy(vop);
  return rc;
}

// getElasticity3d (src/BEM3dElasticity.cpp:604-952): V (3ne x 3ne), K
// (3ne x 3np), D (3np x 3np), M (3ne x 3np), component-blocked rows and
// columns, row-major; each output may be NULL.  V, K and D are the
// matrix-free operators' columns (one apply per column).
int ebh_assemble_elasticity(ebh_mesh *mesh, double nu, int order, double *V, double *K,
                            double *D, double *M)
{
  if (!mesh) { set_error("null mesh"); return EBH_ERR_ARG; }
  const Topology &T = mesh->T;
  const int ne = T.ne, np = T.np;
  EBH_HIP(hipSetDevice(mesh->device));
  if (M) {
    std::fill(M, M + (size_t)9 * ne * np, 0.0);
    for (int i = 0; i < ne; i++) {
      const int *f = &T.tri[3 * i];
      const double *a = &T.nodes[3 * f[0]], *b = &T.nodes[3 * f[1]], *c = &T.nodes[3 * f[2]];
      const double p[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      const double q[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      const double n0 = p[1] * q[2] - p[2] * q[1], n1 = p[2] * q[0] - p[0] * q[2],
                   n2 = p[0] * q[1] - p[1] * q[0];
      const double J = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      for (int a3 = 0; a3 < 3; a3++)
        for (int k = 0; k < 3; k++)
          M[(size_t)(a3 * ne + i) * 3 * np + a3 * np + f[k]] += J / 6.0;
    }
  }
  ebh_op_desc d{};
  d.far_precision = EBH_FP64;
  d.line_rule = d.tri_rule = order;
  d.max_rhs = 1;
  d.nu = nu;
  const int kinds[3] = {EBH_KELVIN_SL, EBH_KELVIN_DL, EBH_KELVIN_HS};
  const int tests[3] = {EBH_P0, EBH_P0, EBH_P1};
  const int trials[3] = {EBH_P0, EBH_P1, EBH_P1};
  double *outs[3] = {V, K, D};
  for (int w = 0; w < 3; w++) {
    if (!outs[w]) continue;
    d.kernel = kinds[w];
    d.test_basis = tests[w];
    d.trial_basis = trials[w];
    ebh_op *op = nullptr;
    EBH_TRY(ebh_op_create(mesh, &d, &op));
    const int rc = dense_by_columns(op, outs[w]);
    ebh_op_destroy(op);
    EBH_TRY(rc);
  }
  return EBH_OK;
}

}  // extern "C"
