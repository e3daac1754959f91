// Laplace hypersingular operator on P1 through surface curls:
//
//   D = sum_{c=1..3} C_c^T V C_c   (+ stab m m^T)
//
// with V the P0 single layer, (C_c)_{i,n} the c-th Cartesian component of
// the surface curl of hat function n on triangle i (getLaplace3dCurl,
// src/BEM3dLaplace.cpp:489-558, and the dense D assembly at :652,769), and m
// the lumped node areas, m_n = sum_{i ni n} |T_i|/3
// (python/examples/laplace3dNeumannGPU.py:36-49 adds m (m . x) to fix the
// constants).  One apply: a curl kernel (three P0 densities per right-hand
// side), the wrapped single layer on 3 nrhs right-hand sides, and a gather
// over each node's triangles; no atomics, a fixed summation order.
#include <algorithm>
#include <cmath>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr int kDotBlocks = 64;

// w[(3r + c) ne + i] = sum_k curl[(3i + c) 3 + k] x[r np + tri[3i + k]]
__global__ __launch_bounds__(256) void k_hs_curl(int ne, int np,
                                                 const double *__restrict__ curl,
                                                 const int *__restrict__ tri,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ w)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= ne) return;
  const double *xr = x + (size_t)r * np;
  const int n0 = tri[3 * i], n1 = tri[3 * i + 1], n2 = tri[3 * i + 2];
  const double x0 = xr[n0], x1 = xr[n1], x2 = xr[n2];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double *o = curl + (3 * (size_t)i + c) * 3;
    w[(3 * (size_t)r + c) * ne + i] = fma(o[0], x0, fma(o[1], x1, o[2] * x2));
  }
}

// part[r kDotBlocks + b] = sum over block b's strided nodes of m_n x_n
__global__ __launch_bounds__(256) void k_hs_dot(int np, const double *__restrict__ m,
                                                const double *__restrict__ x,
                                                double *__restrict__ part)
{
  __shared__ double red[256];
  const int r = blockIdx.y;
  const double *xr = x + (size_t)r * np;
  double s = 0.0;
  for (int n = blockIdx.x * 256 + threadIdx.x; n < np; n += kDotBlocks * 256)
    s = fma(m[n], xr[n], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[r * kDotBlocks + blockIdx.x] = red[0];
}

// y[r np + n] = sum_{(i,k) at n} sum_c curl[(3i + c) 3 + k] z[(3r + c) ne + i]
//               + stab m_n (m . x_r)
__global__ __launch_bounds__(256) void k_hs_gather(
    int ne, int np, double stab, const double *__restrict__ curl,
    const int *__restrict__ ptr, const int *__restrict__ ntri,
    const int *__restrict__ nk, const double *__restrict__ z,
    const double *__restrict__ m, const double *__restrict__ part,
    double *__restrict__ y)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (n >= np) return;
  const double *zr = z + 3 * (size_t)r * ne;
  double acc = 0.0;
  for (int e = ptr[n]; e < ptr[n + 1]; e++) {
    const int i = ntri[e], k = nk[e];
#pragma unroll
    for (int c = 0; c < 3; c++)
      acc = fma(curl[(3 * (size_t)i + c) * 3 + k], zr[(size_t)c * ne + i], acc);
  }
  if (stab != 0.0) {
    double s = 0.0;
    for (int b = 0; b < kDotBlocks; b++) s += part[r * kDotBlocks + b];
    acc = fma(stab * m[n], s, acc);
  }
  y[(size_t)r * np + n] = acc;
}

// Chebyshev iteration for M z = v preconditioned by the lumped mass M_L
// (Saad, Iterative Methods for Sparse Linear Systems, alg. 12.1): for P1
// elements the eigenvalues of M_L^-1 M lie in [1/4, 1] on any triangle
// mesh (each element's are 1, 1/4, 1/4).  A fixed number of steps is a
// fixed polynomial in M_L^-1 M times M_L^-1, a linear operator.
constexpr double kChebLo = 0.25, kChebHi = 1.0;
// Chebyshev steps of the two Calderon preconditioners.  FP64 GMRES(50)
// on V of the L5 icosphere takes 11, 6, 5, 4, 5 iterations with 2..6
// steps, 4 with the exact mass inverse; on D + m m^T of the L4 and L5
// serving meshes 7 with 3 steps and 6 with 5.
constexpr int kCalderonVSteps = 5, kCalderonHsSteps = 5;

// r = v, d = M_L^-1 v / theta, z = d; blockIdx.y: right-hand side
__global__ void k_cheb_init(int n, const double *v, const double *__restrict__ dinv,
                            double inv_theta, double *__restrict__ r,
                            double *__restrict__ d, double *z)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t k = (size_t)blockIdx.y * n + i;
  const double vi = v[k];
  const double di = dinv[i] * vi * inv_theta;
  r[k] = vi;
  d[k] = di;
  z[k] = di;  // z may alias v: read above, by the same thread
}

// One step, row i of the mass matrix per thread: r -= M d, dn = a d +
// b M_L^-1 r, z += dn (d and dn distinct: other rows still read d)
__global__ void k_cheb_step(int n, const int *__restrict__ ptr,
                            const int *__restrict__ col,
                            const double *__restrict__ val,
                            const double *__restrict__ dinv, double a, double b,
                            double *__restrict__ r, const double *__restrict__ d,
                            double *__restrict__ dn, double *__restrict__ z)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t off = (size_t)blockIdx.y * n;
  const double *dr = d + off;
  double t = 0.0;
  for (int e = ptr[i]; e < ptr[i + 1]; e++) t = fma(val[e], dr[col[e]], t);
  const size_t k = off + i;
  const double ri = r[k] - t;
  const double di = fma(a, dr[i], b * dinv[i] * ri);
  r[k] = ri;
  dn[k] = di;
  z[k] += di;
}

}  // namespace

int mass_inv_apply(ebh_op *op, const double *v, double *z, int nrhs, hipStream_t s)
{
  const int n = op->mesh->T.np;
  const dim3 grid((n + 255) / 256, nrhs), blk(256);
  const double th = 0.5 * (kChebHi + kChebLo), de = 0.5 * (kChebHi - kChebLo);
  const double sig = th / de;
  double rho = 1.0 / sig;
  // the search direction alternates between two buffers
  double *d = op->cheb_d.p, *dn = op->cheb_t.p;
  hipLaunchKernelGGL(k_cheb_init, grid, blk, 0, s, n, v, op->dinv.p, 1.0 / th,
                     op->cheb_r.p, d, z);
  EBH_HIP(hipGetLastError());
  for (int k = 1; k < op->pcheb; k++) {
    const double rn = 1.0 / (2.0 * sig - rho);
    hipLaunchKernelGGL(k_cheb_step, grid, blk, 0, s, n, op->mass_ptr.p, op->mass_col.p,
                       op->mass_val.p, op->dinv.p, rn * rho, 2.0 * rn / de, op->cheb_r.p,
                       d, dn, z);
    EBH_HIP(hipGetLastError());
    std::swap(d, dn);
    rho = rn;
  }
  return EBH_OK;
}

// Consistent P1 mass matrix of the operator's mesh as CSR, the inverse
// lumped masses, and Chebyshev work vectors for nb right-hand sides; the
// mass inverse then takes `steps` Chebyshev steps (mass_inv_apply)
int mass_setup(ebh_op *op, int steps, int nb)
{
  const Topology &T = op->mesh->T;
  struct Trip { int r, c; double v; };
  std::vector<Trip> trip;
  trip.reserve((size_t)9 * T.ne);
  std::vector<double> m((size_t)T.np, 0.0);
  for (int i = 0; i < T.ne; i++) {
    const double *P0 = &T.nodes[3 * T.tri[3 * i]];
    const double *P1 = &T.nodes[3 * T.tri[3 * i + 1]];
    const double *P2 = &T.nodes[3 * T.tri[3 * i + 2]];
    double u[3], v[3];
    for (int k = 0; k < 3; k++) { u[k] = P1[k] - P0[k]; v[k] = P2[k] - P0[k]; }
    const double c0 = u[1] * v[2] - u[2] * v[1], c1 = u[2] * v[0] - u[0] * v[2],
                 c2 = u[0] * v[1] - u[1] * v[0];
    const double J = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    for (int k = 0; k < 3; k++) m[T.tri[3 * i + k]] += J / 6.0;
    // consistent P1 mass: |T|/12 (1 + delta_ab) = J/24 (1 + delta_ab)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        trip.push_back({T.tri[3 * i + a], T.tri[3 * i + b], (a == b ? 2.0 : 1.0) * J / 24.0});
  }
  for (double &v : m) v = 1.0 / v;
  // triplets sorted by (row, column), equal ones summed in triangle order
  std::stable_sort(trip.begin(), trip.end(), [](const Trip &x, const Trip &y) {
    return x.r != y.r ? x.r < y.r : x.c < y.c;
  });
  std::vector<int> mptr(T.np + 1, 0), mcol;
  std::vector<double> mval;
  for (size_t k = 0; k < trip.size(); k++) {
    if (k > 0 && trip[k].r == trip[k - 1].r && trip[k].c == trip[k - 1].c) {
      mval.back() += trip[k].v;
      continue;
    }
    mcol.push_back(trip[k].c);
    mval.push_back(trip[k].v);
    mptr[trip[k].r + 1]++;
  }
  for (int r = 0; r < T.np; r++) mptr[r + 1] += mptr[r];
  const size_t nv = (size_t)T.np * std::max(1, nb);
  EBH_TRY(op->dinv.upload(m));
  EBH_TRY(op->mass_ptr.upload(mptr));
  EBH_TRY(op->mass_col.upload(mcol));
  EBH_TRY(op->mass_val.upload(mval));
  EBH_TRY(op->cheb_r.alloc(nv));
  EBH_TRY(op->cheb_d.alloc(nv));
  EBH_TRY(op->cheb_t.alloc(nv));
  op->pcheb = steps;
  return EBH_OK;
}

int hs_build(ebh_op *op)
{
  const Topology &T = op->mesh->T;
  std::vector<double> curl, m((size_t)T.np, 0.0);
  surface_curls(T, curl);
  for (int i = 0; i < T.ne; i++) {
    const double *P0 = &T.nodes[3 * T.tri[3 * i]];
    const double *P1 = &T.nodes[3 * T.tri[3 * i + 1]];
    const double *P2 = &T.nodes[3 * T.tri[3 * i + 2]];
    double u[3], v[3];
    for (int k = 0; k < 3; k++) { u[k] = P1[k] - P0[k]; v[k] = P2[k] - P0[k]; }
    const double c0 = u[1] * v[2] - u[2] * v[1], c1 = u[2] * v[0] - u[0] * v[2],
                 c2 = u[0] * v[1] - u[1] * v[0];
    const double J = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    for (int k = 0; k < 3; k++) m[T.tri[3 * i + k]] += J / 6.0;
  }
  EBH_TRY(op->curl.upload(curl));
  EBH_TRY(op->mvec.upload(m));
  const size_t nb = (size_t)std::max(1, op->d.max_rhs);
  EBH_TRY(op->hsw.alloc(3 * (size_t)T.ne * nb));
  EBH_TRY(op->hsz.alloc(3 * (size_t)T.ne * nb));
  EBH_TRY(op->hsdot.alloc(kDotBlocks * nb));
  return EBH_OK;
}

int hs_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s,
             bool timed)
{
  if (nrhs > std::max(1, op->d.max_rhs)) {
    set_error("nrhs " + std::to_string(nrhs) + " exceeds the operator's max_rhs " +
              std::to_string(op->d.max_rhs));
    return EBH_ERR_ARG;
  }
  const int ne = op->mesh->T.ne, np = op->mesh->T.np;
  if (timed) EBH_HIP(hipEventRecord(op->ev[0], s));
  if (op->hs_scale) {
    // Calderon form with the Chebyshev mass inverse Q on both sides
    EBH_TRY(mass_inv_apply(op, x, op->chebx.p, nrhs, s));
    x = op->chebx.p;
  }
  hipLaunchKernelGGL(k_hs_curl, dim3((ne + 255) / 256, nrhs), dim3(256), 0, s, ne, np,
                     op->curl.p, op->mesh->tri.p, x, op->hsw.p);
  EBH_HIP(hipGetLastError());
  if (op->d.stab != 0.0) {
    hipLaunchKernelGGL(k_hs_dot, dim3(kDotBlocks, nrhs), dim3(256), 0, s, np,
                       op->mvec.p, x, op->hsdot.p);
    EBH_HIP(hipGetLastError());
  }
  if (timed) EBH_HIP(hipEventRecord(op->ev[1], s));
  EBH_TRY(op_apply_multi(op->inner, op->hsw.p, op->hsz.p, 3 * nrhs, s, false));
  if (timed) EBH_HIP(hipEventRecord(op->ev[2], s));
  hipLaunchKernelGGL(k_hs_gather, dim3((np + 255) / 256, nrhs), dim3(256), 0, s, ne,
                     np, op->d.stab, op->curl.p, op->mesh->node_tri_ptr.p,
                     op->mesh->node_tri.p, op->mesh->node_tri_k.p, op->hsz.p,
                     op->mvec.p, op->hsdot.p, y);
  EBH_HIP(hipGetLastError());
  if (op->hs_scale) EBH_TRY(mass_inv_apply(op, y, y, nrhs, s));
  if (timed) EBH_HIP(hipEventRecord(op->ev[3], s));
  return EBH_OK;
}

}  // namespace ebh

using namespace ebh;

extern "C" int ebh_precond_calderon(ebh_mesh *mesh, int far_precision, int line_rule,
                                    int tri_rule, int max_rhs, ebh_op **out)
{
  if (!mesh || !out) { set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  double area = 0.0;
  const Topology &T = mesh->T;
  for (int i = 0; i < T.ne; i++) {
    const double *P0 = &T.nodes[3 * T.tri[3 * i]];
    const double *P1 = &T.nodes[3 * T.tri[3 * i + 1]];
    const double *P2 = &T.nodes[3 * T.tri[3 * i + 2]];
    double u[3], v[3];
    for (int k = 0; k < 3; k++) { u[k] = P1[k] - P0[k]; v[k] = P2[k] - P0[k]; }
    const double c0 = u[1] * v[2] - u[2] * v[1], c1 = u[2] * v[0] - u[0] * v[2],
                 c2 = u[0] * v[1] - u[1] * v[0];
    area += 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
  ebh_op_desc d{};
  d.kernel = EBH_LAPLACE_HS;
  d.test_basis = d.trial_basis = EBH_P1;
  d.far_precision = far_precision;
  d.line_rule = line_rule;
  d.tri_rule = tri_rule;
  d.max_rhs = max_rhs;
  // constants: V 1 ~ 1 on the unit sphere, so the rank-one term maps them
  // to about 1/4 as D V does the other harmonics
  d.stab = 0.25 / area;
  ebh_op *op = nullptr;
  EBH_TRY(ebh_op_create(mesh, &d, &op));
  int rc = mass_setup(op, kCalderonVSteps, max_rhs);
  if (rc == EBH_OK) rc = op->chebx.alloc((size_t)T.np * std::max(1, max_rhs));
  if (rc != EBH_OK) {
    ebh_op_destroy(op);
    return rc;
  }
  op->hs_scale = true;
  *out = op;
  return EBH_OK;
}

extern "C" int ebh_precond_calderon_hs(ebh_mesh *mesh, int far_precision, int line_rule,
                                       int tri_rule, ebh_op **out)
{
  if (!mesh || !out) { set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  const Topology &T = mesh->T;
  ebh_op_desc d{};
  d.kernel = EBH_LAPLACE_SL;
  d.test_basis = d.trial_basis = EBH_P1;
  d.far_precision = far_precision;
  d.line_rule = line_rule;
  d.tri_rule = tri_rule;
  d.max_rhs = 1;
  ebh_op *op = nullptr;
  EBH_TRY(ebh_op_create(mesh, &d, &op));
  int rc = op->hsw.alloc(T.np);
  if (rc == EBH_OK) rc = mass_setup(op, kCalderonHsSteps, 1);
  if (rc != EBH_OK) {
    ebh_op_destroy(op);
    return rc;
  }
  op->pscale = true;
  *out = op;
  return EBH_OK;
}

