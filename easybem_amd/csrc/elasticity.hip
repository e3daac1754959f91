// Elasticity double layer K and hypersingular D (getElasticity3d,
// src/BEM3dElasticity.cpp:604-952), applied through operators this library
// already has.  With the curls c_{j,k} of triangle j's hat functions, V the
// Laplace single layer, K_L the Laplace double layer (P0 x P1), V_E the
// Kelvin single layer (factor (1+nu)/(2(1-nu)), fact 3-4nu) and, per P1
// field x = (x_1, x_2, x_3),
//
//   w(j)      = sum_k c_{j,k} x x(node_k)          (3 values per triangle)
//   G_mb(j)   = sum_k c_{j,k,m} x_b(node_k)        (9 values per triangle)
//
// the reference's pair loop (:760-943) sums to
//
//   K x   = K_L x_a  +  V_E w / (1+nu)  -  V w_a
//   D x   = mu [4 sum_m C_m^T V G_ma - 2 sum_b C_b^T V G_ab
//              - C_a^T V sum_b G_bb]  -  4 mu^2 W^T V_E w,
//
// mu = 1/(2(1+nu)), with (W^T y)_a(node k) = sum_{i at k} (y(i) x c_{i,k})_a.
// The K coefficients factorK (factK I + I_ab) of :851-858 equal
// V_E/(1+nu) - I delta_ab, since factorK (factK - fact) = -1.
#include <algorithm>
#include <cmath>

#include "ebh_internal.h"

namespace ebh {
namespace {

// w[c ne + j] (3 per triangle) and, when G is given, G[(3m + b) ne + j]
__global__ __launch_bounds__(256) void k_el_curl(int ne, int np, int nrhs_off,
                                                 const double *__restrict__ curl,
                                                 const int *__restrict__ tri,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ w,
                                                 double *__restrict__ G)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= ne) return;
  const double *xr = x + (size_t)r * 3 * np;
  double g[3][3];  // g[m][b] = sum_k c_{j,k,m} x_b(node_k)
  for (int m = 0; m < 3; m++)
    for (int b = 0; b < 3; b++) g[m][b] = 0.0;
  for (int k = 0; k < 3; k++) {
    const int n = tri[3 * j + k];
    const double xb[3] = {xr[n], xr[(size_t)np + n], xr[2 * (size_t)np + n]};
    for (int m = 0; m < 3; m++) {
      const double c = curl[(3 * (size_t)j + m) * 3 + k];
      for (int b = 0; b < 3; b++) g[m][b] = fma(c, xb[b], g[m][b]);
    }
  }
  double *wr = w + (size_t)r * 3 * ne;
  // (c x x)_m = g[m+1][m+2] - g[m+2][m+1]
  wr[j] = g[1][2] - g[2][1];
  wr[(size_t)ne + j] = g[2][0] - g[0][2];
  wr[2 * (size_t)ne + j] = g[0][1] - g[1][0];
  if (G) {
    double *Gr = G + (size_t)r * 9 * ne;
    for (int m = 0; m < 3; m++)
      for (int b = 0; b < 3; b++) Gr[(size_t)(3 * m + b) * ne + j] = g[m][b];
  }
  (void)nrhs_off;
}

// y = y + s1 t1 + s2 t2 (n entries)
__global__ void k_el_axpy2(size_t n, double s1, const double *__restrict__ t1, double s2,
                           const double *__restrict__ t2, double *__restrict__ y)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fma(s2, t2[i], fma(s1, t1[i], y[i]));
}

// y_a(node k) = sum_{i at k} mu (4 sum_m c_{i,k,m} VG_ma - 2 sum_b c_{i,k,b} VG_ab
//               - c_{i,k,a} sum_b VG_bb) - 4 mu^2 (Vw(i) x c_{i,k})_a
__global__ __launch_bounds__(256) void k_el_hs_gather(
    int ne, int np, double mu, const double *__restrict__ curl,
    const int *__restrict__ ptr, const int *__restrict__ ntri, const int *__restrict__ nk,
    const double *__restrict__ VG, const double *__restrict__ Vw, double *__restrict__ y)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (n >= np) return;
  const double *VGr = VG + (size_t)r * 9 * ne;
  const double *Vwr = Vw + (size_t)r * 3 * ne;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int e = ptr[n]; e < ptr[n + 1]; e++) {
    const int i = ntri[e], k = nk[e];
    double c[3], vg[3][3], v[3];
    for (int m = 0; m < 3; m++) {
      c[m] = curl[(3 * (size_t)i + m) * 3 + k];
      v[m] = Vwr[(size_t)m * ne + i];
      for (int b = 0; b < 3; b++) vg[m][b] = VGr[(size_t)(3 * m + b) * ne + i];
    }
    const double tr = vg[0][0] + vg[1][1] + vg[2][2];
    const double vxc[3] = {v[1] * c[2] - v[2] * c[1], v[2] * c[0] - v[0] * c[2],
                           v[0] * c[1] - v[1] * c[0]};
    for (int a = 0; a < 3; a++) {
      double s = 0.0;
      for (int m = 0; m < 3; m++) s = fma(4.0 * c[m], vg[m][a], fma(-2.0 * c[m], vg[a][m], s));
      s -= c[a] * tr;
      acc[a] = fma(mu, s, fma(-4.0 * mu * mu, vxc[a], acc[a]));
    }
  }
  double *yr = y + (size_t)r * 3 * np;
  for (int a = 0; a < 3; a++) yr[(size_t)a * np + n] = acc[a];
}

}  // namespace

int kel_build(ebh_op *op)
{
  const Topology &T = op->mesh->T;
  std::vector<double> curl;
  surface_curls(T, curl);
  EBH_TRY(op->curl.upload(curl));
  const size_t nb = (size_t)std::max(1, op->d.max_rhs);
  const size_t ne = T.ne;
  EBH_TRY(op->hsw.alloc(3 * ne * nb));   // w
  EBH_TRY(op->kt1.alloc(3 * ne * nb));   // V_E w
  if (op->d.kernel == EBH_KELVIN_DL) {
    EBH_TRY(op->kt2.alloc(3 * ne * nb));  // V w_a
  } else {
    EBH_TRY(op->hsz.alloc(9 * ne * nb));  // G
    EBH_TRY(op->kt2.alloc(9 * ne * nb));  // V G
  }
  return EBH_OK;
}

int kel_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s, bool timed)
{
  if (nrhs > std::max(1, op->d.max_rhs)) {
    set_error("nrhs " + std::to_string(nrhs) + " exceeds the operator's max_rhs " +
              std::to_string(op->d.max_rhs));
    return EBH_ERR_ARG;
  }
  const int ne = op->mesh->T.ne, np = op->mesh->T.np;
  const double nu = op->d.nu;
  const bool dl = op->d.kernel == EBH_KELVIN_DL;
  if (timed) EBH_HIP(hipEventRecord(op->ev[0], s));
  hipLaunchKernelGGL(k_el_curl, dim3((ne + 255) / 256, nrhs), dim3(256), 0, s, ne, np, 0,
                     op->curl.p, op->mesh->tri.p, x, op->hsw.p, dl ? nullptr : op->hsz.p);
  EBH_HIP(hipGetLastError());
  if (timed) EBH_HIP(hipEventRecord(op->ev[1], s));
  // V_E w: the Kelvin single layer on 3 components per triangle
  EBH_TRY(op_apply_multi(op->inner_kel, op->hsw.p, op->kt1.p, nrhs, s, false));
  if (dl) {
    // K_L x_a for the three components, then V w_a
    EBH_TRY(op_apply_multi(op->inner_dl, x, y, 3 * nrhs, s, false));
    EBH_TRY(op_apply_multi(op->inner, op->hsw.p, op->kt2.p, 3 * nrhs, s, false));
    if (timed) EBH_HIP(hipEventRecord(op->ev[2], s));
    const size_t n = (size_t)3 * ne * nrhs;
    hipLaunchKernelGGL(k_el_axpy2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       1.0 / (1.0 + nu), op->kt1.p, -1.0, op->kt2.p, y);
    EBH_HIP(hipGetLastError());
  } else {
    EBH_TRY(op_apply_multi(op->inner, op->hsz.p, op->kt2.p, 9 * nrhs, s, false));
    if (timed) EBH_HIP(hipEventRecord(op->ev[2], s));
    hipLaunchKernelGGL(k_el_hs_gather, dim3((np + 255) / 256, nrhs), dim3(256), 0, s, ne,
                       np, 0.5 / (1.0 + nu), op->curl.p, op->mesh->node_tri_ptr.p,
                       op->mesh->node_tri.p, op->mesh->node_tri_k.p, op->kt2.p,
                       op->kt1.p, y);
    EBH_HIP(hipGetLastError());
  }
  if (timed) EBH_HIP(hipEventRecord(op->ev[3], s));
  return EBH_OK;
}

int composite_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s,
                    bool timed)
{
  if (op->d.kernel == EBH_LAPLACE_HS) return hs_apply(op, x, y, nrhs, s, timed);
  return kel_apply(op, x, y, nrhs, s, timed);
}

}  // namespace ebh
