// Near-field precompute: FP64 Sauter-Schwab integrals of every touching
// triangle pair (coincident, edge-adjacent, vertex-adjacent), one lane per
// pair, assembled into a CSR matrix that each apply multiplies in FP64.
//
// The reduced rules follow the reference's parameterisations
// (src/BEM3dLaplace.cpp:22-148 single layer, :195-383 double layer;
// src/BEM3dElasticity.cpp:30-286 Kelvin): the same difference vectors d_m
// and the same outer Gauss points.  The shape-function weights of each
// (region, outer point) come from host-built tables (host_geometry.cpp,
// ss_table), so one kernel yields the full 3x3 P1 block and every P0
// reduction of it.
#include <algorithm>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr double kInv4Pi = 0.07957747154594766788;  // 1/(4 pi)

struct V3 { double x, y, z; };
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 load3(const double *p, int n)
{
  return {p[3 * n], p[3 * n + 1], p[3 * n + 2]};
}
// d = a*p + b*q + c*r + e*s
__device__ __forceinline__ V3 lin4(double a, V3 p, double b, V3 q, double c,
                                   V3 r, double e, V3 s)
{
  return {a * p.x + b * q.x + c * r.x + e * s.x,
          a * p.y + b * q.y + c * r.y + e * s.y,
          a * p.z + b * q.z + c * r.z + e * s.z};
}

// Kernel values at difference vector d: KIND 0 -> 1/r, 1 -> d.n/r^3,
// 2 -> (1/r, d_a d_b / r^3) for the 7 Kelvin integrals.
template <int KIND>
struct NearAcc {
  static constexpr int NV = KIND == 2 ? 63 : 9;
  double a[NV];
  __device__ void zero()
  {
#pragma unroll
    for (int e = 0; e < NV; e++) a[e] = 0.0;
  }
  __device__ __forceinline__ void add(const double *Wq, V3 d, V3 n)
  {
    const double r2 = dot(d, d);
    const double r = sqrt(r2);
    if (KIND == 0) {
      const double kv = 1.0 / r;
#pragma unroll
      for (int e = 0; e < 9; e++) a[e] = fma(Wq[e], kv, a[e]);
    } else if (KIND == 1) {
      const double kv = dot(d, n) / (r2 * r);
#pragma unroll
      for (int e = 0; e < 9; e++) a[e] = fma(Wq[e], kv, a[e]);
    } else {
      const double ir = 1.0 / r, ir3 = ir / r2;
      const double kv[7] = {ir, d.x * d.x * ir3, d.y * d.y * ir3,
                            d.z * d.z * ir3, d.x * d.y * ir3,
                            d.x * d.z * ir3, d.y * d.z * ir3};
#pragma unroll
      for (int e = 0; e < 9; e++) {
        const double w = Wq[e];
#pragma unroll
        for (int c = 0; c < 7; c++) a[7 * e + c] = fma(w, kv[c], a[7 * e + c]);
      }
    }
  }
};

// One lane per pair of one class.  pinfo holds (oi[3], oj[3]) per pair,
// global node ids in the pair's local vertex orders; W the class table with
// the outer Gauss weights folded in; outer the outer coordinates.
template <int KIND>
__global__ __launch_bounds__(64) void k_near_pairs(
    int n, const int *__restrict__ ids, const int *__restrict__ pinfo, int cls,
    int nout, int nreg, const double *__restrict__ W,
    const double *__restrict__ outer, const double *__restrict__ nodes,
    double *__restrict__ pairbuf)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int p = ids[k];
  const int *pi = pinfo + 6 * p;
  const V3 I0 = load3(nodes, pi[0]), I1 = load3(nodes, pi[1]),
           I2 = load3(nodes, pi[2]), J0 = load3(nodes, pi[3]),
           J1 = load3(nodes, pi[4]), J2 = load3(nodes, pi[5]);
  V3 u, v, w = {0, 0, 0}, z = {0, 0, 0}, nrm = {0, 0, 0};
  double J;
  if (cls == kCoincident) {  // src/BEM3dLaplace.cpp:643-649
    u = sub(I1, I0);
    v = sub(I2, I1);
    J = dot(cross(u, v), cross(u, v));
  } else if (cls == kEdge) {  // src/BEM3dLaplace.cpp:745-759
    u = sub(J1, J0);
    v = sub(I2, J1);
    w = sub(J2, J1);
    const V3 c = cross(u, w);
    const double jt = sqrt(dot(c, c));
    nrm = {c.x / jt, c.y / jt, c.z / jt};
    J = sqrt(dot(cross(u, v), cross(u, v))) * jt;
  } else {  // vertex, src/BEM3dLaplace.cpp:733-740
    u = sub(I1, I0);
    v = sub(I2, I1);
    w = sub(J1, J0);
    z = sub(J2, J1);
    const V3 c = cross(w, z);
    const double jt = sqrt(dot(c, c));
    nrm = {c.x / jt, c.y / jt, c.z / jt};
    J = sqrt(dot(cross(u, v), cross(u, v))) * jt;
  }
  NearAcc<KIND> acc;
  acc.zero();
  if (!(KIND == 1 && cls == kCoincident)) {  // (x-y).n_y = 0 on a flat panel
    for (int q = 0; q < nout; q++) {
      const double o0 = outer[3 * q], o1 = outer[3 * q + 1], o2 = outer[3 * q + 2];
      for (int m = 0; m < nreg; m++) {
        V3 d;
        if (cls == kVertex) {
          if (m == 0) d = lin4(1.0, u, o0, v, -o1, w, -o1 * o2, z);
          else d = lin4(o1, u, o1 * o2, v, -1.0, w, -o0, z);
        } else if (cls == kEdge) {
          const double e2 = o0, e3 = o1;
          switch (m) {
          case 0: d = lin4(e2 * e3, u, 1.0, v, e2 * e3 - e2, w, 0.0, z); break;
          case 1: d = lin4(-e2, u, 1.0 - e2, v, -e2 * e3, w, 0.0, z); break;
          case 2: d = lin4(-e2 * e3, u, e2 - e2 * e3, v, -1.0, w, 0.0, z); break;
          case 3: d = lin4(-e2 * e3, u, 1.0 - e2 * e3, v, -e2, w, 0.0, z); break;
          default: d = lin4(e2, u, e3, v, e2 - 1.0, w, 0.0, z); break;
          }
        } else {
          const double e = o0;
          if (m == 0) d = lin4(e, u, 1.0, v, 0.0, w, 0.0, z);
          else if (m == 1) d = lin4(1.0, u, e, v, 0.0, w, 0.0, z);
          else d = lin4(e, u, e - 1.0, v, 0.0, w, 0.0, z);
        }
        acc.add(W + ((size_t)m * nout + q) * 9, d, nrm);
      }
    }
  }
  const double f = J * kInv4Pi;
  double *out = pairbuf + (size_t)p * NearAcc<KIND>::NV;
#pragma unroll
  for (int e = 0; e < NearAcc<KIND>::NV; e++) out[e] = acc.a[e] * f;
}

__device__ __forceinline__ double slot_sum(const double *v, int code, int stride)
{
  if (code < 9) return v[stride * code];
  if (code < 12) {
    const int l = code - 9;
    return v[stride * l] + v[stride * (3 + l)] + v[stride * (6 + l)];
  }
  if (code < 15) {
    const int k = code - 12;
    return v[stride * 3 * k] + v[stride * (3 * k + 1)] + v[stride * (3 * k + 2)];
  }
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < 9; e++) s += v[stride * e];
  return s;
}

// One lane per CSR entry; sums its contributions in the fixed host order.
__global__ void k_near_assemble(int nnz, const int *__restrict__ cptr,
                                const int64_t *__restrict__ contrib,
                                const double *__restrict__ pairbuf, int kelvin,
                                double nu, double *__restrict__ val)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  if (!kelvin) {
    double s = 0.0;
    for (int c = cptr[e]; c < cptr[e + 1]; c++) {
      const int64_t t = contrib[c];
      s += slot_sum(pairbuf + (t >> 4) * 9, (int)(t & 15), 1);
    }
    val[e] = s;
    return;
  }
  double I[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int c = cptr[e]; c < cptr[e + 1]; c++) {
    const int64_t t = contrib[c];
    const double *v = pairbuf + (t >> 4) * 63;
#pragma unroll
    for (int k = 0; k < 7; k++) I[k] += slot_sum(v + k, (int)(t & 15), 7);
  }
  // Kelvin block, src/BEM3dElasticity.cpp:612,838-846 (E = 1)
  const double f = (1.0 + nu) / 2.0 / (1.0 - nu), g = 3.0 - 4.0 * nu;
  double *o = val + (size_t)e * 6;
  o[0] = f * (g * I[0] + I[1]);
  o[1] = f * (g * I[0] + I[2]);
  o[2] = f * (g * I[0] + I[3]);
  o[3] = f * I[4];
  o[4] = f * I[5];
  o[5] = f * I[6];
}

}  // namespace

int op_build_near(ebh_op *op)
{
  const Topology &T = op->mesh->T;
  // the adjoint double layer's near field is the transpose of the double
  // layer's (test basis = the adjoint's trial basis, trial basis P1)
  const bool adj = op->d.kernel == EBH_LAPLACE_ADL;
  const int kind = adj ? EBH_LAPLACE_DL : op->d.kernel;
  const int nval = kind == EBH_KELVIN_SL ? 63 : 9;
  const size_t npairs = T.pairs.size();
  // pair info and per-class id lists
  std::vector<int> pinfo(6 * npairs);
  std::vector<int> ids[4];
  for (size_t p = 0; p < npairs; p++) {
    const NearPair &P = T.pairs[p];
    for (int k = 0; k < 3; k++) {
      pinfo[6 * p + k] = P.oi[k];
      pinfo[6 * p + 3 + k] = P.oj[k];
    }
    ids[P.cls].push_back((int)p);
  }
  DevBuf<int> d_pinfo;
  DevBuf<double> d_pairbuf;
  EBH_TRY(d_pinfo.upload(pinfo));
  EBH_TRY(d_pairbuf.alloc(npairs * nval));
  std::vector<double> gp, gw;
  if (!gauss01(op->nq, gp, gw)) {
    set_error("invalid near-field rule");
    return EBH_ERR_ARG;
  }
  for (int cls = kVertex; cls <= kCoincident; cls++) {
    if (ids[cls].empty()) continue;
    const int nq = op->nq, nreg = ss_nregions(cls), nout = ss_nouter(cls, nq);
    std::vector<double> W;
    ss_table(cls, kind == EBH_LAPLACE_DL ? 2 : 1, nq, W);
    std::vector<double> outer(3 * (size_t)nout, 0.0);
    std::vector<double> wq(nout);
    for (int q = 0; q < nout; q++) {
      if (cls == kVertex) {
        int a = q / (nq * nq), b = (q / nq) % nq, c = q % nq;
        outer[3 * q] = gp[a]; outer[3 * q + 1] = gp[b]; outer[3 * q + 2] = gp[c];
        wq[q] = gw[a] * gw[b] * gw[c];
      } else if (cls == kEdge) {
        int a = q / nq, b = q % nq;
        outer[3 * q] = gp[a]; outer[3 * q + 1] = gp[b];
        wq[q] = gw[a] * gw[b];
      } else {
        outer[3 * q] = gp[q];
        wq[q] = gw[q];
      }
    }
    for (int m = 0; m < nreg; m++)
      for (int q = 0; q < nout; q++)
        for (int e = 0; e < 9; e++) W[((size_t)m * nout + q) * 9 + e] *= wq[q];
    DevBuf<int> d_ids;
    DevBuf<double> d_W, d_outer;
    EBH_TRY(d_ids.upload(ids[cls]));
    EBH_TRY(d_W.upload(W));
    EBH_TRY(d_outer.upload(outer));
    const int n = (int)ids[cls].size();
    dim3 grid((n + 63) / 64), blk(64);
    switch (kind) {
    case EBH_LAPLACE_SL:
      hipLaunchKernelGGL(k_near_pairs<0>, grid, blk, 0, op->stream, n, d_ids.p,
                         d_pinfo.p, cls, nout, nreg, d_W.p, d_outer.p,
                         op->mesh->nodes.p, d_pairbuf.p);
      break;
    case EBH_LAPLACE_DL:
      hipLaunchKernelGGL(k_near_pairs<1>, grid, blk, 0, op->stream, n, d_ids.p,
                         d_pinfo.p, cls, nout, nreg, d_W.p, d_outer.p,
                         op->mesh->nodes.p, d_pairbuf.p);
      break;
    default:
      hipLaunchKernelGGL(k_near_pairs<2>, grid, blk, 0, op->stream, n, d_ids.p,
                         d_pinfo.p, cls, nout, nreg, d_W.p, d_outer.p,
                         op->mesh->nodes.p, d_pairbuf.p);
      break;
    }
    EBH_HIP(hipGetLastError());
    EBH_HIP(hipStreamSynchronize(op->stream));
  }
  const std::pair<int, int> bases(adj ? op->d.trial_basis : op->d.test_basis,
                                  adj ? EBH_P1 : op->d.trial_basis);
  std::shared_ptr<const NearCSR> shared;
  {
    std::lock_guard<std::mutex> lk(op->mesh->csr_mu);
    auto it = op->mesh->csr_cache.find(bases);
    if (it != op->mesh->csr_cache.end()) {
      shared = it->second;
    } else {
      auto built = std::make_shared<NearCSR>();
      build_near_csr(T, bases.first, bases.second, *built);
      shared = built;
      op->mesh->csr_cache[bases] = shared;
    }
  }
  const NearCSR &csr = *shared;
  DevBuf<int> d_cptr;
  DevBuf<int64_t> d_contrib;
  EBH_TRY(op->near_ptr.upload(csr.row_ptr));
  EBH_TRY(op->near_col.upload(csr.col));
  EBH_TRY(d_cptr.upload(csr.contrib_ptr));
  EBH_TRY(d_contrib.upload(csr.contrib));
  const int nnz = (int)csr.col.size();
  op->near_nnz = nnz;
  const int kel = kind == EBH_KELVIN_SL;
  EBH_TRY(op->near_val.alloc((size_t)nnz * (kel ? 6 : 1)));
  hipLaunchKernelGGL(k_near_assemble, dim3((nnz + 255) / 256), dim3(256), 0,
                     op->stream, nnz, d_cptr.p, d_contrib.p, d_pairbuf.p, kel,
                     op->d.nu, op->near_val.p);
  EBH_HIP(hipGetLastError());
  EBH_HIP(hipStreamSynchronize(op->stream));
  if (adj) {
    // transpose on the host once: rows become the P1 test nodes
    std::vector<double> val(nnz), tval(nnz);
    EBH_HIP(gated([&] { return hipMemcpy(val.data(), op->near_val.p, sizeof(double) * nnz,
                      hipMemcpyDeviceToHost); }));
    const int nr = csr.nrows, nc = T.np;
    std::vector<int> tptr(nc + 1, 0), tcol(nnz);
    for (int e = 0; e < nnz; e++) tptr[csr.col[e] + 1]++;
    for (int c = 0; c < nc; c++) tptr[c + 1] += tptr[c];
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (int r = 0; r < nr; r++)
      for (int e = csr.row_ptr[r]; e < csr.row_ptr[r + 1]; e++) {
        const int d = fill[csr.col[e]]++;
        tcol[d] = r;
        tval[d] = val[e];
      }
    EBH_TRY(op->near_ptr.upload(tptr));
    EBH_TRY(op->near_col.upload(tcol));
    EBH_TRY(op->near_val.upload(tval));
  }
  return EBH_OK;
}

}  // namespace ebh
