// Near-field sparse approximate inverse of the single layer or of the
// hypersingular operator, a right preconditioner for GMRES (no counterpart
// in the reference).
//
// Row c of M is row c of the inverse of V restricted to a patch around DOF
// c: the DOFs within `rings` adjacency steps of c, P1 nodes being adjacent
// when they share a triangle and P0 triangles when they share a vertex (on
// a regular mesh 19 or 37 nodes for 2 or 3 steps, 13 triangles for 1).
// M V is close to the identity on the high-frequency part of the spectrum,
// where V's eigenvalues decay (1/(2l+1) on the sphere), so the GMRES
// iteration count stays nearly flat under refinement.
//
// Build: the entries V_ab for every pair (a, b) that shares a patch are
// computed exactly -- the near-field CSR of V gives the touching-pair part,
// k_spai_far the far part with V's own triangle rule -- then every patch
// matrix is solved on the host (partial pivoting).  Apply: one sparse
// matrix-vector product in FP64, 8 lanes per row.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr double kInv4PiSpai = 0.07957747154594766788;
constexpr int kSpaiLanes = 8;

// 1/sqrt(x): the hardware FP64 estimate (relative error below 6e-8) and one
// Newton step (below 5e-15)
__device__ __forceinline__ double rsq64(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}

// Far part of V_ab for the pattern entry e = (row[e], col[e]): the sum over
// triangle pairs (i in supp a, j in supp b) that do not touch of
// sum_q sum_p w_q J_i phi_a(x_q) w_p J_j phi_b(y_p) / (4 pi |x_q - y_p|).
// One thread per entry; P0 supports are one triangle with phi = 1.
__global__ __launch_bounds__(256) void k_spai_far(int nent, const int *__restrict__ erow,
                                                  const int *__restrict__ ecol, int p1,
                                                  RuleArgs rule,
                                                  const double *__restrict__ nodes,
                                                  const int *__restrict__ tri,
                                                  const int *__restrict__ ntp,
                                                  const int *__restrict__ ntri,
                                                  const int *__restrict__ ntk,
                                                  const int *__restrict__ touch_ptr,
                                                  const int *__restrict__ touch,
                                                  double *__restrict__ eval)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int a = erow[e], b = ecol[e];
  const int nt = rule.nt;
  const int ia0 = p1 ? ntp[a] : 0, ia1 = p1 ? ntp[a + 1] : 1;
  const int ib0 = p1 ? ntp[b] : 0, ib1 = p1 ? ntp[b + 1] : 1;
  double s = 0.0;
  for (int ia = ia0; ia < ia1; ia++) {
    const int i = p1 ? ntri[ia] : a, ki = p1 ? ntk[ia] : -1;
    double X[16][3], W[16];
    {
      const double *P0 = nodes + 3 * tri[3 * i], *P1 = nodes + 3 * tri[3 * i + 1],
                   *P2 = nodes + 3 * tri[3 * i + 2];
      const double u[3] = {P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2]};
      const double v[3] = {P2[0] - P0[0], P2[1] - P0[1], P2[2] - P0[2]};
      const double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
                   n2 = u[0] * v[1] - u[1] * v[0];
      const double J = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      for (int q = 0; q < nt; q++) {
        const double x1 = rule.x[q], x2 = rule.y[q];
        for (int d = 0; d < 3; d++) X[q][d] = P0[d] + x1 * u[d] + x2 * v[d];
        const double phi = ki < 0 ? 1.0 : ki == 0 ? 1.0 - x1 - x2 : ki == 1 ? x1 : x2;
        W[q] = rule.w[q] * J * phi;
      }
    }
    for (int ib = ib0; ib < ib1; ib++) {
      const int j = p1 ? ntri[ib] : b, kj = p1 ? ntk[ib] : -1;
      bool touching = false;
      for (int c = touch_ptr[i]; c < touch_ptr[i + 1]; c++) touching |= touch[c] == j;
      if (touching) continue;  // in the near-field CSR
      const double *P0 = nodes + 3 * tri[3 * j], *P1 = nodes + 3 * tri[3 * j + 1],
                   *P2 = nodes + 3 * tri[3 * j + 2];
      const double u[3] = {P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2]};
      const double v[3] = {P2[0] - P0[0], P2[1] - P0[1], P2[2] - P0[2]};
      const double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
                   n2 = u[0] * v[1] - u[1] * v[0];
      const double J = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      for (int p = 0; p < nt; p++) {
        const double y1 = rule.x[p], y2 = rule.y[p];
        const double phi = kj < 0 ? 1.0 : kj == 0 ? 1.0 - y1 - y2 : kj == 1 ? y1 : y2;
        const double wp = rule.w[p] * J * phi;
        const double Y[3] = {P0[0] + y1 * u[0] + y2 * v[0], P0[1] + y1 * u[1] + y2 * v[1],
                             P0[2] + y1 * u[2] + y2 * v[2]};
        double t = 0.0;
        for (int q = 0; q < nt; q++) {
          const double dx = X[q][0] - Y[0], dy = X[q][1] - Y[1], dz = X[q][2] - Y[2];
          t += W[q] * rsq64(dx * dx + dy * dy + dz * dz);
        }
        s += wp * t;
      }
    }
  }
  eval[e] = s * kInv4PiSpai;
}

// Entries of the hypersingular operator D = sum_c C_c^T V0 C_c (V0 the P0
// single layer, C_c the surface-curl components of the hat functions) for
// the pattern entries (a, b) of P1 nodes:
//   D_ab = sum_{i in supp a} sum_{j in supp b} (curl_i phi_a . curl_j phi_b) V0_ij,
// V0_ij from V0's near-field CSR for touching triangles, else from the far
// rule.  One thread per entry.
__global__ __launch_bounds__(256) void k_spai_hs(int nent, const int *__restrict__ erow,
                                                 const int *__restrict__ ecol, RuleArgs rule,
                                                 const double *__restrict__ nodes,
                                                 const int *__restrict__ tri,
                                                 const int *__restrict__ ntp,
                                                 const int *__restrict__ ntri,
                                                 const int *__restrict__ ntk,
                                                 const int *__restrict__ touch_ptr,
                                                 const int *__restrict__ touch,
                                                 const double *__restrict__ curl,
                                                 const int *__restrict__ vptr,
                                                 const int *__restrict__ vcol,
                                                 const double *__restrict__ vval,
                                                 double *__restrict__ eval)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int a = erow[e], b = ecol[e];
  const int nt = rule.nt;
  double s = 0.0;
  for (int ia = ntp[a]; ia < ntp[a + 1]; ia++) {
    const int i = ntri[ia], ki = ntk[ia];
    double ci[3];
    for (int c = 0; c < 3; c++) ci[c] = curl[(3 * (size_t)i + c) * 3 + ki];
    double X[16][3], W[16];
    {
      const double *P0 = nodes + 3 * tri[3 * i], *P1 = nodes + 3 * tri[3 * i + 1],
                   *P2 = nodes + 3 * tri[3 * i + 2];
      const double u[3] = {P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2]};
      const double v[3] = {P2[0] - P0[0], P2[1] - P0[1], P2[2] - P0[2]};
      const double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
                   n2 = u[0] * v[1] - u[1] * v[0];
      const double J = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      for (int q = 0; q < nt; q++) {
        const double x1 = rule.x[q], x2 = rule.y[q];
        for (int d = 0; d < 3; d++) X[q][d] = P0[d] + x1 * u[d] + x2 * v[d];
        W[q] = rule.w[q] * J;
      }
    }
    for (int ib = ntp[b]; ib < ntp[b + 1]; ib++) {
      const int j = ntri[ib], kj = ntk[ib];
      double dot = 0.0;
      for (int c = 0; c < 3; c++) dot += ci[c] * curl[(3 * (size_t)j + c) * 3 + kj];
      bool touching = false;
      for (int c = touch_ptr[i]; c < touch_ptr[i + 1]; c++) touching |= touch[c] == j;
      double vij = 0.0;
      if (touching) {
        for (int k = vptr[i]; k < vptr[i + 1]; k++)
          if (vcol[k] == j) vij = vval[k];
      } else {
        const double *P0 = nodes + 3 * tri[3 * j], *P1 = nodes + 3 * tri[3 * j + 1],
                     *P2 = nodes + 3 * tri[3 * j + 2];
        const double u[3] = {P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2]};
        const double v[3] = {P2[0] - P0[0], P2[1] - P0[1], P2[2] - P0[2]};
        const double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
                     n2 = u[0] * v[1] - u[1] * v[0];
        const double J = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        for (int p = 0; p < nt; p++) {
          const double y1 = rule.x[p], y2 = rule.y[p];
          const double Y[3] = {P0[0] + y1 * u[0] + y2 * v[0], P0[1] + y1 * u[1] + y2 * v[1],
                               P0[2] + y1 * u[2] + y2 * v[2]};
          double t = 0.0;
          for (int q = 0; q < nt; q++) {
            const double dx = X[q][0] - Y[0], dy = X[q][1] - Y[1], dz = X[q][2] - Y[2];
            t += W[q] * rsq64(dx * dx + dy * dy + dz * dz);
          }
          vij += rule.w[p] * J * t;
        }
        vij *= kInv4PiSpai;
      }
      s += dot * vij;
    }
  }
  eval[e] = s;
}

__device__ __forceinline__ double spai_row_sum(double v)
{
#pragma unroll
  for (int m = 1; m < kSpaiLanes; m <<= 1) v += __shfl_xor(v, m, kSpaiLanes);
  return v;
}

// y = M x, 8 lanes per row; blockIdx.y selects the right-hand side
__global__ __launch_bounds__(256) void k_spai_apply(int n, const int *__restrict__ ptr,
                                                    const int *__restrict__ col,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ y)
{
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = gid / kSpaiLanes, sub = gid % kSpaiLanes;
  if (r >= n) return;  // uniform over the row's lanes
  x += (size_t)blockIdx.y * n;
  y += (size_t)blockIdx.y * n;
  double s = 0.0;
  for (int e = ptr[r] + sub; e < ptr[r + 1]; e += kSpaiLanes) s = fma(val[e], x[col[e]], s);
  s = spai_row_sum(s);
  if (sub == 0) y[r] = s;
}

// Solves B x = rhs in place (B m x m, row-major, overwritten) by Gaussian
// elimination with partial pivoting; false if a pivot vanishes.
bool solve_inplace(int m, double *B, double *rhs)
{
  for (int k = 0; k < m; k++) {
    int piv = k;
    for (int r = k + 1; r < m; r++)
      if (std::fabs(B[(size_t)r * m + k]) > std::fabs(B[(size_t)piv * m + k])) piv = r;
    if (B[(size_t)piv * m + k] == 0.0) return false;
    if (piv != k) {
      for (int c = 0; c < m; c++) std::swap(B[(size_t)k * m + c], B[(size_t)piv * m + c]);
      std::swap(rhs[k], rhs[piv]);
    }
    const double d = B[(size_t)k * m + k];
    for (int r = k + 1; r < m; r++) {
      const double f = B[(size_t)r * m + k] / d;
      if (f == 0.0) continue;
      for (int c = k; c < m; c++) B[(size_t)r * m + c] -= f * B[(size_t)k * m + c];
      rhs[r] -= f * rhs[k];
    }
  }
  for (int k = m - 1; k >= 0; k--) {
    double v = rhs[k];
    for (int c = k + 1; c < m; c++) v -= B[(size_t)k * m + c] * rhs[c];
    rhs[k] = v / B[(size_t)k * m + k];
  }
  return true;
}

}  // namespace

int spai_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s, bool timed)
{
  const int n = op->nrows;
  if (timed) {
    EBH_HIP(hipEventRecord(op->ev[0], s));
    EBH_HIP(hipEventRecord(op->ev[1], s));
    EBH_HIP(hipEventRecord(op->ev[2], s));
  }
  const dim3 grid((unsigned)(((size_t)n * kSpaiLanes + 255) / 256), (unsigned)nrhs);
  hipLaunchKernelGGL(k_spai_apply, grid, dim3(256), 0, s, n, op->near_ptr.p,
                     op->near_col.p, op->near_val.p, x, y);
  EBH_HIP(hipGetLastError());
  if (timed) EBH_HIP(hipEventRecord(op->ev[3], s));
  return EBH_OK;
}

}  // namespace ebh

using namespace ebh;

extern "C" int ebh_precond_spai(ebh_op *A, int rings, int max_rhs, ebh_op **out)
{
  if (!A || !out) { set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  if (rings < 1 || rings > 4) {
    set_error("patch radius must be 1..4 adjacency steps");
    return EBH_ERR_ARG;
  }
  // the single layer with equal bases, or the hypersingular operator
  // D + stab m m^T (not its Calderon-scaled form)
  const bool hs = A->d.kernel == EBH_LAPLACE_HS && A->inner && !A->hs_scale;
  if (!hs && (A->inner || A->d.kernel != EBH_LAPLACE_SL ||
              A->d.test_basis != A->d.trial_basis || A->nc != 1)) {
    set_error("the sparse approximate inverse needs a Laplace single layer with equal "
              "test and trial bases or the hypersingular operator");
    return EBH_ERR_UNSUPPORTED;
  }
  // the patch quadrature keeps a triangle's rule points in registers
  if (far_op(A)->rule.nt > 16) {
    set_error("the sparse approximate inverse takes far rules of at most 16 points "
              "(triangle rule index 7)");
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipSetDevice(A->mesh->device));
  EBH_HIP(hipStreamSynchronize(A->stream));
  const Topology &T = A->mesh->T;
  const bool p1 = A->d.test_basis == EBH_P1;
  const int n = p1 ? T.np : T.ne;

  // DOF adjacency: P1 nodes sharing a triangle, P0 triangles sharing a
  // vertex
  std::vector<std::vector<int>> adj(n);
  if (p1) {
    for (int a = 0; a < n; a++) {
      for (int p = T.node_tri_ptr[a]; p < T.node_tri_ptr[a + 1]; p++)
        for (int k = 0; k < 3; k++) {
          const int b = T.tri[3 * T.node_tri[p] + k];
          if (b != a) adj[a].push_back(b);
        }
      std::sort(adj[a].begin(), adj[a].end());
      adj[a].erase(std::unique(adj[a].begin(), adj[a].end()), adj[a].end());
    }
  } else {
    for (int i = 0; i < n; i++)
      for (int c = T.touch_ptr[i]; c < T.touch_ptr[i + 1]; c++)
        if (T.touch[c] != i) adj[i].push_back(T.touch[c]);
  }
  // host threads over rows (row c goes to thread c mod nthreads; every row
  // is independent, so results do not depend on the split)
  const int nthreads =
      (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto on_threads = [&](auto fn) {
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; t++) pool.emplace_back(fn, t);
    fn(0);
    for (auto &th : pool) th.join();
  };
  // DOFs within `depth` adjacency steps of c, sorted (breadth-first search
  // with a visit stamp per DOF and thread)
  struct Bfs {
    std::vector<int> stamp, frontier, next;
    int tick = 0;
  };
  std::vector<Bfs> bfs(nthreads);
  auto ball = [&](Bfs &w, int c, int depth, std::vector<int> &out) {
    const int mark = ++w.tick;
    out.assign(1, c);
    w.stamp[c] = mark;
    w.frontier.assign(1, c);
    for (int d = 0; d < depth; d++) {
      w.next.clear();
      for (int a : w.frontier)
        for (int b : adj[a])
          if (w.stamp[b] != mark) {
            w.stamp[b] = mark;
            w.next.push_back(b);
            out.push_back(b);
          }
      w.frontier.swap(w.next);
    }
    std::sort(out.begin(), out.end());
  };
  // patches, and the pattern of every pair that shares a patch: two DOFs
  // share a patch of radius k exactly when they are 2k steps apart at most
  std::vector<std::vector<int>> patch(n), prow(n);
  on_threads([&](int tid) {
    Bfs &w = bfs[tid];
    w.stamp.assign(n, 0);
    for (int c = tid; c < n; c += nthreads) {
      ball(w, c, rings, patch[c]);
      ball(w, c, 2 * rings, prow[c]);
    }
  });
  std::vector<int> eptr(n + 1, 0);
  for (int a = 0; a < n; a++) eptr[a + 1] = eptr[a] + (int)prow[a].size();
  const int nent = eptr[n];
  std::vector<int> ecol(nent), erow(nent);
  on_threads([&](int tid) {
    for (int a = tid; a < n; a += nthreads) {
      std::copy(prow[a].begin(), prow[a].end(), ecol.begin() + eptr[a]);
      std::fill(erow.begin() + eptr[a], erow.begin() + eptr[a + 1], a);
      std::vector<int>().swap(prow[a]);
    }
  });
  // per thread: position of a DOF in the row or patch at hand, else -1
  std::vector<std::vector<int>> posv(nthreads);

  // entries on the device: the hypersingular operator's whole, or the
  // single layer's far part
  std::vector<double> ev(nent, 0.0);
  if (hs) {
    const ebh_op *V0 = A->inner;
    DevBuf<int> drow, dcol;
    DevBuf<double> dval;
    EBH_TRY(drow.upload(erow));
    EBH_TRY(dcol.upload(ecol));
    EBH_TRY(dval.alloc(nent));
    hipLaunchKernelGGL(k_spai_hs, dim3((nent + 255) / 256), dim3(256), 0, A->stream, nent,
                       drow.p, dcol.p, V0->rule, A->mesh->nodes.p, A->mesh->tri.p,
                       A->mesh->node_tri_ptr.p, A->mesh->node_tri.p, A->mesh->node_tri_k.p,
                       A->mesh->touch_ptr.p, A->mesh->touch.p, A->curl.p, V0->near_ptr.p,
                       V0->near_col.p, V0->near_val.p, dval.p);
    EBH_HIP(hipGetLastError());
    EBH_HIP(hipStreamSynchronize(A->stream));
    EBH_HIP(gated([&] { return hipMemcpy(ev.data(), dval.p, sizeof(double) * nent, hipMemcpyDeviceToHost); }));
    if (A->d.stab != 0.0) {
      // + stab m_a m_b, m the lumped node areas
      std::vector<double> m(n);
      EBH_HIP(gated([&] { return hipMemcpy(m.data(), A->mvec.p, sizeof(double) * n, hipMemcpyDeviceToHost); }));
      for (int a = 0; a < n; a++)
        for (int k = eptr[a]; k < eptr[a + 1]; k++) ev[k] += A->d.stab * m[a] * m[ecol[k]];
    }
  } else {
    DevBuf<int> drow, dcol;
    DevBuf<double> dval;
    EBH_TRY(drow.upload(erow));
    EBH_TRY(dcol.upload(ecol));
    EBH_TRY(dval.alloc(nent));
    hipLaunchKernelGGL(k_spai_far, dim3((nent + 255) / 256), dim3(256), 0, A->stream, nent,
                       drow.p, dcol.p, p1 ? 1 : 0, A->rule, A->mesh->nodes.p,
                       A->mesh->tri.p, A->mesh->node_tri_ptr.p, A->mesh->node_tri.p,
                       A->mesh->node_tri_k.p, A->mesh->touch_ptr.p, A->mesh->touch.p,
                       dval.p);
    EBH_HIP(hipGetLastError());
    EBH_HIP(hipStreamSynchronize(A->stream));
    EBH_HIP(gated([&] { return hipMemcpy(ev.data(), dval.p, sizeof(double) * nent, hipMemcpyDeviceToHost); }));
  }
  // touching parts: V's near-field CSR
  if (!hs) {
    std::vector<int> nptr(A->nrows + 1), ncol(A->near_nnz);
    std::vector<double> nval(A->near_nnz);
    EBH_HIP(gated([&] { return hipMemcpy(nptr.data(), A->near_ptr.p, sizeof(int) * nptr.size(),
                      hipMemcpyDeviceToHost); }));
    EBH_HIP(gated([&] { return hipMemcpy(ncol.data(), A->near_col.p, sizeof(int) * ncol.size(),
                      hipMemcpyDeviceToHost); }));
    EBH_HIP(gated([&] { return hipMemcpy(nval.data(), A->near_val.p, sizeof(double) * nval.size(),
                      hipMemcpyDeviceToHost); }));
    on_threads([&](int tid) {
      std::vector<int> &pos = posv[tid];
      pos.assign(n, -1);
      for (int a = tid; a < n; a += nthreads) {
        for (int k = eptr[a]; k < eptr[a + 1]; k++) pos[ecol[k]] = k;
        for (int e = nptr[a]; e < nptr[a + 1]; e++) {
          const int k = pos[ncol[e]];
          if (k >= 0) ev[k] += nval[e];
        }
        for (int k = eptr[a]; k < eptr[a + 1]; k++) pos[ecol[k]] = -1;
      }
    });
  }

  // rows of M: patch inverses, rows split over host threads (each row is
  // independent; the result does not depend on the split)
  std::vector<int> mptr(n + 1, 0);
  for (int c = 0; c < n; c++) mptr[c + 1] = mptr[c] + (int)patch[c].size();
  std::vector<int> mcol(mptr[n]);
  std::vector<double> mval(mptr[n]);
  std::vector<int> bad(nthreads, -1);
  auto rows_of = [&](int tid) {
    std::vector<double> B, rhs;
    std::vector<int> &loc = posv[tid];
    loc.assign(n, -1);
    for (int c = tid; c < n; c += nthreads) {
      const std::vector<int> &p = patch[c];
      const int m = (int)p.size();
      // B = transpose of V restricted to the patch: B[q][r] = V[p_r][p_q]
      // from row p_r of the pattern (which holds every DOF of the patch)
      B.assign((size_t)m * m, 0.0);
      rhs.assign(m, 0.0);
      for (int r = 0; r < m; r++) loc[p[r]] = r;
      for (int r = 0; r < m; r++) {
        if (p[r] == c) rhs[r] = 1.0;
        for (int k = eptr[p[r]]; k < eptr[p[r] + 1]; k++) {
          const int q = loc[ecol[k]];
          if (q >= 0) B[(size_t)q * m + r] = ev[k];
        }
      }
      for (int r = 0; r < m; r++) loc[p[r]] = -1;
      bool ok = solve_inplace(m, B.data(), rhs.data());
      for (int r = 0; r < m && ok; r++) ok = std::isfinite(rhs[r]);
      if (!ok) {
        if (bad[tid] < 0) bad[tid] = c;
        continue;
      }
      std::copy(p.begin(), p.end(), mcol.begin() + mptr[c]);
      std::copy(rhs.begin(), rhs.end(), mval.begin() + mptr[c]);
    }
  };
  on_threads(rows_of);
  for (int t = 0; t < nthreads; t++)
    if (bad[t] >= 0) {
      set_error("singular or non-finite patch inverse at DOF " + std::to_string(bad[t]));
      return EBH_ERR_NONFINITE;
    }

  ebh_op *op = new ebh_op();
  op->mesh = A->mesh;
  op->d = A->d;
  op->d.kernel = EBH_SPARSE_INVERSE;
  op->d.max_rhs = std::max(1, max_rhs);
  op->rule = A->rule;
  op->nq = A->nq;
  op->nc = 1;
  op->nrows = op->ncols = n;
  op->near_nnz = (int64_t)mcol.size();
  int rc = EBH_OK;
  if (gated([&] { return hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking); }) !=
      hipSuccess) {
    delete op;
    set_error("hipStreamCreate failed");
    return EBH_ERR_HIP;
  }
  for (auto &e : op->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      ebh_op_destroy(op);
      set_error("hipEventCreate failed");
      return EBH_ERR_HIP;
    }
  if ((rc = op->near_ptr.upload(mptr)) || (rc = op->near_col.upload(mcol)) ||
      (rc = op->near_val.upload(mval))) {
    ebh_op_destroy(op);
    return rc;
  }
  *out = op;
  return EBH_OK;
}
