// Off-surface potential by the representation formula, evaluateLaplace3d
// (src/BEM3dLaplace.cpp:784-843):
//
//   u(x) = sum_j J_j/(4 pi) sum_k w_k [ t_j / |x - y_k|
//                                     - u_h(y_k) (x - y_k).n_j / |x - y_k|^3 ]
//
// t P0 (Neumann data per triangle), u_h P1 (Dirichlet data at the nodes),
// y_k = A + xi1 (B - A) + xi2 (C - A) the triangle rule's points.  An N-body
// sum with point targets: k_eval_fold writes one record per source point
// (position, single-layer strength, double-layer dipole), k_eval sums over
// LDS tiles of records with one lane per target point, FP64 partial sums
// per source chunk go to a [chunk][point] buffer and k_eval_reduce adds
// them in a fixed order.  No singular quadrature: the points must lie off
// the surface, as in the reference.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr double kInv4PiE = 0.07957747154594766788;
constexpr int kTileE = 256;  // source points per LDS tile

template <typename R>
__device__ __forceinline__ R rsqe(R x);
template <>
__device__ __forceinline__ float rsqe<float>(float x) { return __builtin_amdgcn_rsqf(x); }
// FP64 hardware estimate (relative error below 6e-8) refined by one
// Newton-Raphson step (below 5e-15)
template <>
__device__ __forceinline__ double rsqe<double>(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// Record of a source point: y (3), sigma = w J t / 4 pi, p = w J u_h(y) n /
// 4 pi (3), 0; FP32 adds the rounding remainder y - (float)y (3) and 0, so
// that x - y keeps its relative accuracy at points near the surface.
template <typename R> struct ERec {
  static constexpr int n = sizeof(R) == 4 ? 12 : 8;
};

template <typename R>
__global__ __launch_bounds__(256) void k_eval_fold(int ne, int nt, RuleArgs rule,
                                                   const double *__restrict__ nodes,
                                                   const int *__restrict__ tri,
                                                   const double *__restrict__ t,
                                                   const double *__restrict__ u,
                                                   R *__restrict__ rec)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ne * nt) return;
  const int j = idx / nt, k = idx - j * nt;
  const int a = tri[3 * j], b = tri[3 * j + 1], c = tri[3 * j + 2];
  double A[3], U[3], V[3];
  for (int d = 0; d < 3; d++) {
    A[d] = nodes[3 * a + d];
    U[d] = nodes[3 * b + d] - A[d];
    V[d] = nodes[3 * c + d] - A[d];
  }
  double n[3] = {U[1] * V[2] - U[2] * V[1], U[2] * V[0] - U[0] * V[2],
                 U[0] * V[1] - U[1] * V[0]};
  const double J = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const double e1 = rule.x[k], e2 = rule.y[k];
  const double wj = rule.w[k] * J * kInv4PiE;
  const double uh = u[a] + (u[b] - u[a]) * e1 + (u[c] - u[a]) * e2;
  R *o = rec + (size_t)idx * ERec<R>::n;
  for (int d = 0; d < 3; d++) {
    const double y = A[d] + U[d] * e1 + V[d] * e2;
    o[d] = (R)y;
    if constexpr (ERec<R>::n == 12) o[8 + d] = (R)(y - (double)o[d]);
  }
  o[3] = (R)(wj * t[j]);
  for (int d = 0; d < 3; d++) o[4 + d] = (R)(wj * uh * n[d] / J);
  o[7] = (R)0;
  if constexpr (ERec<R>::n == 12) o[11] = (R)0;
}

template <typename R>
__global__ __launch_bounds__(256) void k_eval(int npts, int nsrc, int chunks, int chunk,
                                              const double *__restrict__ pts,
                                              const R *__restrict__ rec,
                                              double *__restrict__ part)
{
  constexpr int REC = ERec<R>::n;
  __shared__ __align__(16) R tile[kTileE * REC];
  const int s = blockIdx.x % chunks;
  const int p = (blockIdx.x / chunks) * blockDim.x + threadIdx.x;
  const int pp = p < npts ? p : npts - 1;
  const double Xd = pts[3 * (size_t)pp], Yd = pts[3 * (size_t)pp + 1],
               Zd = pts[3 * (size_t)pp + 2];
  const R X = (R)Xd, Y = (R)Yd, Z = (R)Zd;
  // FP32: remainders of the target point, paired with the sources'
  const R Xl = (R)(Xd - (double)X), Yl = (R)(Yd - (double)Y), Zl = (R)(Zd - (double)Z);
  const int j0 = s * chunk, j1 = min(nsrc, j0 + chunk);
  double acc64 = 0.0;
  for (int js = j0; js < j1; js += kTileE) {
    const int n = min(kTileE, j1 - js);
    __syncthreads();
    {
      const uint4 *g = reinterpret_cast<const uint4 *>(rec + (size_t)js * REC);
      uint4 *l = reinterpret_cast<uint4 *>(tile);
      const int nv = n * REC * (int)sizeof(R) / 16;
      for (int k = threadIdx.x; k < nv; k += blockDim.x) l[k] = g[k];
    }
    __syncthreads();
    R acc = R(0);
    for (int b = 0; b < n; b++) {
      const R *r = tile + b * REC;
      R dx = X - r[0], dy = Y - r[1], dz = Z - r[2];
      if constexpr (REC == 12) {
        dx += Xl - r[8];
        dy += Yl - r[9];
        dz += Zl - r[10];
      }
      const R ri = rsqe<R>(fma(dx, dx, fma(dy, dy, dz * dz)));
      const R dp = fma(dx, r[4], fma(dy, r[5], dz * r[6]));
      acc = fma(ri, fma(-dp, ri * ri, r[3]), acc);
    }
    acc64 += (double)acc;
  }
  if (p < npts) part[(size_t)s * npts + p] = acc64;
}

__global__ void k_eval_reduce(int npts, int chunks, const double *__restrict__ part,
                              double *__restrict__ out)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  double s = 0.0;
  for (int c = 0; c < chunks; c++) s += part[(size_t)c * npts + p];
  out[p] = s;
}

// Source chunks of one evaluation: about 4096 workgroups, so that few
// points split the sources into more chunks; at most max_chunks.
void eval_chunks(int npts, int nsrc, size_t max_chunks, int *chunks_out, int *chunk_out)
{
  const int tb = (npts + 255) / 256;
  size_t want = (size_t)std::max(1, std::min((4096 + tb - 1) / tb, (nsrc + 63) / 64));
  int chunks = (int)std::max<size_t>(1, std::min(want, max_chunks));
  const int chunk = (nsrc + chunks - 1) / chunks;
  *chunks_out = (nsrc + chunk - 1) / chunk;
  *chunk_out = chunk;
}

// The three launches of one evaluation on stream s with the source records
// rec (nsrc records) and the partial sums part (chunks x npts).
template <typename R>
int eval_launch(const ebh_mesh *m, const RuleArgs &rule, int npts, const double *pts,
                const double *t, const double *u, double *out, hipStream_t s, R *rec,
                double *part, int chunks, int chunk)
{
  const int ne = m->T.ne, nt = rule.nt, nsrc = ne * nt;
  hipLaunchKernelGGL((k_eval_fold<R>), dim3((nsrc + 255) / 256), dim3(256), 0, s, ne,
                     nt, rule, m->nodes.p, m->tri.p, t, u, rec);
  EBH_HIP(hipGetLastError());
  const int tb = (npts + 255) / 256;
  hipLaunchKernelGGL((k_eval<R>), dim3(tb * chunks), dim3(256), 0, s, npts, nsrc,
                     chunks, chunk, pts, rec, part);
  EBH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_eval_reduce, dim3((npts + 255) / 256), dim3(256), 0, s, npts,
                     chunks, part, out);
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

template <typename R>
int eval_run(const ebh_mesh *m, const RuleArgs &rule, int npts, const double *pts,
             const double *t, const double *u, double *out, hipStream_t s)
{
  const int nsrc = m->T.ne * rule.nt;
  DevBuf<R> rec;
  DevBuf<double> part;
  EBH_TRY(rec.alloc((size_t)nsrc * ERec<R>::n));
  int chunks, chunk;
  eval_chunks(npts, nsrc, SIZE_MAX, &chunks, &chunk);
  EBH_TRY(part.alloc((size_t)chunks * npts));
  EBH_TRY(eval_launch<R>(m, rule, npts, pts, t, u, out, s, rec.p, part.p, chunks, chunk));
  // the scratch buffers are freed on return
  EBH_HIP(hipStreamSynchronize(s));
  return EBH_OK;
}

// Workspace of ebh_evaluate_laplace_noalloc: the source records, then the
// partial sums, in one file-scope device array that each GPU holds until
// the process ends.  Host code never uses the array's name as an address:
// its device address is looked up once per device (hipGetSymbolAddress,
// with that device current, from ebh_mesh_create) and kept in g_scratch.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// validated rule arguments of an evaluation call
int eval_args(const ebh_mesh *mesh, int tri_rule_idx, int far_precision, int npts,
              const double *points, const double *t, const double *u,
              const double *result, RuleArgs &rule)
{
  if (!mesh || !points || !t || !u || !result || npts < 1) {
    set_error("bad evaluation arguments");
    return EBH_ERR_ARG;
  }
  if (far_precision != EBH_FP64 && far_precision != EBH_FP32) {
    set_error("far_precision must be EBH_FP64 or EBH_FP32");
    return EBH_ERR_ARG;
  }
  std::vector<double> rx, ry, rw;
  if (!tri_rule(tri_rule_idx, rx, ry, rw)) {
    set_error("tri_rule index must be 0..19");
    return EBH_ERR_UNSUPPORTED;
  }
  rule = RuleArgs{};
  rule.nt = (int)rw.size();
  for (int q = 0; q < rule.nt; q++) {
    rule.x[q] = rx[q];
    rule.y[q] = ry[q];
    rule.w[q] = rw[q];
  }
  return EBH_OK;
}

template <typename R>
int eval_noalloc(const ebh_mesh *m, const RuleArgs &rule, int npts, const double *pts,
                 const double *t, const double *u, double *out, hipStream_t s,
                 unsigned char *base)
{
  const size_t nsrc = (size_t)m->T.ne * rule.nt;
  // records first; the partial sums start at the next 256-byte boundary
  const size_t rec_bytes = (nsrc * ERec<R>::n * sizeof(R) + 255) / 256 * 256;
  const size_t row = (size_t)npts * sizeof(double);
  if (rec_bytes + row > kScratchBytes) {
    set_error("evaluation needs " + std::to_string(rec_bytes + row) +
              " bytes of workspace (source records of " + std::to_string(nsrc) +
              " points and one partial sum per point); the allocation-free "
              "entry point holds " + std::to_string(kScratchBytes));
    return EBH_ERR_ARG;
  }
  int chunks, chunk;
  eval_chunks(npts, (int)nsrc, (kScratchBytes - rec_bytes) / row, &chunks, &chunk);
  return eval_launch<R>(m, rule, npts, pts, t, u, out, s, reinterpret_cast<R *>(base),
                        reinterpret_cast<double *>(base + rec_bytes), chunks, chunk);
}

}  // namespace

// Device address of the workspace array on device dev (the current device).
int eval_scratch_resolve(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index " + std::to_string(dev) + " out of range");
    return EBH_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return EBH_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("hipGetSymbolAddress of the evaluation workspace failed: ") +
              hipGetErrorString(e));
    return EBH_ERR_HIP;
  }
  g_scratch[dev] = ptr;
  return EBH_OK;
}

}  // namespace ebh

using namespace ebh;

extern "C" {

int ebh_evaluate_laplace(ebh_mesh *mesh, int tri_rule_idx, int far_precision,
                         int npts, const double *points, const double *t,
                         const double *u, double *result, void *stream)
{
  RuleArgs rule;
  EBH_TRY(eval_args(mesh, tri_rule_idx, far_precision, npts, points, t, u, result, rule));
  EBH_HIP(hipSetDevice(mesh->device));
  hipStream_t s = (hipStream_t)stream;
  return far_precision == EBH_FP32
             ? eval_run<float>(mesh, rule, npts, points, t, u, result, s)
             : eval_run<double>(mesh, rule, npts, points, t, u, result, s);
}

int ebh_evaluate_laplace_noalloc(ebh_mesh *mesh, int tri_rule_idx, int far_precision,
                                 int npts, const double *points, const double *t,
                                 const double *u, double *result, void *stream)
{
  RuleArgs rule;
  EBH_TRY(eval_args(mesh, tri_rule_idx, far_precision, npts, points, t, u, result, rule));
  EBH_HIP(hipSetDevice(mesh->device));
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (mesh->device >= 0 && mesh->device < kMaxDevices) base = g_scratch[mesh->device];
  }
  if (!base) {
    set_error("evaluation workspace not set up on this device (ebh_mesh_create sets it up)");
    return EBH_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  auto *b = static_cast<unsigned char *>(base);
  return far_precision == EBH_FP32
             ? eval_noalloc<float>(mesh, rule, npts, points, t, u, result, s, b)
             : eval_noalloc<double>(mesh, rule, npts, points, t, u, result, s, b);
}

int ebh_evaluate_laplace_host(ebh_mesh *mesh, int tri_rule_idx, int far_precision,
                              int npts, const double *points, const double *t,
                              const double *u, double *result)
{
  if (!mesh || !points || !t || !u || !result || npts < 1) {
    set_error("bad evaluation arguments");
    return EBH_ERR_ARG;
  }
  const Topology &T = mesh->T;
  for (size_t k = 0; k < 3 * (size_t)npts; k++)
    if (!std::isfinite(points[k])) { set_error("non-finite point"); return EBH_ERR_NONFINITE; }
  DevBuf<double> dp, dt, du, dr;
  EBH_TRY(dp.upload(points, 3 * (size_t)npts));
  EBH_TRY(dt.upload(t, T.ne));
  EBH_TRY(du.upload(u, T.np));
  EBH_TRY(dr.alloc(npts));
  EBH_TRY(ebh_evaluate_laplace(mesh, tri_rule_idx, far_precision, npts, dp.p, dt.p,
                               du.p, dr.p, nullptr));
  EBH_HIP(gated([&] { return hipMemcpy(result, dr.p, sizeof(double) * npts, hipMemcpyDeviceToHost); }));
  return EBH_OK;
}

}  // extern "C"
