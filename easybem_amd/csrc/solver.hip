// Device-resident GMRES(m) in FP64 for a batch of nb systems sharing one
// operator, and mixed-precision iterative refinement (FP32 far-field operator
// inside, FP64 operator for the outer residual).
//
// The reference has no Krylov solver: its examples call scipy's CG on the
// GPU apply (python/examples/laplace3dDirichletGPU.py:64-81).  Here the
// operator, the Krylov bases, the Hessenberg matrices, the Givens rotations
// and the residual estimates all stay on the device:
//   * one Arnoldi step of all nb systems = one multi-RHS operator apply,
//     classical Gram-Schmidt twice (batched dot / update kernels), a norm,
//     a Givens update and a normalisation -- a fixed kernel sequence that is
//     captured once per (operator, step j, nb) as a hipGraph and replayed;
//   * the host reads the per-system convergence flags every `check` steps
//     (one small copy), never per kernel;
//   * systems that converge go inactive (their Krylov vectors are zeroed)
//     while the others continue in lockstep.
// Reductions are deterministic: fixed per-block partial sums, reduced in a
// fixed order by one wave per output.
#include <rocblas/rocblas.h>

#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr int kDotBlocks = 64;   // partial sums per dot product
constexpr int kThreads = 256;

// partial[((r * k) + j) * kDotBlocks + b] = sum over block b's rows of
// V_j[r][i] * W[r][i];  V_j[r] = V0 + j*jstride + r*n, W[r] = W + r*n.
__global__ __launch_bounds__(kThreads) void k_bdots(int n, int k,
                                                    const double *__restrict__ V0,
                                                    size_t jstride,
                                                    const double *__restrict__ W,
                                                    const int *__restrict__ active,
                                                    double *__restrict__ partial)
{
  __shared__ double red[kThreads / 64];
  const int r = blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double *w = W + (size_t)r * n;
  const bool on = active[r] != 0;
  for (int j = 0; j < k; j++) {
    const double *v = V0 + j * jstride + (size_t)r * n;
    double s = 0.0;
    if (on)
      for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
        s = fma(v[i], w[i], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kThreads / 64; q++) t += red[q];
      partial[((size_t)r * k + j) * kDotBlocks + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// out[r*ldo + j] (= or +=) sum_b partial; one wave per (r, j).
__global__ void k_bdots_finish(int k, int nb, const double *__restrict__ partial,
                               double *__restrict__ out, int ldo, int add)
{
  const int wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wid >= k * nb) return;
  const int r = wid / k, j = wid - r * k;
  double s = lane < kDotBlocks ? partial[(size_t)wid * kDotBlocks + lane] : 0.0;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) {
    double *o = out + (size_t)r * ldo + j;
    *o = add ? *o + s : s;
  }
}

// W[r][i] += alpha * sum_j C[r*ldc + j] V_j[r][i], for j < kr (kr per system
// from kvec when given, else k).
__global__ void k_bgemv(int n, int k, const int *__restrict__ kvec,
                        const double *__restrict__ V0, size_t jstride,
                        const double *__restrict__ C, int ldc, double alpha,
                        double *__restrict__ W, const int *__restrict__ active)
{
  const int r = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !active[r]) return;
  const int kr = kvec ? kvec[r] : k;
  const double *c = C + (size_t)r * ldc;
  const double *v = V0 + (size_t)r * n + i;
  double s = 0.0;
  for (int j = 0; j < kr; j++) s = fma(c[j], v[j * jstride], s);
  W[(size_t)r * n + i] = fma(alpha, s, W[(size_t)r * n + i]);
}

// One Arnoldi step's small work for every system: h = hc[r][0..j] (after
// CGS2), hc[r][j+1] = ||w||^2.  Applies the stored rotations, builds the new
// one, updates g and the residual estimate, stores the column of R, sets the
// normalisation factor of w and retires converged systems.
__global__ void k_bgivens(int nb, int m, int j, double *__restrict__ hc,
                          double *__restrict__ Rm, double *__restrict__ cs,
                          double *__restrict__ sn, double *__restrict__ g,
                          double *__restrict__ est, const double *__restrict__ atol,
                          double *__restrict__ scale, int *__restrict__ active,
                          int *__restrict__ jdone)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  if (!active[r]) { scale[r] = 0.0; return; }
  double *h = hc + (size_t)r * (m + 2);
  const double hn = sqrt(fmax(h[j + 1], 0.0));
  h[j + 1] = hn;
  double *c = cs + (size_t)r * m, *s = sn + (size_t)r * m, *gg = g + (size_t)r * (m + 1);
  for (int k = 0; k < j; k++) {
    const double a = h[k], b = h[k + 1];
    h[k] = c[k] * a + s[k] * b;
    h[k + 1] = -s[k] * a + c[k] * b;
  }
  const double a = h[j], b = h[j + 1];
  const double rr = hypot(a, b);
  c[j] = rr > 0.0 ? a / rr : 1.0;
  s[j] = rr > 0.0 ? b / rr : 0.0;
  h[j] = rr;
  gg[j + 1] = -s[j] * gg[j];
  gg[j] = c[j] * gg[j];
  double *col = Rm + (size_t)r * (m + 1) * m;
  for (int k = 0; k <= j; k++) col[(size_t)k * m + j] = h[k];
  est[r] = fabs(gg[j + 1]);
  scale[r] = hn > 0.0 ? 1.0 / hn : 0.0;
  jdone[r] = j + 1;
  if (est[r] <= atol[r] || !(hn > 0.0)) active[r] = 0;
}

// dst[r] = (src ? src[r] : dst[r]) * scale[r]
__global__ void k_bscale(int n, const double *__restrict__ src,
                         double *__restrict__ dst, const double *__restrict__ scale)
{
  const int r = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = scale[r];
  dst[(size_t)r * n + i] = (src ? src[(size_t)r * n + i] : dst[(size_t)r * n + i]) * f;
}

// Back substitution R y = g for each system over its jdone[r] columns.
__global__ void k_bbacksolve(int nb, int m, const double *__restrict__ Rm,
                             const double *__restrict__ g, const int *__restrict__ jdone,
                             double *__restrict__ y)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  const int jn = jdone[r];
  const double *R = Rm + (size_t)r * (m + 1) * m;
  const double *gg = g + (size_t)r * (m + 1);
  double *yy = y + (size_t)r * m;
  for (int k = jn - 1; k >= 0; k--) {
    double s = gg[k];
    for (int l = k + 1; l < jn; l++) s -= R[(size_t)k * m + l] * yy[l];
    yy[k] = s / R[(size_t)k * m + k];
  }
}

// y = a*x + b*y over n entries
__global__ void k_axpby(size_t n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ void k_copy(size_t n, const double *__restrict__ x, double *__restrict__ y)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

__global__ void k_fill(size_t n, double v, double *__restrict__ y)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// Restart set-up from beta2 = ||r||^2 per system.
__global__ void k_brestart(int nb, int m, const double *__restrict__ beta2,
                           const double *__restrict__ atol, double *__restrict__ g,
                           double *__restrict__ est, double *__restrict__ scale,
                           int *__restrict__ active, int *__restrict__ jdone)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  const double b = sqrt(fmax(beta2[r], 0.0));
  double *gg = g + (size_t)r * (m + 1);
  for (int k = 0; k <= m; k++) gg[k] = 0.0;
  gg[0] = b;
  est[r] = b;
  jdone[r] = 0;
  active[r] = b > atol[r] ? 1 : 0;
  scale[r] = b > 0.0 ? 1.0 / b : 0.0;
}

// ---- block GMRES: column Gram-Schmidt inside one block of p vectors ------
// partial[(k - k0) * kDotBlocks + b] = sum over block b's rows of W_k . W_c,
// k = k0 .. k0 + gridDim.y - 1 (columns of n entries back to back).
__global__ __launch_bounds__(kThreads) void k_coldots(int n, int c, int k0,
                                                      const double *__restrict__ W,
                                                      double *__restrict__ partial)
{
  __shared__ double red[kThreads / 64];
  const int k = k0 + (int)blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double *a = W + (size_t)k * n, *b = W + (size_t)c * n;
  double s = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    s = fma(a[i], b[i], s);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; q++) t += red[q];
    partial[(size_t)blockIdx.y * kDotBlocks + blockIdx.x] = t;
  }
}

// W_c -= sum_{k < c} coef[k] W_k
__global__ void k_colupdate(int n, int c, const double *__restrict__ coef,
                            double *__restrict__ W)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < c; k++) s = fma(coef[k], W[(size_t)k * n + i], s);
  W[(size_t)c * n + i] -= s;
}

// S[k + c p] = coef[k] + coef2[k] (k < c), the CGS2 coefficients of column c
__global__ void k_colcoef(int c, int p, const double *__restrict__ coef,
                          const double *__restrict__ coef2, double *__restrict__ S)
{
  const int k = threadIdx.x;
  if (k < p) S[k + (size_t)c * p] = k < c ? coef[k] + coef2[k] : 0.0;
}

// Normalises column c: S[c + c p] = ||W_c||, W_c /= ||W_c||; a column whose
// norm fell below 1e-12 of its norm before the orthogonalisation (norm0sq)
// depends on the others and is set to zero with S[c + c p] = 0.
__global__ void k_colnormalize(int n, int c, int p, const double *__restrict__ nsq,
                               const double *__restrict__ norm0sq, double *__restrict__ S,
                               double *__restrict__ W)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double nrm = sqrt(fmax(nsq[0], 0.0));
  const bool keep = nrm > 1e-12 * sqrt(fmax(norm0sq[0], 0.0)) && nrm > 0.0;
  if (i == 0) S[c + (size_t)c * p] = keep ? nrm : 0.0;
  if (i < n) W[(size_t)c * n + i] = keep ? W[(size_t)c * n + i] / nrm : 0.0;
}

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// H = V^T W, the block Gram-Schmidt projection: V (n x K) and W (n x p),
// column-major with leading dimension n, H (K x p) column-major.  The
// reduction runs over the long dimension n, so it is split: wave (ib, c,
// jb), one 64-thread workgroup, forms the 64 x 64 block of columns 64 ib ..
// of V against columns 64 jb .. of W over rows c ch .. c ch + ch - 1, in
// FP64 MFMA 16x16x4 blocks (A = V^T: lane i + 16 k holds row k, column i of
// a 4 x 16 slice of V; B = W: lane j + 16 k holds row k, column j; D:
// column i, column j at lane j + 16 (i % 4), item i / 4), into part[c];
// k_gemm_tn_sum adds the chunks in order.  Rows and columns past the
// matrices read as 0.
typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_gemm_tn(int n, int K, int p, int ch,
                                                const double *__restrict__ V,
                                                const double *__restrict__ W,
                                                double *__restrict__ part)
{
  const int ib = blockIdx.x, c = blockIdx.y, jb = blockIdx.z, l = threadIdx.x;
  const int li = l & 15, lk = l >> 4;
  const int r0 = c * ch, r1 = min(n, r0 + ch);
  const double *va[4], *wb[4];
  bool vok[4], wok[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int ci = ib * 64 + q * 16 + li, cj = jb * 64 + q * 16 + li;
    vok[q] = ci < K;
    wok[q] = cj < p;
    va[q] = V + (size_t)(vok[q] ? ci : 0) * n;
    wb[q] = W + (size_t)(wok[q] ? cj : 0) * n;
  }
  v4d acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  // 2 steps (8 rows) a block: their operands first, then 32 MFMAs
  for (int r = r0; r < r1; r += 8) {
    double av[2][4], bv[2][4];
#pragma unroll
    for (int st = 0; st < 2; st++) {
      const int row = r + 4 * st + lk;
      const bool rok = row < r1;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        av[st][q] = rok && vok[q] ? va[q][row] : 0.0;
        bv[st][q] = rok && wok[q] ? wb[q][row] : 0.0;
      }
    }
#pragma unroll
    for (int st = 0; st < 2; st++)
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st][a], bv[st][b], acc[a][b],
                                                           0, 0, 0);
  }
  // part[c] is K x p column-major, like H
  double *o = part + (size_t)c * K * p;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = ib * 64 + a * 16 + 4 * v + lk, j = jb * 64 + b * 16 + li;
        if (i < K && j < p) o[i + (size_t)j * K] = acc[a][b][v];
      }
}

// H = sum of the chunks' parts in chunk order
__global__ void k_gemm_tn_sum(int kp, int nch, const double *__restrict__ part,
                              double *__restrict__ H)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kp) return;
  double s = 0.0;
  for (int c = 0; c < nch; c++) s += part[(size_t)c * kp + e];
  H[e] = s;
}

// chunks of the split reduction for K x p: about 2048 waves, chunks of at
// least 256 rows (a multiple of 16)
inline void gemm_tn_shape(int n, int K, int p, int *nch, int *ch)
{
  const int nib = ((K + 63) / 64) * ((p + 63) / 64);
  int c = std::max(1, std::min(2048 / nib, (n + 255) / 256));
  int h = (n + c - 1) / c;
  h = (h + 15) / 16 * 16;
  *ch = h;
  *nch = (n + h - 1) / h;
}

}  // namespace
}  // namespace ebh

using namespace ebh;

struct ebh_solver {
  ebh_op *a64 = nullptr, *a32 = nullptr;
  // right preconditioners for the FP64 and FP32 operators (may be null)
  ebh_op *p64 = nullptr, *p32 = nullptr;
  DevBuf<double> Zt, Ut;             // nbmax x n: preconditioned vectors
  // with a preconditioner, GMRES is flexible: z_j = P v_j is kept (m x
  // nbmax x n) and the update is Z += sum_j y_j z_j, so P need not be
  // linear (an FP32 P is linear only to its rounding) and the update needs
  // no further P apply
  DevBuf<double> ZV;
  int precond_applies = 0;
  int n = 0, m = 50, nbmax = 1;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DevBuf<double> V;                  // (m+1) x nbmax x n
  DevBuf<double> Rv, D, X2;          // nbmax x n
  DevBuf<double> hc, Rm, cs, sn, g, est, atol, scale, yv, beta2, partial;
  DevBuf<int> active, jdone, allon;
  double *h_est = nullptr;           // pinned host mirrors
  int *h_active = nullptr, *h_jdone = nullptr;
  std::map<std::tuple<const ebh_op *, const ebh_op *, int, int, int, int>, hipGraphExec_t>
      graphs;
  bool use_graphs = true;
  int capture_failures = 0;  // captures the runtime invalidated (then eager)
  int graph_steps = 0;       // steps of the current solve launched as graphs
  int check = 4;
  int matvec64 = 0, matvec32 = 0;
  // block GMRES (batched solves with `block` set): one block Krylov space of
  // nb columns per restart cycle, packed densely in V (column c at V + c n);
  // Hb, H2b: block Gram-Schmidt coefficients ((m+1) nb x nb); Sb: the factor
  // of the new block (nb x nb); Yb: the cycle's coefficients; cw: column
  // Gram-Schmidt scratch; blas: rocBLAS handle on `stream` with workspace
  bool block = false;
  rocblas_handle blas = nullptr;
  DevBuf<double> Hb, H2b, Sb, Yb, cw;
  DevBuf<unsigned char> blas_ws;
  // chunk parts of the projections V^T W (k_gemm_tn), sized for K = (m+1) nb
  DevBuf<double> tn_part;
  // host side of the block least-squares problem: R (the block Hessenberg
  // matrix reduced by Householder reflectors, column-major, leading
  // dimension (m+1) nb), the reflectors (one of nb+1 entries per column,
  // starting at the column's own row) and the reduced right-hand sides
  std::vector<double> hR, hV, hbeta, hG, hH, hS;

  // Captured steps hold the addresses of the operators' and the solver's
  // buffers (the kept z_j among them); they go whenever the
  // preconditioners change.  The stream must be idle.
  void drop_graphs()
  {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }

  ~ebh_solver()
  {
    if (stream) (void)hipStreamSynchronize(stream);
    if (blas) (void)rocblas_destroy_handle(blas);
    drop_graphs();
    if (h_est) (void)gated([&] { return hipHostFree(h_est); });
    if (h_active) (void)gated([&] { return hipHostFree(h_active); });
    if (h_jdone) (void)gated([&] { return hipHostFree(h_jdone); });
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }

  size_t vstride() const { return (size_t)nbmax * n; }

  int apply(ebh_op *op, const double *x, double *y, int nb)
  {
    EBH_TRY(op_apply_multi(op, x, y, nb, stream, false));
    if (op == a64) matvec64 += nb; else matvec32 += nb;
    return EBH_OK;
  }

  // dots of W[r] against k vectors V0 + j*jstride, into out[r*ldo + j]
  int bdots(int nb, int k, const double *V0, size_t jstride, const double *W,
            const int *act, double *out, int ldo, int add)
  {
    hipLaunchKernelGGL(k_bdots, dim3(kDotBlocks, nb), dim3(kThreads), 0, stream, n, k,
                       V0, jstride, W, act, partial.p);
    const int waves = k * nb;
    hipLaunchKernelGGL(k_bdots_finish, dim3((waves + 3) / 4), dim3(256), 0, stream, k,
                       nb, partial.p, out, ldo, add);
    EBH_HIP(hipGetLastError());
    return EBH_OK;
  }

  // second CGS coefficients live after the first in hc
  double *h2() { return hc.p + (size_t)nbmax * (m + 2); }

  ebh_op *precond_for(const ebh_op *A) const { return A == a64 ? p64 : p32; }

  // One Arnoldi step j for nb systems (kernels only; graph-capturable):
  // w = A z_j with z_j = P v_j (P the right preconditioner, if any, z_j kept
  // for the update), orthogonalised.
  int step_kernels(ebh_op *A, int j, int nb)
  {
    const size_t vs = vstride();
    double *vj = V.p + (size_t)j * vs, *w = V.p + (size_t)(j + 1) * vs;
    ebh_op *P = precond_for(A);
    if (P) {
      double *zj = ZV.p + (size_t)j * vs;
      EBH_TRY(op_apply_multi(P, vj, zj, nb, stream, false));
      EBH_TRY(op_apply_multi(A, zj, w, nb, stream, false));
    } else {
      EBH_TRY(op_apply_multi(A, vj, w, nb, stream, false));
    }
    const int ld = m + 2;
    const dim3 gv((n + 255) / 256, nb);
    // classical Gram-Schmidt, twice
    EBH_TRY(bdots(nb, j + 1, V.p, vs, w, active.p, hc.p, ld, 0));
    hipLaunchKernelGGL(k_bgemv, gv, dim3(256), 0, stream, n, j + 1, (const int *)nullptr,
                       V.p, vs, hc.p, ld, -1.0, w, active.p);
    EBH_TRY(bdots(nb, j + 1, V.p, vs, w, active.p, h2(), ld, 0));
    hipLaunchKernelGGL(k_bgemv, gv, dim3(256), 0, stream, n, j + 1, (const int *)nullptr,
                       V.p, vs, h2(), ld, -1.0, w, active.p);
    hipLaunchKernelGGL(k_axpby, grid1((size_t)nb * ld), dim3(256), 0, stream,
                       (size_t)nb * ld, 1.0, h2(), 1.0, hc.p);
    // ||w||^2 into hc[r][j+1]
    EBH_TRY(bdots(nb, 1, w, 0, w, active.p, hc.p + j + 1, ld, 0));
    hipLaunchKernelGGL(k_bgivens, dim3((nb + 63) / 64), dim3(64), 0, stream, nb, m, j,
                       hc.p, Rm.p, cs.p, sn.p, g.p, est.p, atol.p, scale.p, active.p,
                       jdone.p);
    hipLaunchKernelGGL(k_bscale, gv, dim3(256), 0, stream, n, (const double *)nullptr, w,
                       scale.p);
    EBH_HIP(hipGetLastError());
    return EBH_OK;
  }

  using GraphKey = std::tuple<const ebh_op *, const ebh_op *, int, int, int, int>;

  // Replays the graph recorded under key, capturing it from `kernels` the
  // first time.  The capture holds the capture gate exclusively, so no
  // library allocation on another host thread overlaps it.  A capture that
  // fails (the runtime invalidated it) runs the kernels eagerly instead,
  // and this solver launches eagerly from then on: the result is the same.
  template <typename F>
  int replay(const GraphKey &key, F kernels)
  {
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      hipGraph_t gr = nullptr;
      hipError_t e;
      int rc;
      {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
          (void)hipGetLastError();
          use_graphs = false;
          return kernels();
        }
        rc = kernels();
        e = hipStreamEndCapture(stream, &gr);
      }
      if (rc != EBH_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      hipGraphExec_t ex = nullptr;
      if (e == hipSuccess) e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      if (gr) (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        capture_failures++;
        use_graphs = false;
        return kernels();
      }
      it = graphs.emplace(key, ex).first;
    }
    EBH_HIP(hipGraphLaunch(it->second, stream));
    graph_steps++;
    return EBH_OK;
  }

  int step(ebh_op *A, int j, int nb)
  {
    // per-launch event profiling of the operator needs eager launches
    ebh_op *P = precond_for(A);
    if (P) precond_applies += nb;
    if (!use_graphs || A->prof_on) {
      EBH_TRY(step_kernels(A, j, nb));
    } else {
      auto key = std::make_tuple((const ebh_op *)A, (const ebh_op *)P, j, nb,
                                 far_op(A)->gen, P ? far_op(P)->gen : -1);
      EBH_TRY(replay(key, [&]() { return step_kernels(A, j, nb); }));
    }
    if (A == a64) matvec64 += nb; else matvec32 += nb;
    return EBH_OK;
  }

  int read_state(int nb)
  {
    EBH_HIP(hipMemcpyAsync(h_active, active.p, nb * sizeof(int), hipMemcpyDeviceToHost,
                           stream));
    EBH_HIP(hipMemcpyAsync(h_est, est.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                           stream));
    EBH_HIP(hipMemcpyAsync(h_jdone, jdone.p, nb * sizeof(int), hipMemcpyDeviceToHost,
                           stream));
    EBH_HIP(hipStreamSynchronize(stream));
    return EBH_OK;
  }

  // Rv = B - A Z (Z zero when z_zero), beta2 = ||Rv||^2 per system
  int residual(ebh_op *A, const double *B, const double *Z, bool z_zero, int nb)
  {
    const size_t tot = (size_t)nb * n;
    if (z_zero) {
      hipLaunchKernelGGL(k_copy, grid1(tot), dim3(256), 0, stream, tot, B, Rv.p);
    } else {
      EBH_TRY(apply(A, Z, Rv.p, nb));
      hipLaunchKernelGGL(k_axpby, grid1(tot), dim3(256), 0, stream, tot, 1.0, B, -1.0,
                         Rv.p);
    }
    return bdots(nb, 1, Rv.p, 0, Rv.p, allon.p, beta2.p, 1, 0);
  }

  // ---------------------------------------------------------- block GMRES
  int blas_setup(int p)
  {
    const size_t m1 = (size_t)m + 1;
    EBH_TRY(Hb.alloc(m1 * p * p));
    EBH_TRY(H2b.alloc(m1 * p * p));
    EBH_TRY(Sb.alloc((size_t)p * p));
    EBH_TRY(Yb.alloc((size_t)m * p * p));
    EBH_TRY(cw.alloc(2 * (size_t)p + 2));
    {
      size_t most = 0;
      for (int j = 0; j <= m; j++) {
        int nch, ch;
        const int K = (j + 1) * p;
        gemm_tn_shape(n, K, p, &nch, &ch);
        most = std::max(most, (size_t)nch * K * p);
      }
      EBH_TRY(tn_part.alloc(most));
    }
    if (!blas) {
      if (rocblas_create_handle(&blas) != rocblas_status_success) {
        blas = nullptr;
        set_error("rocblas_create_handle failed");
        return EBH_ERR_HIP;
      }
      // a workspace of its own, so that no call allocates (graph capture)
      EBH_TRY(blas_ws.alloc(size_t(32) << 20));
      if (rocblas_set_workspace(blas, blas_ws.p, blas_ws.bytes()) !=
              rocblas_status_success ||
          rocblas_set_pointer_mode(blas, rocblas_pointer_mode_host) !=
              rocblas_status_success) {
        set_error("rocBLAS set-up failed");
        return EBH_ERR_HIP;
      }
    }
    if (rocblas_set_stream(blas, stream) != rocblas_status_success) {
      set_error("rocblas_set_stream failed");
      return EBH_ERR_HIP;
    }
    return EBH_OK;
  }

  // H = V^T W for the first K columns of V and the p columns of W
  int gemm_tn(int K, int p, const double *W, double *H)
  {
    int nch, ch;
    gemm_tn_shape(n, K, p, &nch, &ch);
    hipLaunchKernelGGL(k_gemm_tn, dim3((K + 63) / 64, nch, (p + 63) / 64), dim3(64), 0,
                       stream, n, K, p, ch, V.p, W, tn_part.p);
    hipLaunchKernelGGL(k_gemm_tn_sum, grid1((size_t)K * p), dim3(256), 0, stream, K * p, nch,
                       tn_part.p, H);
    EBH_HIP(hipGetLastError());
    return EBH_OK;
  }

  int dgemm(rocblas_operation ta, int rows, int cols, int kk, double alpha, const double *A,
            int lda, const double *B, int ldb, double beta, double *C, int ldc)
  {
    if (rocblas_dgemm(blas, ta, rocblas_operation_none, rows, cols, kk, &alpha, A, lda, B,
                      ldb, &beta, C, ldc) != rocblas_status_success) {
      set_error("rocblas_dgemm failed");
      return EBH_ERR_HIP;
    }
    return EBH_OK;
  }

  // dots of columns k0 .. k0+cnt-1 of W with column c into out[0 .. cnt-1]
  int coldots(const double *W, int c, int k0, int cnt, double *out)
  {
    hipLaunchKernelGGL(k_coldots, dim3(kDotBlocks, cnt), dim3(kThreads), 0, stream, n, c,
                       k0, W, partial.p);
    hipLaunchKernelGGL(k_bdots_finish, dim3((cnt + 3) / 4), dim3(256), 0, stream, cnt, 1,
                       partial.p, out, 0, 0);
    EBH_HIP(hipGetLastError());
    return EBH_OK;
  }

  // W = Q S for the p columns at W: Q in place (orthonormal columns, a
  // dependent column set to zero), S (p x p, column-major, upper triangular)
  // on the device.  Classical Gram-Schmidt twice per column, deterministic
  // reductions; kernels only.
  int block_qr(double *W, int p, double *S)
  {
    double *coef = cw.p, *coef2 = cw.p + p, *n0 = cw.p + 2 * p, *nsq = n0 + 1;
    for (int c = 0; c < p; c++) {
      EBH_TRY(coldots(W, c, c, 1, n0));
      if (c > 0) {
        EBH_TRY(coldots(W, c, 0, c, coef));
        hipLaunchKernelGGL(k_colupdate, grid1(n), dim3(256), 0, stream, n, c, coef, W);
        EBH_TRY(coldots(W, c, 0, c, coef2));
        hipLaunchKernelGGL(k_colupdate, grid1(n), dim3(256), 0, stream, n, c, coef2, W);
      }
      hipLaunchKernelGGL(k_colcoef, dim3(1), dim3(std::max(p, 64)), 0, stream, c, p, coef,
                         coef2, S);
      EBH_TRY(coldots(W, c, c, 1, nsq));
      hipLaunchKernelGGL(k_colnormalize, grid1(n), dim3(256), 0, stream, n, c, p, nsq, n0,
                         S, W);
      EBH_HIP(hipGetLastError());
    }
    return EBH_OK;
  }

  // Block step j (kernels only, graph-capturable): W = A P V_j, block
  // Gram-Schmidt twice against V_0 .. V_j into Hb, then W = V_{j+1} S.
  int block_step_kernels(ebh_op *A, int j, int p)
  {
    const size_t bs = (size_t)p * n;
    double *vj = V.p + (size_t)j * bs, *w = V.p + (size_t)(j + 1) * bs;
    ebh_op *P = precond_for(A);
    if (P) {
      EBH_TRY(op_apply_multi(P, vj, Zt.p, p, stream, false));
      EBH_TRY(op_apply_multi(A, Zt.p, w, p, stream, false));
    } else {
      EBH_TRY(op_apply_multi(A, vj, w, p, stream, false));
    }
    const int K = (j + 1) * p;
    EBH_TRY(gemm_tn(K, p, w, Hb.p));
    EBH_TRY(dgemm(rocblas_operation_none, n, p, K, -1.0, V.p, n, Hb.p, K, 1.0, w, n));
    EBH_TRY(gemm_tn(K, p, w, H2b.p));
    EBH_TRY(dgemm(rocblas_operation_none, n, p, K, -1.0, V.p, n, H2b.p, K, 1.0, w, n));
    hipLaunchKernelGGL(k_axpby, grid1((size_t)K * p), dim3(256), 0, stream, (size_t)K * p,
                       1.0, H2b.p, 1.0, Hb.p);
    EBH_HIP(hipGetLastError());
    return block_qr(w, p, Sb.p);
  }

  int block_step(ebh_op *A, int j, int p)
  {
    ebh_op *P = precond_for(A);
    if (P) precond_applies += p;
    if (!use_graphs || A->prof_on) {
      EBH_TRY(block_step_kernels(A, j, p));
    } else {
      // block steps are keyed by -p in the batch slot
      auto key = std::make_tuple((const ebh_op *)A, (const ebh_op *)P, j, -p,
                                 far_op(A)->gen, P ? far_op(P)->gen : -1);
      EBH_TRY(replay(key, [&]() { return block_step_kernels(A, j, p); }));
    }
    if (A == a64) matvec64 += p; else matvec32 += p;
    return EBH_OK;
  }

  // Householder reflector of x[0..len-1] (x := alpha e_1): v in place of x
  // with v[0] = x[0] - alpha, returns beta = 2 / v.v (0 for x = 0)
  static double house(double *x, int len, double *alpha)
  {
    double ss = 0.0;
    for (int i = 1; i < len; i++) ss += x[i] * x[i];
    const double nrm = std::sqrt(x[0] * x[0] + ss);
    if (nrm == 0.0) { *alpha = 0.0; return 0.0; }
    *alpha = x[0] >= 0.0 ? -nrm : nrm;
    x[0] -= *alpha;
    return 2.0 / (x[0] * x[0] + ss);
  }
  static void apply_house(const double *v, double beta, int len, double *y)
  {
    if (beta == 0.0) return;
    double d = 0.0;
    for (int i = 0; i < len; i++) d += v[i] * y[i];
    d *= beta;
    for (int i = 0; i < len; i++) y[i] -= d * v[i];
  }

  // Host half of block step j: the new block column [Hb; Sb] of the block
  // Hessenberg matrix, reduced by the earlier reflectors and p new ones
  // (column g's reflector acts on rows g .. g + p); the same reflectors act
  // on the right-hand sides hG.  est[r] = || rows (j+1)p .. (j+2)p-1 of
  // column r of hG ||, the residual of system r after this step.
  int block_lsq(int j, int p, std::vector<double> &est)
  {
    const int K = (j + 1) * p;
    const size_t LD = (size_t)(m + 1) * p;
    hH.resize((size_t)K * p);
    hS.resize((size_t)p * p);
    EBH_HIP(hipMemcpyAsync(hH.data(), Hb.p, hH.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream));
    EBH_HIP(hipMemcpyAsync(hS.data(), Sb.p, hS.size() * sizeof(double),
                           hipMemcpyDeviceToHost, stream));
    EBH_HIP(hipStreamSynchronize(stream));
    for (int c = 0; c < p; c++) {
      double *col = &hR[(size_t)(j * p + c) * LD];
      std::fill(col, col + LD, 0.0);
      for (int i = 0; i < K; i++) col[i] = hH[(size_t)c * K + i];
      for (int i = 0; i <= c; i++) col[K + i] = hS[(size_t)c * p + i];
      for (int g = 0; g < j * p; g++)
        apply_house(&hV[(size_t)g * (p + 1)], hbeta[g], p + 1, col + g);
    }
    for (int c = 0; c < p; c++) {
      const int g = j * p + c;
      double *col = &hR[(size_t)g * LD];
      double *v = &hV[(size_t)g * (p + 1)];
      for (int i = 0; i <= p; i++) v[i] = col[g + i];
      double alpha;
      hbeta[g] = house(v, p + 1, &alpha);
      col[g] = alpha;
      for (int i = 1; i <= p; i++) col[g + i] = 0.0;
      for (int c2 = c + 1; c2 < p; c2++)
        apply_house(v, hbeta[g], p + 1, &hR[(size_t)(j * p + c2) * LD + g]);
      for (int r = 0; r < p; r++) apply_house(v, hbeta[g], p + 1, &hG[(size_t)r * LD + g]);
    }
    est.resize(p);
    for (int r = 0; r < p; r++) {
      double t = 0.0;
      for (int i = K; i < K + p; i++) t += hG[(size_t)r * LD + i] * hG[(size_t)r * LD + i];
      est[r] = std::sqrt(t);
    }
    return EBH_OK;
  }

  // Block GMRES(m) on A Z = B for the p = nb columns of B: restart cycles of
  // up to m block steps in one shared block Krylov space (each step one
  // p-column apply of A P), until every column's residual estimate meets
  // its tolerance; iters receives the block steps (the same for every
  // system).
  int block_gmres(ebh_op *A, const double *B, double *Z, bool z_zero, int p,
                  const std::vector<double> &atol_h, int maxit, std::vector<int> &iters)
  {
    EBH_TRY(blas_setup(p));
    const size_t LD = (size_t)(m + 1) * p;
    hR.assign(LD * m * p, 0.0);
    hV.assign((size_t)m * p * (p + 1), 0.0);
    hbeta.assign((size_t)m * p, 0.0);
    hG.assign(LD * p, 0.0);
    iters.assign(p, 0);
    const size_t bs = (size_t)p * n;
    int total = 0;
    std::vector<double> est(p), y;
    while (true) {
      EBH_TRY(residual(A, B, Z, z_zero, p));
      z_zero = false;
      hipLaunchKernelGGL(k_copy, grid1(bs), dim3(256), 0, stream, bs, Rv.p, V.p);
      EBH_TRY(block_qr(V.p, p, Sb.p));
      hS.resize((size_t)p * p);
      EBH_HIP(hipMemcpyAsync(hS.data(), Sb.p, hS.size() * sizeof(double),
                             hipMemcpyDeviceToHost, stream));
      EBH_HIP(hipStreamSynchronize(stream));
      std::fill(hG.begin(), hG.end(), 0.0);
      bool met = true;
      for (int r = 0; r < p; r++) {
        double t = 0.0;
        for (int i = 0; i <= r; i++) {
          hG[(size_t)r * LD + i] = hS[(size_t)r * p + i];
          t += hS[(size_t)r * p + i] * hS[(size_t)r * p + i];
        }
        met &= std::sqrt(t) <= atol_h[r];
      }
      if (met || total >= maxit) break;
      int J = 0;
      for (int j = 0; j < m && total < maxit; j++) {
        EBH_TRY(block_step(A, j, p));
        EBH_TRY(block_lsq(j, p, est));
        total++;
        J = j + 1;
        bool all = true;
        for (int r = 0; r < p; r++) all &= est[r] <= atol_h[r];
        if (all) break;
      }
      for (int r = 0; r < p; r++) iters[r] += J;
      // Y = R^-1 G over the J p columns (a zero pivot, from a dependent
      // column, gives a zero coefficient)
      const int Kc = J * p;
      y.assign((size_t)Kc * p, 0.0);
      double rmax = 0.0;
      for (int g = 0; g < Kc; g++) rmax = std::max(rmax, std::fabs(hR[(size_t)g * LD + g]));
      for (int r = 0; r < p; r++) {
        double *yy = &y[(size_t)r * Kc];
        const double *gg = &hG[(size_t)r * LD];
        for (int k = Kc - 1; k >= 0; k--) {
          double t = gg[k];
          for (int l = k + 1; l < Kc; l++) t -= hR[(size_t)l * LD + k] * yy[l];
          const double d = hR[(size_t)k * LD + k];
          yy[k] = std::fabs(d) > 1e-14 * rmax ? t / d : 0.0;
        }
      }
      EBH_HIP(hipMemcpyAsync(Yb.p, y.data(), y.size() * sizeof(double),
                             hipMemcpyHostToDevice, stream));
      // Z += P V Y
      ebh_op *P = precond_for(A);
      if (P) {
        EBH_TRY(dgemm(rocblas_operation_none, n, p, Kc, 1.0, V.p, n, Yb.p, Kc, 0.0, Ut.p, n));
        EBH_TRY(op_apply_multi(P, Ut.p, Zt.p, p, stream, false));
        precond_applies += p;
        hipLaunchKernelGGL(k_axpby, grid1(bs), dim3(256), 0, stream, bs, 1.0, Zt.p, 1.0, Z);
        EBH_HIP(hipGetLastError());
      } else {
        EBH_TRY(dgemm(rocblas_operation_none, n, p, Kc, 1.0, V.p, n, Yb.p, Kc, 1.0, Z, n));
      }
      EBH_HIP(hipStreamSynchronize(stream));
      bool all = true;
      for (int r = 0; r < p; r++) all &= est[r] <= atol_h[r];
      if (all || total >= maxit) break;
    }
    return EBH_OK;
  }

  // GMRES(m) on A Z = B for nb systems with absolute tolerances atol_h;
  // iters receives the Arnoldi steps per system.
  int gmres(ebh_op *A, const double *B, double *Z, bool z_zero, int nb,
            const std::vector<double> &atol_h, int maxit, std::vector<int> &iters)
  {
    EBH_HIP(hipMemcpyAsync(atol.p, atol_h.data(), nb * sizeof(double),
                           hipMemcpyHostToDevice, stream));
    EBH_HIP(hipStreamSynchronize(stream));
    iters.assign(nb, 0);
    // the kept z_j of flexible GMRES (allocated once, before any step is
    // captured, so captured steps keep their addresses)
    if (precond_for(A) && !ZV.p) EBH_TRY(ZV.alloc((size_t)m * vstride()));
    int total = 0;
    while (true) {
      EBH_TRY(residual(A, B, Z, z_zero, nb));
      z_zero = false;
      hipLaunchKernelGGL(k_brestart, dim3((nb + 63) / 64), dim3(64), 0, stream, nb, m,
                         beta2.p, atol.p, g.p, est.p, scale.p, active.p, jdone.p);
      hipLaunchKernelGGL(k_bscale, dim3((n + 255) / 256, nb), dim3(256), 0, stream, n,
                         Rv.p, V.p, scale.p);
      EBH_HIP(hipGetLastError());
      EBH_TRY(read_state(nb));
      bool any = false;
      for (int r = 0; r < nb; r++) any |= h_active[r] != 0;
      if (!any || total >= maxit) break;
      for (int j = 0; j < m && total < maxit; j++) {
        EBH_TRY(step(A, j, nb));
        total++;
        if ((j + 1) % check == 0 || j + 1 == m) {
          EBH_TRY(read_state(nb));
          bool on = false;
          for (int r = 0; r < nb; r++) on |= h_active[r] != 0;
          if (!on) break;
        }
      }
      EBH_TRY(read_state(nb));
      bool all_met = true;
      for (int r = 0; r < nb; r++) {
        iters[r] += h_jdone[r];
        all_met &= h_est[r] <= atol_h[r];
      }
      // Z += sum_j y_j z_j (z_j = P v_j, or v_j without a preconditioner)
      // with each system's own number of steps
      hipLaunchKernelGGL(k_bbacksolve, dim3((nb + 63) / 64), dim3(64), 0, stream, nb, m,
                         Rm.p, g.p, jdone.p, yv.p);
      hipLaunchKernelGGL(k_bgemv, dim3((n + 255) / 256, nb), dim3(256), 0, stream, n, m,
                         jdone.p, precond_for(A) ? ZV.p : V.p, vstride(), yv.p, m, 1.0, Z,
                         allon.p);
      EBH_HIP(hipGetLastError());
      // the Arnoldi estimate met every tolerance: no restart residual needed
      // (callers that need the true residual compute it themselves)
      if (all_met || total >= maxit) break;
    }
    return EBH_OK;
  }
};

namespace {

int read_norms(ebh_solver *s, int nb, std::vector<double> &out)
{
  out.resize(nb);
  EBH_HIP(hipMemcpyAsync(out.data(), s->beta2.p, nb * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream));
  EBH_HIP(hipStreamSynchronize(s->stream));
  for (double &v : out) v = std::sqrt(std::fmax(v, 0.0));
  return EBH_OK;
}

int solve_impl(ebh_solver *s, int nb, const double *b, double *x,
               const ebh_solve_opts *o, ebh_solve_stats *st)
{
  ebh_solve_opts opt = *o;
  if (opt.tol <= 0) opt.tol = 1e-8;
  if (opt.max_iters <= 0) opt.max_iters = 1000;
  if (opt.inner_tol <= 0) opt.inner_tol = 1e-5;
  if (opt.max_outer <= 0) opt.max_outer = 10;
  if (opt.mixed && !s->a32) {
    set_error("mixed-precision solve needs an FP32 operator");
    return EBH_ERR_ARG;
  }
  if (nb < 1 || nb > s->nbmax) {
    set_error("batch size must be 1.." + std::to_string(s->nbmax));
    return EBH_ERR_ARG;
  }
  s->matvec64 = s->matvec32 = s->precond_applies = s->graph_steps = 0;
  EBH_HIP(hipEventRecord(s->e0, s->stream));
  const size_t tot = (size_t)nb * s->n;
  std::vector<double> bn;
  EBH_TRY(s->bdots(nb, 1, b, 0, b, s->allon.p, s->beta2.p, 1, 0));
  EBH_TRY(read_norms(s, nb, bn));
  for (double v : bn)
    if (!std::isfinite(v)) { set_error("non-finite right-hand side"); return EBH_ERR_NONFINITE; }
  std::vector<int> iters(nb, 0), it;
  int outer = 0;
  // a zero initial guess needs no operator apply for its residual
  std::vector<double> xn;
  EBH_TRY(s->bdots(nb, 1, x, 0, x, s->allon.p, s->beta2.p, 1, 0));
  EBH_TRY(read_norms(s, nb, xn));
  bool x_zero = true;
  for (double v : xn) x_zero &= v == 0.0;
  std::vector<double> rn;   // FP64 residual norms of the returned iterates
  if (!opt.mixed) {
    std::vector<double> atol(nb);
    for (int r = 0; r < nb; r++) atol[r] = opt.tol * bn[r];
    EBH_TRY(s->block && nb > 1
                ? s->block_gmres(s->a64, b, x, x_zero, nb, atol, opt.max_iters, iters)
                : s->gmres(s->a64, b, x, x_zero, nb, atol, opt.max_iters, iters));
    // true FP64 residual of the returned iterates
    EBH_TRY(s->apply(s->a64, x, s->D.p, nb));
    hipLaunchKernelGGL(k_axpby, grid1(tot), dim3(256), 0, s->stream, tot, 1.0, b, -1.0,
                       s->D.p);
    EBH_TRY(s->bdots(nb, 1, s->D.p, 0, s->D.p, s->allon.p, s->beta2.p, 1, 0));
    EBH_TRY(read_norms(s, nb, rn));
  } else {
    int total = 0;
    while (true) {
      // FP64 residual of the current iterates into D (the loop always ends
      // right after this, so rn belongs to the returned iterates)
      if (x_zero) {
        hipLaunchKernelGGL(k_copy, grid1(tot), dim3(256), 0, s->stream, tot, b, s->D.p);
        x_zero = false;
      } else {
        EBH_TRY(s->apply(s->a64, x, s->D.p, nb));
        hipLaunchKernelGGL(k_axpby, grid1(tot), dim3(256), 0, s->stream, tot, 1.0, b,
                           -1.0, s->D.p);
      }
      EBH_TRY(s->bdots(nb, 1, s->D.p, 0, s->D.p, s->allon.p, s->beta2.p, 1, 0));
      EBH_TRY(read_norms(s, nb, rn));
      bool all = true;
      std::vector<double> atol(nb);
      for (int r = 0; r < nb; r++) {
        const double rel = bn[r] > 0 ? rn[r] / bn[r] : 0.0;
        all &= rel <= opt.tol;
        // systems already at the target get a tolerance their residual meets;
        // an inner solve need not go below a tenth of the remaining
        // reduction (inner_tol below 0.1 tol makes the first pass solve to
        // it, so that an accurate FP32 operator needs one FP64 residual)
        const double itol =
            rel <= opt.tol ? 2.0 : std::max(opt.inner_tol, 0.1 * opt.tol / rel);
        atol[r] = itol * rn[r];
      }
      if (all || outer >= opt.max_outer || total >= opt.max_iters) break;
      hipLaunchKernelGGL(k_fill, grid1(tot), dim3(256), 0, s->stream, tot, 0.0, s->X2.p);
      EBH_TRY(s->block && nb > 1
                  ? s->block_gmres(s->a32, s->D.p, s->X2.p, true, nb, atol,
                                   opt.max_iters - total, it)
                  : s->gmres(s->a32, s->D.p, s->X2.p, true, nb, atol, opt.max_iters - total,
                             it));
      int mx = 0;
      for (int r = 0; r < nb; r++) { iters[r] += it[r]; mx = std::max(mx, it[r]); }
      total += mx;
      hipLaunchKernelGGL(k_axpby, grid1(tot), dim3(256), 0, s->stream, tot, 1.0, s->X2.p,
                         1.0, x);
      outer++;
    }
  }
  EBH_HIP(hipEventRecord(s->e1, s->stream));
  EBH_HIP(hipEventSynchronize(s->e1));
  float ms = 0.f;
  EBH_HIP(hipEventElapsedTime(&ms, s->e0, s->e1));
  bool conv = true;
  double worst = 0.0;
  for (int r = 0; r < nb; r++) {
    const double rel = bn[r] > 0 ? rn[r] / bn[r] : 0.0;
    worst = std::max(worst, rel);
    conv &= rel <= opt.tol;
    if (st) {
      st[r].iters = iters[r];
      st[r].outer = outer;
      st[r].matvec64 = s->matvec64 / nb;
      st[r].matvec32 = s->matvec32 / nb;
      st[r].precond = s->precond_applies / nb;
      st[r].relres = rel;
      st[r].ms = ms;
      st[r].converged = rel <= opt.tol;
      st[r].graph_steps = s->graph_steps;
    }
  }
  if (!conv) {
    set_error("solver stopped at relative residual " + std::to_string(worst));
    return EBH_ERR_NOCONV;
  }
  return EBH_OK;
}

}  // namespace

extern "C" {

int ebh_solver_create_batched(ebh_op *a64, ebh_op *a32, int restart, int max_batch,
                              ebh_solver **out)
{
  if (!a64 || !out) { set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  const int n = a64->nrows * a64->nc;
  if (a64->ncols * a64->nc != n) {
    set_error("solver needs a square operator (test and trial spaces equal)");
    return EBH_ERR_ARG;
  }
  if (a32 && (a32->nrows != a64->nrows || a32->ncols != a64->ncols ||
              a32->nc != a64->nc)) {
    set_error("a32 and a64 differ in shape");
    return EBH_ERR_ARG;
  }
  if (max_batch < 1) max_batch = 1;
  if (max_batch > a64->d.max_rhs || (a32 && max_batch > a32->d.max_rhs)) {
    set_error("batch size exceeds the operators' max_rhs");
    return EBH_ERR_ARG;
  }
  if (restart < 1) restart = 50;
  ebh_solver *s = new ebh_solver();
  s->a64 = a64;
  s->a32 = a32;
  s->n = n;
  s->m = restart;
  s->nbmax = max_batch;
  s->stream = a64->stream;
  // convergence polling interval: every step when a step is expensive
  // (large meshes), every 4th when launch latency dominates
  s->check = a64->mesh->T.ne >= 4096 ? 1 : 4;
  const size_t nb = max_batch, m = restart;
  int rc = EBH_OK;
  if ((rc = s->V.alloc((m + 1) * nb * n)) || (rc = s->Rv.alloc(nb * n)) ||
      (rc = s->D.alloc(nb * n)) || (rc = s->X2.alloc(nb * n)) ||
      (rc = s->hc.alloc(2 * nb * (m + 2))) || (rc = s->Rm.alloc(nb * (m + 1) * m)) ||
      (rc = s->cs.alloc(nb * m)) || (rc = s->sn.alloc(nb * m)) ||
      (rc = s->g.alloc(nb * (m + 1))) || (rc = s->est.alloc(nb)) ||
      (rc = s->atol.alloc(nb)) || (rc = s->scale.alloc(nb)) ||
      (rc = s->yv.alloc(nb * m)) || (rc = s->beta2.alloc(nb)) ||
      (rc = s->partial.alloc((size_t)kDotBlocks * nb * (m + 2))) ||
      (rc = s->active.alloc(nb)) || (rc = s->jdone.alloc(nb)) ||
      (rc = s->allon.alloc(nb))) {
    delete s;
    return rc;
  }
  std::vector<int> ones(nb, 1);
  if ((rc = s->allon.upload(ones))) { delete s; return rc; }
  // defined contents for the Hessenberg scratch (entries past the current
  // step are carried along unused)
  if (hipMemsetAsync(s->hc.p, 0, s->hc.bytes(), s->stream) != hipSuccess ||
      hipMemsetAsync(s->V.p, 0, s->V.bytes(), s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    delete s;
    set_error("solver initialisation failed");
    return EBH_ERR_HIP;
  }
  if (gated([&] { return hipHostMalloc((void **)&s->h_est, nb * sizeof(double)); }) !=
          hipSuccess ||
      gated([&] { return hipHostMalloc((void **)&s->h_active, nb * sizeof(int)); }) !=
          hipSuccess ||
      gated([&] { return hipHostMalloc((void **)&s->h_jdone, nb * sizeof(int)); }) !=
          hipSuccess ||
      hipEventCreate(&s->e0) != hipSuccess || hipEventCreate(&s->e1) != hipSuccess) {
    delete s;
    set_error("solver host/event allocation failed");
    return EBH_ERR_HIP;
  }
  *out = s;
  return EBH_OK;
}

int ebh_solver_create(ebh_op *a64, ebh_op *a32, int restart, ebh_solver **out)
{
  return ebh_solver_create_batched(a64, a32, restart, 1, out);
}

int ebh_solver_set_precond(ebh_solver *s, ebh_op *p64, ebh_op *p32)
{
  if (!s) { set_error("null solver"); return EBH_ERR_ARG; }
  for (ebh_op *P : {p64, p32}) {
    if (!P) continue;
    if (P->nrows * P->nc != s->n || P->ncols * P->nc != s->n) {
      set_error("preconditioner size does not match the system");
      return EBH_ERR_ARG;
    }
    if (std::max(1, P->d.max_rhs) < s->nbmax) {
      set_error("preconditioner max_rhs is below the solver's batch size");
      return EBH_ERR_ARG;
    }
  }
  EBH_HIP(hipStreamSynchronize(s->stream));
  // steps captured with the old preconditioners (or the old z_j buffer,
  // released below) must not replay: a preconditioner set again, or a new
  // operator at a freed one's address, would match their keys
  s->drop_graphs();
  const size_t tot = (size_t)s->nbmax * s->n;
  if (p64 || p32) {
    EBH_TRY(s->Zt.alloc(tot));
    EBH_TRY(s->Ut.alloc(tot));
  } else {
    s->ZV.release();
  }
  s->p64 = p64;
  s->p32 = p32;
  return EBH_OK;
}

int ebh_solver_set_block(ebh_solver *s, int block)
{
  if (!s) { set_error("null solver"); return EBH_ERR_ARG; }
  if (block && s->nbmax < 2) {
    set_error("block GMRES needs a batched solver (max_batch >= 2)");
    return EBH_ERR_ARG;
  }
  EBH_HIP(hipStreamSynchronize(s->stream));
  s->block = block != 0;
  return EBH_OK;
}

int ebh_solver_set(ebh_solver *s, int use_graphs, int check_every)
{
  if (!s) { set_error("null solver"); return EBH_ERR_ARG; }
  s->use_graphs = use_graphs != 0;
  if (check_every > 0) s->check = check_every;
  return EBH_OK;
}

int ebh_solver_solve(ebh_solver *s, const double *b, double *x,
                     const ebh_solve_opts *opts, ebh_solve_stats *stats)
{
  if (!s || !b || !x || !opts) { set_error("null argument"); return EBH_ERR_ARG; }
  return solve_impl(s, 1, b, x, opts, stats);
}

int ebh_solver_solve_batched(ebh_solver *s, int nb, const double *b, double *x,
                             const ebh_solve_opts *opts, ebh_solve_stats *stats)
{
  if (!s || !b || !x || !opts) { set_error("null argument"); return EBH_ERR_ARG; }
  return solve_impl(s, nb, b, x, opts, stats);
}

void ebh_solver_destroy(ebh_solver *s) { delete s; }

}  // extern "C"
