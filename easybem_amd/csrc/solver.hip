// Device-resident GMRES(m) in FP64 and mixed-precision iterative refinement
// (FP32 far-field operator inside, FP64 operator for the outer residual).
//
// The reference has no Krylov solver: its examples call scipy's CG on the
// GPU apply (python/examples/laplace3dDirichletGPU.py:64-81).  Here the
// vectors, the Krylov basis and the operator stay on the device; only the
// small Hessenberg column (<= restart+1 numbers) crosses to the host for the
// Givens update and the convergence test.
//
// All reductions are deterministic: fixed per-block partial sums reduced in
// a fixed order by one block.
#include <cmath>
#include <vector>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr int kDotBlocks = 120;
constexpr int kThreads = 256;

// partial[b*k + j] = sum over block b's rows of V[j][i] * w[i]
__global__ __launch_bounds__(kThreads) void k_dots(int n, int k,
                                                   const double *__restrict__ V,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ partial)
{
  __shared__ double red[kThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = 0; j < k; j++) {
    const double *v = V + (size_t)j * n;
    double s = 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
      s = fma(v[i], w[i], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kThreads / 64; q++) t += red[q];
      partial[(size_t)blockIdx.x * k + j] = t;
    }
    __syncthreads();
  }
}

// out[j] = sum_b partial[b*k + j] (fixed order)
__global__ void k_dots_finish(int nb, int k, const double *__restrict__ partial,
                              double *__restrict__ out)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += partial[(size_t)b * k + j];
  out[j] = s;
}

// w[i] += alpha * sum_j h[j] V[j][i]
__global__ void k_gemv_n(int n, int k, const double *__restrict__ V,
                         const double *__restrict__ h, double alpha,
                         double *__restrict__ w)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; j++) s = fma(h[j], V[(size_t)j * n + i], s);
  w[i] = fma(alpha, s, w[i]);
}

// y[i] = a*x[i] + b*y[i]
__global__ void k_axpby(int n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ void k_fill(int n, double v, double *__restrict__ y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

__global__ void k_copy(int n, const double *__restrict__ x, double *__restrict__ y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

}  // namespace
}  // namespace ebh

using namespace ebh;

struct ebh_solver {
  ebh_op *a64 = nullptr, *a32 = nullptr;
  int n = 0, m = 50;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DevBuf<double> V;         // (m+1) x n Krylov basis
  DevBuf<double> w, r, d, t;
  DevBuf<double> partial, hd;
  std::vector<double> hh;   // host staging
  int matvec64 = 0, matvec32 = 0;

  int dots(int k, const double *Vp, const double *wp, double *out)
  {
    hipLaunchKernelGGL(k_dots, dim3(kDotBlocks), dim3(kThreads), 0, stream, n, k,
                       Vp, wp, partial.p);
    hipLaunchKernelGGL(k_dots_finish, dim3((k + 63) / 64), dim3(64), 0, stream,
                       kDotBlocks, k, partial.p, hd.p);
    EBH_HIP(hipGetLastError());
    EBH_HIP(hipMemcpyAsync(out, hd.p, k * sizeof(double), hipMemcpyDeviceToHost,
                           stream));
    EBH_HIP(hipStreamSynchronize(stream));
    return EBH_OK;
  }
  int norm(const double *x, double *out)
  {
    double v;
    EBH_TRY(dots(1, x, x, &v));
    *out = std::sqrt(v);
    return EBH_OK;
  }
  int gemv_n(int k, const double *h_host, double alpha, double *wp)
  {
    EBH_HIP(hipMemcpyAsync(hd.p, h_host, k * sizeof(double), hipMemcpyHostToDevice,
                           stream));
    hipLaunchKernelGGL(k_gemv_n, dim3((n + 255) / 256), dim3(256), 0, stream, n, k,
                       V.p, hd.p, alpha, wp);
    EBH_HIP(hipGetLastError());
    // hd is reused by the next dots(): keep the copy ordered before it
    return EBH_OK;
  }
  int axpby(double a, const double *x, double b, double *y)
  {
    hipLaunchKernelGGL(k_axpby, dim3((n + 255) / 256), dim3(256), 0, stream, n, a,
                       x, b, y);
    EBH_HIP(hipGetLastError());
    return EBH_OK;
  }
  int apply(ebh_op *op, const double *x, double *y)
  {
    EBH_TRY(op_apply_one(op, x, y, stream, false));
    if (op == a64) matvec64++; else matvec32++;
    return EBH_OK;
  }

  // GMRES(m) on A z = rhs from z (in/out).  Stops when the Arnoldi residual
  // estimate drops below atol, or after maxit Arnoldi steps.
  int gmres(ebh_op *A, const double *rhs, double *z, bool z_zero,
            bool true_residual, double atol, int maxit, int *iters, double *est)
  {
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 1),
        h2(m + 1), y(m);
    int it = 0;
    double beta = 0.0;
    while (true) {
      // r = rhs - A z
      if (z_zero) {
        hipLaunchKernelGGL(k_copy, dim3((n + 255) / 256), dim3(256), 0, stream,
                           n, rhs, r.p);
        z_zero = false;
      } else {
        EBH_TRY(apply(A, z, r.p));
        EBH_TRY(axpby(1.0, rhs, -1.0, r.p));
      }
      EBH_TRY(norm(r.p, &beta));
      *est = beta;
      if (beta <= atol || it >= maxit || !(beta > 0.0)) break;
      hipLaunchKernelGGL(k_copy, dim3((n + 255) / 256), dim3(256), 0, stream, n,
                         r.p, V.p);
      EBH_TRY(axpby(0.0, r.p, 1.0 / beta, V.p));
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int j = 0;
      for (; j < m && it < maxit; j++, it++) {
        double *vj1 = V.p + (size_t)(j + 1) * n;
        EBH_TRY(apply(A, V.p + (size_t)j * n, vj1));
        // classical Gram-Schmidt, twice
        EBH_TRY(dots(j + 1, V.p, vj1, h.data()));
        EBH_TRY(gemv_n(j + 1, h.data(), -1.0, vj1));
        EBH_TRY(dots(j + 1, V.p, vj1, h2.data()));
        EBH_TRY(gemv_n(j + 1, h2.data(), -1.0, vj1));
        for (int k = 0; k <= j; k++) h[k] += h2[k];
        double hn;
        EBH_TRY(norm(vj1, &hn));
        h[j + 1] = hn;
        if (hn > 0.0) EBH_TRY(axpby(0.0, vj1, 1.0 / hn, vj1));
        // Givens rotations on the new column
        for (int k = 0; k < j; k++) {
          const double a = h[k], b = h[k + 1];
          h[k] = cs[k] * a + sn[k] * b;
          h[k + 1] = -sn[k] * a + cs[k] * b;
        }
        const double a = h[j], b = h[j + 1];
        const double rr = std::hypot(a, b);
        cs[j] = rr > 0.0 ? a / rr : 1.0;
        sn[j] = rr > 0.0 ? b / rr : 0.0;
        h[j] = rr;
        h[j + 1] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        for (int k = 0; k <= j; k++) H[(size_t)k * m + j] = h[k];
        *est = std::fabs(g[j + 1]);
        if (*est <= atol || hn == 0.0) { j++; it++; break; }
      }
      // back substitution and update z += V_j y
      for (int k = j - 1; k >= 0; k--) {
        double s = g[k];
        for (int l = k + 1; l < j; l++) s -= H[(size_t)k * m + l] * y[l];
        y[k] = s / H[(size_t)k * m + k];
      }
      EBH_TRY(gemv_n(j, y.data(), 1.0, z));
      if (*est <= atol || it >= maxit) {
        if (!true_residual) break;
        // true residual of the returned iterate
        EBH_TRY(apply(A, z, r.p));
        EBH_TRY(axpby(1.0, rhs, -1.0, r.p));
        EBH_TRY(norm(r.p, &beta));
        *est = beta;
        break;
      }
    }
    *iters = it;
    return EBH_OK;
  }
};

extern "C" {

int ebh_solver_create(ebh_op *a64, ebh_op *a32, int restart, ebh_solver **out)
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
  if (restart < 1) restart = 50;
  ebh_solver *s = new ebh_solver();
  s->a64 = a64;
  s->a32 = a32;
  s->n = n;
  s->m = restart;
  s->stream = a64->stream;
  int rc = EBH_OK;
  if ((rc = s->V.alloc((size_t)(restart + 1) * n)) || (rc = s->w.alloc(n)) ||
      (rc = s->r.alloc(n)) || (rc = s->d.alloc(n)) || (rc = s->t.alloc(n)) ||
      (rc = s->partial.alloc((size_t)kDotBlocks * (restart + 2))) ||
      (rc = s->hd.alloc(restart + 2))) {
    delete s;
    return rc;
  }
  if (hipEventCreate(&s->e0) != hipSuccess || hipEventCreate(&s->e1) != hipSuccess) {
    ebh_solver_destroy(s);
    set_error("hipEventCreate failed");
    return EBH_ERR_HIP;
  }
  *out = s;
  return EBH_OK;
}

int ebh_solver_solve(ebh_solver *s, const double *b, double *x,
                     const ebh_solve_opts *o, ebh_solve_stats *st)
{
  if (!s || !b || !x || !o) { set_error("null argument"); return EBH_ERR_ARG; }
  ebh_solve_opts opt = *o;
  if (opt.tol <= 0) opt.tol = 1e-8;
  if (opt.max_iters <= 0) opt.max_iters = 1000;
  if (opt.inner_tol <= 0) opt.inner_tol = 1e-5;
  if (opt.max_outer <= 0) opt.max_outer = 10;
  if (opt.mixed && !s->a32) {
    set_error("mixed-precision solve needs an FP32 operator");
    return EBH_ERR_ARG;
  }
  s->matvec64 = s->matvec32 = 0;
  EBH_HIP(hipEventRecord(s->e0, s->stream));
  double bn;
  EBH_TRY(s->norm(b, &bn));
  if (!std::isfinite(bn)) { set_error("non-finite right-hand side"); return EBH_ERR_NONFINITE; }
  ebh_solve_stats stats{};
  double relres = 0.0;
  if (bn == 0.0) {
    hipLaunchKernelGGL(k_fill, dim3((s->n + 255) / 256), dim3(256), 0, s->stream,
                       s->n, 0.0, x);
    stats.converged = 1;
  } else if (!opt.mixed) {
    int it = 0;
    double est = 0.0;
    EBH_TRY(s->gmres(s->a64, b, x, false, true, opt.tol * bn, opt.max_iters,
                     &it, &est));
    stats.iters = it;
    relres = est / bn;
    stats.converged = relres <= opt.tol;
  } else {
    int total = 0, outer = 0;
    while (true) {
      // FP64 residual of the current iterate
      EBH_TRY(s->apply(s->a64, x, s->w.p));
      EBH_TRY(s->axpby(1.0, b, -1.0, s->w.p));
      double rn;
      EBH_TRY(s->norm(s->w.p, &rn));
      relres = rn / bn;
      if (relres <= opt.tol || outer >= opt.max_outer || total >= opt.max_iters) break;
      // inner FP32-operator solve of A d = r from d = 0
      hipLaunchKernelGGL(k_fill, dim3((s->n + 255) / 256), dim3(256), 0,
                         s->stream, s->n, 0.0, s->d.p);
      int it = 0;
      double est = 0.0;
      // no need to solve further than the target demands
      const double itol = std::max(opt.inner_tol, 0.5 * opt.tol / relres);
      EBH_TRY(s->gmres(s->a32, s->w.p, s->d.p, true, false, itol * rn,
                       opt.max_iters - total, &it, &est));
      total += it;
      EBH_TRY(s->axpby(1.0, s->d.p, 1.0, x));
      outer++;
    }
    stats.iters = total;
    stats.outer = outer;
    stats.converged = relres <= opt.tol;
  }
  EBH_HIP(hipEventRecord(s->e1, s->stream));
  EBH_HIP(hipEventSynchronize(s->e1));
  float ms = 0.f;
  EBH_HIP(hipEventElapsedTime(&ms, s->e0, s->e1));
  stats.relres = relres;
  stats.ms = ms;
  stats.matvec64 = s->matvec64;
  stats.matvec32 = s->matvec32;
  if (st) *st = stats;
  if (!stats.converged) {
    set_error("solver stopped at relative residual " + std::to_string(relres));
    return EBH_ERR_NOCONV;
  }
  return EBH_OK;
}

void ebh_solver_destroy(ebh_solver *s)
{
  if (!s) return;
  if (s->e0) (void)hipEventDestroy(s->e0);
  if (s->e1) (void)hipEventDestroy(s->e1);
  delete s;
}

}  // extern "C"
