// Far-field apply: an N-body sum over triangle quadrature points.
//
//   y_i = sum_{a in i} W_a phi_k(a) sum_{j not touching i} sum_{b in j}
//         G(x_a - y_b) sigma_b
//
// with W_a = w_a |J_i| and the trial density folded into the source
// strength sigma_b = w_b |J_j| (density at y_b) / (4 pi) (src/BEM3dLaplace.cpp:
// 151-192 single layer, :386-437 double layer; src/BEM3dElasticity.cpp:
// 289-348 Kelvin).  Touching pairs are excluded here and handled by the
// FP64 near-field CSR.
//
// Kernel layout (gfx950, wave64):
//   * one lane = one target triangle; its NT quadrature points live in
//     registers, so each source point read from LDS feeds NT interactions;
//   * a block walks one contiguous chunk of source triangles in LDS tiles of
//     up to 64 triangles; all lanes read the same LDS record (broadcast);
//   * touching source triangles are masked per lane with a 64-bit tile mask
//     built from the sorted touching list (a moving cursor, no inner-loop
//     memory traffic);
//   * FP32 variant: per-tile FP32 partial sums, folded into FP64
//     accumulators after every tile;
//   * partial row sums of every chunk go to a [chunk][out][target] buffer and
//     a second kernel reduces them in a fixed order (deterministic, no
//     atomics).
// Block b works on chunk b % chunks, so with chunks a multiple of 8 every
// block of one source chunk runs on the same XCD (blocks are dealt to the 8
// XCDs round-robin).
//
// The single layer with equal test and trial bases also has a pair-symmetric
// schedule (k_far_sym, below): one rsqrt per point pair for both directions,
// source triangles rotating through the lanes of a wave.
#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

#include "ebh_internal.h"

namespace ebh {
namespace {

constexpr double kInv4Pi = 0.07957747154594766788;

template <typename R> struct Eps;
// Added to r^2 so an excluded coincident point (x_a == y_b, strength 0)
// yields 0 * finite; far below the squared distance of any two distinct
// quadrature points of non-touching triangles.
template <> struct Eps<float> { static constexpr float v = 1e-20f; };
template <> struct Eps<double> { static constexpr double v = 1e-30; };

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
// FP64 reciprocal square root: the FP64 hardware estimate (relative error
// below 6e-8 on gfx950, measured over 1e-6 <= x <= 1e2) refined by one
// Newton-Raphson step y <- y + y(1 - x y^2)/2, which squares the relative
// error (below 5e-15).
__device__ __forceinline__ double rsq(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

}  // namespace

// rule sizes whose points a lane keeps in registers (k_far, k_far_multi and
// the multi-RHS kernels); other rules (16 to 79 points) run k_far_pt
bool reg_rule(int nt)
{
  return nt == 1 || nt == 3 || nt == 4 || nt == 6 || nt == 7 || nt == 12 || nt == 13;
}

// The pair-symmetric schedule's partial sums are reduced inside the
// near-field pass (k_reduce_finish) when that is requested and every node
// row block's sums fit that pass's LDS stage.
bool fused_pass(const ebh_op *op)
{
  return op->sym && op->nc == 1 && op->mesh->fin_ok && op->fused_req && op->pass_m.size() <= 2;
}

namespace {

template <typename R>
struct FarArgs {
  int ne, chunks, chunk, tile, nout;  // nout: 1 (P0 test) or 3 (P1 test)
  int ksplit;                         // pair-symmetric: step ranges per group pair
  R kappa;                            // Kelvin: 3 - 4 nu
  const R *tq;                        // [nt][3][ne] (FP32 double layer: offsets
                                      // from the triangle centroid)
  const R *tc;                        // FP32 double layer: centroid hi/lo [6][ne]
  const R *tsl;                       // single layer target frame, [rows][ne]:
                                      // median rules (nt <= 7): C, 2(P0-C),
                                      // 2(P1-C), c_q^2 |P_k-C|^2 per point;
                                      // others: C, then 2(x_q-C), |x_q-C|^2
  R rc[16];                           // median rules: c_q of point q
  R kap[4];                           // symmetric kernel: -c c' / 2 for the
                                      // median orbits (o, o'), [2 o + o'],
                                      // FP64 times 4 (doubled positions)
  const double *tn;                   // adjoint double layer: unit normals [ne][3]
  const R *src;                       // [ne*nt][REC]
  const int *touch_ptr, *touch;
  const double *tw;                   // [nt][ne]
  double *partial;                    // [chunks][nout*nc][ne]
  const int *pos;                     // pair-symmetric, P1 test, fused pass:
                                      // the sums of (t, k) go to padded
                                      // incidence position pos[k ne + t] of
                                      // a partial row, so that a node row
                                      // block's sums are contiguous
  int prow;                           // pair-symmetric: values per partial
                                      // row and right-hand side (nout ne, or
                                      // the padded length with pos)
  // Pair-symmetric schedule in passes of block rows (bounded partial sums):
  // the launch covers block rows m0 .. (its grid); target-side sums of its
  // triangles t0 .. t0+tcols-1 go to partial, row c - c0, and source-side
  // sums of triangles t0 .. t0+scols-1 to spart (sprow values per row and
  // right-hand side).  One pass: m0 = t0 = c0 = 0, tcols = scols = ne,
  // spart = partial, sprow = prow.
  int m0, t0, c0, tcols, scols, sprow;
  double *spart;
  int pf;                             // pair-symmetric single layer, P0 test
                                      // and trial: the P0-factored form
  double phi[16][3];                  // test shape functions at the points
};

// Source record lengths: single layer (x, y, z, sigma); double layer and
// Kelvin (x, y, z, s0, s1, s2, -, -); FP32 double layer keeps the source
// point as a hi + lo float pair: (hi xyz, lo xyz, p xyz, -, -, -).
template <typename R, int KIND> struct Rec {
  static constexpr int n =
      KIND == 0 || KIND == 3 ? 4 : (KIND == 1 && sizeof(R) == 4 ? 12 : 8);
};

// One workgroup's share of the far field: target triangles
// tblk*blockDim .. +blockDim against source chunk s, partial sums to
// a.partial[s].  Every thread of the workgroup calls it (it uses barriers).
template <typename R, int KIND, int NT>
__device__ __forceinline__ void far_block(const FarArgs<R> &a, int s, int tblk,
                                          R *tile)
{
  // Single layer with a median-symmetric rule (the point set is the
  // centroid C and/or orbits C + c (P_k - C), k = 0, 1, 2): with D = C - y,
  // r^2 = |D|^2 + c G_k + c^2 |P_k - C|^2, G_k = 2 D.(P_k - C), and
  // G_0 + G_1 + G_2 = 0.  Per source point: D, |D|^2, G_0, G_1, G_2; per
  // interaction: one add and one FMA (none for the centroid) before rsqrt.
  //
  // Adjoint double layer (KIND 3): the same target frame, with the kernel
  // (y - x).n_x / r^3 at target point x.  Every rule point lies in the
  // target's plane, so (y - x).n_x = -(C - y).n_x is one value per source
  // point and target triangle.
  constexpr bool SLF = KIND == 0 || KIND == 3;  // single-layer target frame
  constexpr bool RAY = SLF && NT <= 7;
  constexpr int C0 = NT % 3;  // 1 when the centroid leads the rule
  constexpr int REC = Rec<R, KIND>::n;
  constexpr int NC = KIND == 2 ? 3 : 1;
  constexpr bool F32 = sizeof(R) == 4;
  // FP32 double layer: (x - y).n is a near-cancelling difference for nearby
  // panels of a smooth surface, so x - y is formed as
  // (c_hi - y_hi) + (c_lo - y_lo) + (x - c) with c the target centroid.
  constexpr bool COMP = F32 && KIND == 1;
  const int ne = a.ne;
  const int t = tblk * blockDim.x + threadIdx.x;
  const int tt = t < ne ? t : ne - 1;
  const int j0 = s * a.chunk;
  const int j1 = min(ne, j0 + a.chunk);

  // Single layer: r^2 = |D|^2 + (2 D.d_q + |d_q|^2) with D = C - y formed
  // once per source point and d_q = x_q - C; double layer and Kelvin keep
  // the points (x_q) themselves.
  R X[NT], Y[NT], Z[NT], E[NT];
  R Cx = 0, Cy = 0, Cz = 0;
  R F0x = 0, F0y = 0, F0z = 0, F1x = 0, F1y = 0, F1z = 0;
  if constexpr (RAY) {
    Cx = a.tsl[0 * (size_t)ne + tt];
    Cy = a.tsl[1 * (size_t)ne + tt];
    Cz = a.tsl[2 * (size_t)ne + tt];
    F0x = a.tsl[3 * (size_t)ne + tt];
    F0y = a.tsl[4 * (size_t)ne + tt];
    F0z = a.tsl[5 * (size_t)ne + tt];
    F1x = a.tsl[6 * (size_t)ne + tt];
    F1y = a.tsl[7 * (size_t)ne + tt];
    F1z = a.tsl[8 * (size_t)ne + tt];
#pragma unroll
    for (int q = 0; q < NT; q++) {
      X[q] = Y[q] = Z[q] = 0;
      E[q] = a.tsl[(9 + q) * (size_t)ne + tt];
    }
  } else if constexpr (SLF) {
    Cx = a.tsl[0 * (size_t)ne + tt];
    Cy = a.tsl[1 * (size_t)ne + tt];
    Cz = a.tsl[2 * (size_t)ne + tt];
#pragma unroll
    for (int q = 0; q < NT; q++) {
      X[q] = a.tsl[(3 + 4 * q) * (size_t)ne + tt];
      Y[q] = a.tsl[(4 + 4 * q) * (size_t)ne + tt];
      Z[q] = a.tsl[(5 + 4 * q) * (size_t)ne + tt];
      E[q] = a.tsl[(6 + 4 * q) * (size_t)ne + tt];
    }
  } else {
#pragma unroll
    for (int q = 0; q < NT; q++) {
      X[q] = a.tq[(3 * q + 0) * (size_t)ne + tt];
      Y[q] = a.tq[(3 * q + 1) * (size_t)ne + tt];
      Z[q] = a.tq[(3 * q + 2) * (size_t)ne + tt];
      E[q] = 0;
    }
  }
  R Nx = 0, Ny = 0, Nz = 0;
  if constexpr (KIND == 3) {
    Nx = (R)a.tn[3 * (size_t)tt];
    Ny = (R)a.tn[3 * (size_t)tt + 1];
    Nz = (R)a.tn[3 * (size_t)tt + 2];
  }
  R Ch[3] = {0, 0, 0}, Cl[3] = {0, 0, 0};
  if constexpr (COMP) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Ch[k] = a.tc[k * (size_t)ne + tt];
      Cl[k] = a.tc[(3 + k) * (size_t)ne + tt];
    }
  }
  double acc64[NT * NC];
#pragma unroll
  for (int k = 0; k < NT * NC; k++) acc64[k] = 0.0;

  int cur = a.touch_ptr[tt];
  const int cend = a.touch_ptr[tt + 1];
  while (cur < cend && a.touch[cur] < j0) cur++;
  int nextx = cur < cend ? a.touch[cur] : INT_MAX;
  const R eps = Eps<R>::v;

  for (int js = j0; js < j1; js += a.tile) {
    const int nt = min(a.tile, j1 - js);
    __syncthreads();
    {
      const uint4 *g = reinterpret_cast<const uint4 *>(a.src + (size_t)js * NT * REC);
      uint4 *l = reinterpret_cast<uint4 *>(tile);
      const int nv = nt * NT * REC * (int)sizeof(R) / 16;
      for (int k = threadIdx.x; k < nv; k += blockDim.x) l[k] = g[k];
    }
    __syncthreads();
    uint64_t mask = 0;
    while (nextx < js + nt) {
      mask |= 1ull << (nextx - js);
      ++cur;
      nextx = cur < cend ? a.touch[cur] : INT_MAX;
    }
    R acc[NT * NC];
#pragma unroll
    for (int k = 0; k < NT * NC; k++) acc[k] = R(0);
    for (int jj = 0; jj < nt; jj++) {
      // 0 for a touching source triangle (handled by the near field), else 1
      const R keep = (R)(1u - (unsigned)((mask >> jj) & 1ull));
      const R *rec = tile + jj * NT * REC;
#pragma unroll
      for (int b = 0; b < NT; b++, rec += REC) {
        const R sx = rec[0], sy = rec[1], sz = rec[2];
        if constexpr (RAY) {
          const R Dx = Cx - sx, Dy = Cy - sy, Dz = Cz - sz;
          // single layer: sigma; adjoint double layer: sigma (y - x).n_x
          const R sg = KIND == 3 ? -rec[3] * keep * fma(Dx, Nx, fma(Dy, Ny, Dz * Nz))
                                 : rec[3] * keep;
          // an excluded (touching) triangle gets r^2 + 1 with strength 0
          const R A = fma(Dx, Dx, fma(Dy, Dy, fma(Dz, Dz, R(1) - keep)));
          R G[3];
          G[0] = fma(Dx, F0x, fma(Dy, F0y, Dz * F0z));
          G[1] = fma(Dx, F1x, fma(Dy, F1y, Dz * F1z));
          G[2] = -(G[0] + G[1]);
#pragma unroll
          for (int q = 0; q < NT; q++) {
            const R r2 = q < C0 ? A : fma(a.rc[q], G[(q - C0) % 3], A + E[q]);
            const R ri = rsq(r2);
            if constexpr (KIND == 3) acc[q] = fma(sg, ri * ri * ri, acc[q]);
            else acc[q] = fma(sg, ri, acc[q]);
          }
        } else if constexpr (SLF) {
          const R Dx = Cx - sx, Dy = Cy - sy, Dz = Cz - sz;
          const R sg = KIND == 3 ? -rec[3] * keep * fma(Dx, Nx, fma(Dy, Ny, Dz * Nz))
                                 : rec[3] * keep;
          // an excluded (touching) triangle gets |D|^2 + 1: its strength is
          // 0 and r^2 stays positive, also at a coincident point
          const R A = fma(Dx, Dx, fma(Dy, Dy, fma(Dz, Dz, R(1) - keep)));
#pragma unroll
          for (int q = 0; q < NT; q++) {
            const R B = fma(Dx, X[q], fma(Dy, Y[q], fma(Dz, Z[q], E[q])));
            const R ri = rsq(A + B);
            if constexpr (KIND == 3) acc[q] = fma(sg, ri * ri * ri, acc[q]);
            else acc[q] = fma(sg, ri, acc[q]);
          }
        } else if constexpr (COMP) {
          const R px = rec[6] * keep, py = rec[7] * keep, pz = rec[8] * keep;
          const R Dx = (Ch[0] - sx) + (Cl[0] - rec[3]);
          const R Dy = (Ch[1] - sy) + (Cl[1] - rec[4]);
          const R Dz = (Ch[2] - sz) + (Cl[2] - rec[5]);
#pragma unroll
          for (int q = 0; q < NT; q++) {
            const R dx = Dx + X[q], dy = Dy + Y[q], dz = Dz + Z[q];
            const R r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
            const R ri = rsq(r2);
            const R dp = fma(dx, px, fma(dy, py, dz * pz));
            acc[q] = fma(dp, ri * ri * ri, acc[q]);
          }
        } else if constexpr (KIND == 1) {
          const R px = rec[3] * keep, py = rec[4] * keep, pz = rec[5] * keep;
#pragma unroll
          for (int q = 0; q < NT; q++) {
            const R dx = X[q] - sx, dy = Y[q] - sy, dz = Z[q] - sz;
            const R r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
            const R ri = rsq(r2);
            const R dp = fma(dx, px, fma(dy, py, dz * pz));
            acc[q] = fma(dp, ri * ri * ri, acc[q]);
          }
        } else {
          const R gx = rec[3] * keep, gy = rec[4] * keep, gz = rec[5] * keep;
          const R kx = a.kappa * gx, ky = a.kappa * gy, kz = a.kappa * gz;
#pragma unroll
          for (int q = 0; q < NT; q++) {
            const R dx = X[q] - sx, dy = Y[q] - sy, dz = Z[q] - sz;
            const R r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
            const R ri = rsq(r2);
            const R ds = fma(dx, gx, fma(dy, gy, dz * gz)) * (ri * ri);
            acc[3 * q + 0] = fma(ri, fma(dx, ds, kx), acc[3 * q + 0]);
            acc[3 * q + 1] = fma(ri, fma(dy, ds, ky), acc[3 * q + 1]);
            acc[3 * q + 2] = fma(ri, fma(dz, ds, kz), acc[3 * q + 2]);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NT * NC; k++) acc64[k] += (double)acc[k];
  }
  if (t < ne) {
    const size_t stride = (size_t)ne;
    double *out = a.partial + (size_t)s * a.nout * NC * stride + t;
    if (a.nout == 1) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < NT; q++) v = fma(a.tw[q * stride + t], acc64[q * NC + c], v);
        out[c * stride] = v;
      }
    } else {
      double wq[NT];
#pragma unroll
      for (int q = 0; q < NT; q++) wq[q] = a.tw[q * stride + t];
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
          double v = 0.0;
#pragma unroll
          for (int q = 0; q < NT; q++)
            v = fma(wq[q] * a.phi[q][k], acc64[q * NC + c], v);
          out[(k * NC + c) * stride] = v;
        }
    }
  }
  (void)F32;
}

template <typename R, int KIND, int NT>
__global__ __launch_bounds__(256) void k_far(FarArgs<R> a)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  far_block<R, KIND, NT>(a, blockIdx.x % a.chunks, blockIdx.x / a.chunks,
                         reinterpret_cast<R *>(smem_raw));
}

// Rules whose points do not fit a lane's registers (triangle rule indices
// 7..19, 16 to 79 points; the reference's `order` 7 and above): one lane per
// target quadrature point, a workgroup of 256 lanes holds TB = 256 / nt whole
// target triangles, and source triangles pass through LDS in tiles of TS
// triangles (all lanes read the same record: a broadcast).  Per interaction
// the point forms of far_block: x - y, r^2, rsqrt, the kernel's factor; an
// excluded (touching) source triangle has strength 0 and r^2 + eps.  The NT
// lanes of a triangle then reduce their weighted point sums through LDS in a
// fixed order into partial[chunk][out][triangle], as far_block writes them.
template <typename R>
struct FarPtArgs {
  int ne, nt, chunks, chunk, ts, nout;  // ts: source triangles per LDS tile
  R kappa;                              // Kelvin: 3 - 4 nu
  const R *tq;                          // [nt][3][ne] (FP32 double layer:
                                        // offsets from the centroid)
  const R *tc;                          // FP32 double layer: centroid hi/lo [6][ne]
  const double *tn;                     // adjoint double layer: normals [ne][3]
  const R *src;                         // [ne*nt][REC]
  const int *touch_ptr, *touch;
  const double *tw;                     // [nt][ne]
  const double *phi;                    // [nt][3] test shape functions
  double *partial;                      // [chunks][nout*nc][ne]
};

template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_far_pt(FarPtArgs<R> a)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int REC = Rec<R, KIND>::n;
  constexpr int NC = KIND == 2 ? 3 : 1;
  constexpr bool COMP = sizeof(R) == 4 && KIND == 1;
  R *tile = reinterpret_cast<R *>(smem_raw);
  double *red = reinterpret_cast<double *>(smem_raw);  // reused after the tiles
  const int ne = a.ne, nt = a.nt, TB = 256 / nt;
  const int s = blockIdx.x % a.chunks, tb = blockIdx.x / a.chunks;
  const int lt = threadIdx.x / nt, q = threadIdx.x - lt * nt;
  const bool lane_ok = lt < TB;
  const int t = tb * TB + lt;
  const int tt = (lane_ok && t < ne) ? t : ne - 1;
  const int qq = lane_ok ? q : 0;
  const int j0 = s * a.chunk, j1 = min(ne, j0 + a.chunk);
  const R X = a.tq[(3 * qq + 0) * (size_t)ne + tt];
  const R Y = a.tq[(3 * qq + 1) * (size_t)ne + tt];
  const R Z = a.tq[(3 * qq + 2) * (size_t)ne + tt];
  R Ch[3] = {0, 0, 0}, Cl[3] = {0, 0, 0};
  if constexpr (COMP) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Ch[k] = a.tc[k * (size_t)ne + tt];
      Cl[k] = a.tc[(3 + k) * (size_t)ne + tt];
    }
  }
  R Nx = 0, Ny = 0, Nz = 0;
  if constexpr (KIND == 3) {
    Nx = (R)a.tn[3 * (size_t)tt];
    Ny = (R)a.tn[3 * (size_t)tt + 1];
    Nz = (R)a.tn[3 * (size_t)tt + 2];
  }
  double acc64[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) acc64[c] = 0.0;
  int cur = a.touch_ptr[tt];
  const int cend = a.touch_ptr[tt + 1];
  while (cur < cend && a.touch[cur] < j0) cur++;
  int nextx = cur < cend ? a.touch[cur] : INT_MAX;
  const R eps = Eps<R>::v;
  for (int js = j0; js < j1; js += a.ts) {
    const int nts = min(a.ts, j1 - js);
    __syncthreads();
    {
      const uint4 *g = reinterpret_cast<const uint4 *>(a.src + (size_t)js * nt * REC);
      uint4 *l = reinterpret_cast<uint4 *>(tile);
      const int nv = nts * nt * REC * (int)sizeof(R) / 16;
      for (int k = threadIdx.x; k < nv; k += blockDim.x) l[k] = g[k];
    }
    __syncthreads();
    uint64_t mask = 0;
    while (nextx < js + nts) {
      mask |= 1ull << (nextx - js);
      ++cur;
      nextx = cur < cend ? a.touch[cur] : INT_MAX;
    }
    R acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = R(0);
    for (int jj = 0; jj < nts; jj++) {
      const R keep = (R)(1u - (unsigned)((mask >> jj) & 1ull));
      const R *rec = tile + jj * nt * REC;
      for (int b = 0; b < nt; b++, rec += REC) {
        R dx, dy, dz;
        if constexpr (COMP) {
          dx = ((Ch[0] - rec[0]) + (Cl[0] - rec[3])) + X;
          dy = ((Ch[1] - rec[1]) + (Cl[1] - rec[4])) + Y;
          dz = ((Ch[2] - rec[2]) + (Cl[2] - rec[5])) + Z;
        } else {
          dx = X - rec[0];
          dy = Y - rec[1];
          dz = Z - rec[2];
        }
        const R ri = rsq(fma(dx, dx, fma(dy, dy, fma(dz, dz, eps))));
        if constexpr (KIND == 0) {
          acc[0] = fma(rec[3] * keep, ri, acc[0]);
        } else if constexpr (KIND == 3) {
          // (y - x).n_x sigma / r^3
          const R dn = fma(dx, Nx, fma(dy, Ny, dz * Nz));
          acc[0] = fma(-rec[3] * keep * dn, ri * ri * ri, acc[0]);
        } else if constexpr (KIND == 1) {
          const R *p = rec + (COMP ? 6 : 3);
          const R dp = fma(dx, p[0], fma(dy, p[1], dz * p[2])) * keep;
          acc[0] = fma(dp, ri * ri * ri, acc[0]);
        } else {
          const R gx = rec[3] * keep, gy = rec[4] * keep, gz = rec[5] * keep;
          const R ds = fma(dx, gx, fma(dy, gy, dz * gz)) * (ri * ri);
          acc[0] = fma(ri, fma(dx, ds, a.kappa * gx), acc[0]);
          acc[1] = fma(ri, fma(dy, ds, a.kappa * gy), acc[1]);
          acc[2] = fma(ri, fma(dz, ds, a.kappa * gz), acc[2]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) acc64[c] += (double)acc[c];
  }
  // weighted point sums of each triangle, one output at a time
  const double wq = lane_ok ? a.tw[(size_t)qq * ne + tt] : 0.0;
  for (int k = 0; k < a.nout; k++)
    for (int c = 0; c < NC; c++) {
      __syncthreads();
      const double f = a.nout == 1 ? 1.0 : a.phi[3 * qq + k];
      red[threadIdx.x] = lane_ok ? wq * f * acc64[c] : 0.0;
      __syncthreads();
      if ((int)threadIdx.x < TB) {
        const int t2 = tb * TB + (int)threadIdx.x;
        if (t2 < ne) {
          double v = 0.0;
          for (int q2 = 0; q2 < nt; q2++) v += red[threadIdx.x * nt + q2];
          a.partial[((size_t)s * a.nout * NC + k * NC + c) * ne + t2] = v;
        }
      }
    }
}

// Persistent multi-mesh far field: one launch for several operators of one
// kind.  Work items (operator k, target block of 256 triangles, source
// chunk s) are split into 8 queues by s mod 8; workgroup w starts on queue
// w mod 8 -- the XCD it was dispatched to, so the workgroups sharing a
// chunk's source records share an L2 -- and moves on to the other queues
// when its own is drained.  Each workgroup takes items from an atomic
// counter until every queue is empty; there is no waiting between
// workgroups.  off[x][k] holds the first item of operator k in queue x.
template <typename R, int KIND, int NT>
__global__ __launch_bounds__(256) void k_far_multi(const FarArgs<R> *__restrict__ descs,
                                                   const int *__restrict__ off,
                                                   int nops, int *__restrict__ counters)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R *tile = reinterpret_cast<R *>(smem_raw);
  __shared__ int sh_item;
  const int home = blockIdx.x & 7;
  for (int c = 0; c < 8; c++) {
    const int x = (home + c) & 7;
    const int *o = off + x * (nops + 1);
    const int total = o[nops];
    while (true) {
      __syncthreads();  // sh_item and the LDS tile of the last item are free
      if (threadIdx.x == 0) sh_item = atomicAdd(&counters[x], 1);
      __syncthreads();
      const int item = sh_item;
      if (item >= total) break;
      int k = 0;
      while (item >= o[k + 1]) k++;
      const FarArgs<R> &a = descs[k];
      const int local = item - o[k];
      const int cx = (a.chunks - x + 7) / 8;  // chunks s = x + 8 j of this queue
      far_block<R, KIND, NT>(a, x + 8 * (local % cx), local / cx, tile);
    }
  }
}

// Pair-symmetric single-layer far field, for V with equal test and trial
// bases and a median rule.  G(x - y) = G(y - x), so one rsqrt serves the
// point pair (a of triangle i, b of triangle j) in both directions: the lane
// of target triangle i adds sigma_b / r to its point a, and sigma_a / r to
// an accumulator of source point b, sigma_a being i's own folded strength.
//
// Triangles form groups of 64 (one wave).  A wave takes the group pair
// (I, J), I < J, in 64 steps: at step k lane L pairs its target triangle
// 64 I + L with source triangle 64 J + (L + k) mod 64, read from LDS, and
// after every step the source accumulators move down one lane (ds_bpermute
// from lane L + 1), so after 64 steps lane L holds the sums of source
// triangle 64 J + L.  Pairs inside one group (I == J) take the target side
// only.  Contributions to the targets of group I from the sources of group
// J go to partial slot J (target side), those to the targets of group J
// from group I to slot I (source side): each slot of a target is written
// exactly once, and k_reduce_wide adds the ng slots of a target in a fixed
// order.  A workgroup is 4 waves, target groups
// 4M .. 4M+3, with source group J >= 4M staged in LDS; a wave with I > J has
// no work (its pair is the source side of the (J, I) wave).  For small
// meshes the 64 steps split into ksplit ranges, one workgroup each, with
// ksplit partial slots per group (slot J ksplit + range); a launch with NR
// right-hand sides keeps NR of them per slot, so the buffer is
// [ng ksplit][NR][nout][ne].
//
// With a median rule every point of a triangle is its centroid or C + c_q
// (P_k - C), k = 0, 1, 2, so for target point q (direction k) and source
// point b (direction l) of triangles with centroids C, C' and D = C - C':
//   r^2 = |D + c_q m_k - c_b m'_l|^2
//       = A_b + B_q + kappa_qb W_kl,
//   A_b = |D|^2 - c_b D.F'_l + E'_b,  B_q = c_q D.F_k + E_q,
//   W_kl = F_k.F'_l,  kappa_qb = -c_q c_b / 2,
// with m = P - C, F = 2m and E = c^2 |m|^2 (terms of a centroid point are 0).
// Since m_0 + m_1 + m_2 = 0, the sums for k = 2 or l = 2 follow from those
// for 0 and 1: per triangle pair the kernel forms D, |D|^2, 2 + 2 + 4 dot
// products and a few sums; per point pair one add and one FMA.

// value of lane L + 1 (mod 64) in lane L: DPP wave_rol:1 (control 0x134),
// every lane of the wave active (the rotation reads no out-of-wave lane, so
// the "old" operand is never used; passing the value itself avoids a copy)
__device__ __forceinline__ int lane_rol1(int v)
{
  return __builtin_amdgcn_update_dpp(v, v, 0x134, 0xF, 0xF, false);
}
__device__ __forceinline__ float lane_rol1(float v)
{
  return __int_as_float(lane_rol1(__float_as_int(v)));
}
__device__ __forceinline__ double lane_rol1(double v)
{
  const long long u = __double_as_longlong(v);
  const int lo = lane_rol1((int)u), hi = lane_rol1((int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Added to r^2 of an excluded (touching) pair in the symmetric kernel.
template <typename R> struct Excl;
template <> struct Excl<float> { static constexpr float v = 0x1p100f; };
template <> struct Excl<double> { static constexpr double v = 1e200; };

// source record of the one-point symmetric kernel: x, y, z and NR strengths
// (NR = 3: the Kelvin fold's layout)
template <typename R, int NR>
struct alignas(4 * sizeof(R)) SymRec {
  R v[NR == 1 ? 4 : 8];
};

// LDS layout of the symmetric kernel's source group (64 triangles).
// One-point single-layer rules: one SymRec per triangle.  Otherwise one
// record per triangle of GW values, the source triangle's frame in the
// target frame's form (centroid C', F'_0 = 2 (P'_0 - C'), F'_1 = 2 (P'_1 -
// C'), E'_b = c_b^2 |P'_l - C'|^2 of its NT - C0 median points), for the
// double layer its unit normal n', then the NT x NR strengths.  GW is a
// multiple of 4 with GW / 4 odd, so the 16-byte reads of 16 lanes at
// rotating records spread over the LDS banks.
// HL (FP32): the record also holds the rounding remainder of the source
// centroid, C' - (float)C', at offset LO.  PF (the P0-factored single
// layer): in place of the NT x NR strengths, the NT point weights w_b J'
// and then the NR densities (times 1/4 pi) of the triangle.
template <int NT, int NR, bool DL = false, bool HL = false, bool PF = false>
struct SymLayout {
  static constexpr bool ROT = NT == 1 && !DL;
  static constexpr int C0 = NT % 3;
  static constexpr int NO = NT - C0;
  static constexpr int NRM = 9 + NO;              // offset of n' (double layer)
  static constexpr int SIG = NRM + (DL ? 3 : 0);  // offset of the strengths
  static constexpr int LO = SIG + (PF ? NT + NR : NT * NR);  // centroid remainder (HL)
  static constexpr int USED = ROT ? (NR == 1 ? 4 : 8) : LO + (HL ? 3 : 0);
  static constexpr int Q4 = (USED + 3) / 4;
  static constexpr int GW = ROT ? USED : 4 * (Q4 % 2 == 1 ? Q4 : Q4 + 1);
  static constexpr int LDS = 64 * GW;
};
// the layout of sym_block<R, NT, NR, DK> (DK 1, 2: double layers with the
// normal; FP32 records read from LDS by index, every kernel but the
// one-point single layers: the centroid remainder; DK 4: the P0-factored
// single layer)
template <typename R, int NT, int NR, int DK>
using SymLay = SymLayout<NT, NR, DK == 1 || DK == 2,
                         sizeof(R) == 4 && (NT != 1 || DK == 1 || DK == 2), DK == 4>;

// Workgroup bid of the symmetric schedule (see above); lds holds the source
// group in the SymLayout form.  Every thread of the workgroup calls it (one
// barrier).
//
// DK = 1: the double layer (x - y).n_y / r^3 with the same point sets.  Rule
// points lie in their triangle's plane, so for target point q of triangle i
// and source point b of triangle j, tf_q = (x_q - y_b).n_j = D.n_j + c_q
// m_k.n_j does not depend on b, and sf_b = (y_b - x_q).n_i = -D.n_i + c_b
// m'_l.n_i not on q: each direction is one FMA per point pair (strength
// times r^-3) plus one per point and step for that factor.  DK = 2: the
// adjoint double layer (y - x).n_x / r^3, whose factors are the same two
// with the roles swapped: target point q takes sigma_b sf_b / r^3, source
// point b takes sigma_q tf_q / r^3, so the strengths are scaled once per
// step.  The strengths are the single layer's (w J u_h / 4 pi per point).
//
// QS = 2 (the FP64 Kelvin single layer with 6 or 7 points, whose target and
// source sums would not fit in 256 registers): a workgroup takes the target
// points Q0 .. Q0 + NQ - 1 of its triangles (QH = 0: the first half, 1: the
// rest) against every source point; the two halves are two step-range slots
// of the partial sums (slot index rng = step range * QS + QH), which the
// reduction adds like any other.
template <typename R, int NT, int NR, int DK = 0, int QS = 1, int QH = 0>
__device__ __forceinline__ void sym_block(const FarArgs<R> &a, int bid,
                                          R *__restrict__ lds)
{
  constexpr bool DL = DK == 1 || DK == 2, ADJ = DK == 2, KEL = DK == 3, PF = DK == 4;
  static_assert(!DL || NR == 1, "double layer: one right-hand side");
  static_assert(!KEL || NR == 3, "Kelvin: three components");
  static_assert(!PF || NT > 1, "P0-factored form: rules of several points");
  static_assert(QS == 1 || (KEL && NT > 1), "target-point halves: Kelvin only");
  // target points of this workgroup: Q0 .. Q0 + NQ - 1 (those below NT)
  constexpr int NQ = (NT + QS - 1) / QS, Q0 = QH * NQ;
  // point sums per side: NQ target and NT source points, or one per triangle
  // in the P0-factored form
  constexpr int NA = PF ? 1 : NQ, NB = PF ? 1 : NT;
  constexpr int RW = NR == 1 ? 4 : 8;
  using RecT = SymRec<R, NR>;
  using Lay = SymLay<R, NT, NR, DK>;
  constexpr int C0 = Lay::C0, NO = Lay::NO, GW = Lay::GW;
  // FP32 (records read from LDS): the centroid difference D = C - C' is
  // formed from hi and lo parts, (C_hi - C'_hi) + (C_lo - C'_lo), so that it
  // carries no rounding of the positions (which would be systematic for each
  // triangle pair; for the double layers D.n is a near-cancelling difference
  // between nearby panels), and with one right-hand side every step's point
  // sums (FP32 over the step's source points) go to FP64 accumulators.
  // Measured at L5 P1 single layer: FP32-vs-FP64 apply error 6e-8 -> 8e-9
  // (random density) and 7e-9 (the solution of the bench's solve); flushing
  // every 1, 4 or 8 steps costs the same time, and an FP32 Newton step on
  // rsqrt costs 60% and helps nothing.
  constexpr bool HL = sizeof(R) == 4 && !Lay::ROT;
  constexpr bool FL = HL && NR == 1;
  constexpr int KF = 1;
  // One-point single-layer rules move the source records between lanes (DPP
  // wave_rol:1, lane L reads lane L + 1 -- VALU moves instead of LDS reads,
  // which would otherwise bound a step of so little arithmetic); otherwise
  // the source triangle's record is read from LDS by index.  The
  // accumulators move with DPP for one-point rules and with ds_bpermute
  // otherwise.
  constexpr bool ROT = Lay::ROT;
  const int ks = a.ksplit, ng = a.chunks / ks, ne = a.ne;
  int rem = bid / ks, M = a.m0;
  const int rng = bid - rem * ks;
  while (rem >= ng - 4 * M) { rem -= ng - 4 * M; M++; }
  const int J = 4 * M + rem;
  // step range rng / QS of ks / QS (slot rng: the range and the target half)
  const int k0 = rng / QS * 64 / (ks / QS), k1 = (rng / QS + 1) * 64 / (ks / QS);
  // FP64 works with doubled positions, so the distances it forms are 2r and
  // rsqrt gives 1/(2r); a Newton step from that estimate needs one multiply
  // fewer (see below).  FP32 uses the positions as they are.  Centroids and
  // the radii c_q scale by 2, E and kappa by 4, F and W not at all.
  constexpr bool F64 = sizeof(R) == 8;
  const R ps = F64 ? R(2) : R(1);
  if constexpr (ROT) {
    // source group J: positions and strengths of its 64 points; points past
    // the mesh get strength 0, far away
    const RecT *g = reinterpret_cast<const RecT *>(a.src);
    RecT *l = reinterpret_cast<RecT *>(lds);
    const int p0 = J * 64, pend = ne;
    for (int p = threadIdx.x; p < 64; p += blockDim.x) {
      RecT r;
      if (p0 + p < pend) {
        r = g[p0 + p];
        if constexpr (F64) {
          r.v[0] *= ps;
          r.v[1] *= ps;
          r.v[2] *= ps;
        }
      } else {
#pragma unroll
        for (int k = 0; k < RW; k++) r.v[k] = k < 3 ? R(1e18) : R(0);
      }
      l[p] = r;
    }
  } else {
    // source group J: frames and strengths of its 64 triangles, one field of
    // 64 triangles per pass (coalesced rows of the frame); triangles past
    // the mesh get strength 0 and a centroid far away
    const int j0 = J * 64;
    for (int i = threadIdx.x; i < 64 * Lay::USED; i += blockDim.x) {
      const int tri = i & 63, f = i >> 6, j = j0 + tri;
      R v;
      if (j < ne) {
        if (f < 3) {
          v = ps * a.tsl[f * (size_t)ne + j];
        } else if (f < 9) {
          v = a.tsl[f * (size_t)ne + j];
        } else if (f < Lay::NRM) {
          v = ps * ps * a.tsl[(f + C0) * (size_t)ne + j];
        } else if (f < Lay::SIG) {
          v = (R)a.tn[3 * (size_t)j + (f - Lay::NRM)];
        } else if (f < Lay::LO) {
          const int sr = f - Lay::SIG;
          if constexpr (PF) {
            // point weights w_b J', then the densities: point 0's strength
            // over its weight (the fold's strengths are w_b J' sigma / 4 pi)
            if (sr < NT) {
              v = (R)a.tw[(size_t)sr * ne + j];
            } else {
              const double w0 = a.tw[j];
              v = w0 != 0.0 ? (R)((double)a.src[RW * (size_t)j * NT + 3 + (sr - NT)] / w0)
                            : R(0);
            }
          } else {
            const int b = sr / NR, r = sr - b * NR;
            v = a.src[RW * ((size_t)j * NT + b) + 3 + r];
          }
        } else {
          v = a.tsl[(9 + NT + f - Lay::LO) * (size_t)ne + j];
        }
      } else {
        v = f < 3 ? R(1e18) : R(0);
      }
      lds[tri * GW + f] = v;
    }
  }
  __syncthreads();
  const int L = threadIdx.x & 63;
  const int I = 4 * M + (int)(threadIdx.x >> 6);
  if (I > J) return;  // wave-uniform; no barrier follows in this function
  const int t = I * 64 + L;
  const int tt = t < ne ? t : ne - 1;
  const R Cx = ps * a.tsl[0 * (size_t)ne + tt], Cy = ps * a.tsl[1 * (size_t)ne + tt],
          Cz = ps * a.tsl[2 * (size_t)ne + tt];
  const R F0x = a.tsl[3 * (size_t)ne + tt], F0y = a.tsl[4 * (size_t)ne + tt],
          F0z = a.tsl[5 * (size_t)ne + tt];
  const R F1x = a.tsl[6 * (size_t)ne + tt], F1y = a.tsl[7 * (size_t)ne + tt],
          F1z = a.tsl[8 * (size_t)ne + tt];
  R Clx = 0, Cly = 0, Clz = 0;
  if constexpr (HL) {
    Clx = a.tsl[(9 + NT + 0) * (size_t)ne + tt];
    Cly = a.tsl[(9 + NT + 1) * (size_t)ne + tt];
    Clz = a.tsl[(9 + NT + 2) * (size_t)ne + tt];
  }
  // (sq: the strengths of this workgroup's target points Q0 + qi)
  R E[NT], sq[NQ][NR], rc[NT], hc[NT];
  // P0-factored form: the target's point weights w_q J and densities
  R wt[PF ? NT : 1], st[PF ? NR : 1];
#pragma unroll
  for (int q = 0; q < NT; q++) {
    E[q] = ps * ps * a.tsl[(9 + q) * (size_t)ne + tt];
    rc[q] = ps * a.rc[q];
    hc[q] = R(0.5) * a.rc[q];  // double layer: c_q m.n = (c_q / 2) F.n
    if constexpr (PF) wt[q] = (R)a.tw[(size_t)q * ne + tt];
  }
  if constexpr (!PF) {
#pragma unroll
    for (int qi = 0; qi < NQ; qi++) {
      const int q = Q0 + qi;
#pragma unroll
      for (int r = 0; r < NR; r++)
        sq[qi][r] = t < ne && q < NT ? a.src[RW * ((size_t)tt * NT + q) + 3 + r] : R(0);
    }
  }
  if constexpr (PF) {
    const double w0 = a.tw[tt];
#pragma unroll
    for (int r = 0; r < NR; r++)
      st[r] = t < ne && w0 != 0.0
                  ? (R)((double)a.src[RW * (size_t)tt * NT + 3 + r] / w0) : R(0);
  }
  // touching triangles of group J (a lane past the mesh excludes them all)
  uint64_t mask = ~0ull;
  if (t < ne) {
    mask = 0;
    const int j0 = J * 64;
    for (int c = a.touch_ptr[tt]; c < a.touch_ptr[tt + 1]; c++) {
      const unsigned j = (unsigned)(a.touch[c] - j0);
      if (j < 64u) mask |= 1ull << j;
    }
  }
  // double layer: the target triangle's unit normal; D and c scale by ps
  R Nix = 0, Niy = 0, Niz = 0;
  if constexpr (DL) {
    Nix = (R)a.tn[3 * (size_t)tt];
    Niy = (R)a.tn[3 * (size_t)tt + 1];
    Niz = (R)a.tn[3 * (size_t)tt + 2];
  }
  const R hps = R(1) / ps;
  // Kelvin: target point q = C + o_q with o_q = (c_q / 2) F_k (doubled in
  // FP64, as C), and kappa times its strengths (the source side's
  // kappa-diagonal term)
  R oq[KEL ? NQ : 1][3], ksq[KEL ? NQ : 1][3];
  if constexpr (KEL) {
#pragma unroll
    for (int qi = 0; qi < NQ; qi++) {
      const int q = Q0 + qi < NT ? Q0 + qi : NT - 1;
      const int k = q < C0 ? 0 : (q - C0) % 3;
      const R h = q < C0 ? R(0) : R(0.5) * rc[q];
      oq[qi][0] = h * (k == 0 ? F0x : k == 1 ? F1x : -(F0x + F1x));
      oq[qi][1] = h * (k == 0 ? F0y : k == 1 ? F1y : -(F0y + F1y));
      oq[qi][2] = h * (k == 0 ? F0z : k == 1 ? F1z : -(F0z + F1z));
#pragma unroll
      for (int r = 0; r < 3; r++) ksq[qi][r] = a.kappa * sq[qi][r];
    }
  }
  R acc[NA][NR], S[NB][NR];
  // FP64 sums of the flushed steps (FL); the source side's stay aligned
  // with S: after a flush at step k, lane L holds source (L + k) mod 64 in both
  double acc64[NA][NR], S64[NB][NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
#pragma unroll
    for (int q = 0; q < NA; q++) {
      acc[q][r] = R(0);
      acc64[q][r] = 0.0;
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      S[b][r] = R(0);
      S64[b][r] = 0.0;
    }
  }
  // FL: moves the FP32 sums into the FP64 ones; the source side's FP64 sums
  // first move down `shift` lanes (the steps since the last flush)
  static_assert(!FL || NA == NB, "flushed sums: one target and source point set");
  auto flush = [&](int shift, auto sym_tag) {
    constexpr bool SYM = decltype(sym_tag)::value;
    const int from = ((L + shift) & 63) << 2;
#pragma unroll
    for (int q = 0; q < NA; q++)
#pragma unroll
      for (int r = 0; r < NR; r++) {
        acc64[q][r] += (double)acc[q][r];
        acc[q][r] = R(0);
        if constexpr (SYM) {
          const long long v = __double_as_longlong(S64[q][r]);
          const int lo = __builtin_amdgcn_ds_bpermute(from, (int)v);
          const int hi = __builtin_amdgcn_ds_bpermute(from, (int)(v >> 32));
          S64[q][r] = __longlong_as_double(((long long)hi << 32) | (unsigned)lo) +
                      (double)S[q][r];
          S[q][r] = R(0);
        }
      }
  };

  // 1/r from r2 (= r^2, FP64: (2r)^2) for the NT target points against one
  // source point, then both directions' sums.  Double layer: it[q] gathers
  // sigma_b / r^3 for target point q over the step's source points, and
  // source point b's sum of sigma_q / r^3 is scaled by its factor sfb.
  auto inv_r = [&](const R *r2, R *ri) {
    if constexpr (F64) {
      // r2 = 4 r^2: h = rsqrt(r2) ~ 1/(2r), and the Newton step for 1/r
      // from 2h is h (3 - r2 h^2): 3 instructions after rsqrt.  All
      // estimates first, so each result is not needed right after its
      // instruction.
      R h[NT];
#pragma unroll
      for (int q = 0; q < NT; q++) h[q] = __builtin_amdgcn_rsq(r2[q]);
#pragma unroll
      for (int q = 0; q < NT; q++) ri[q] = h[q] * fma(-r2[q], h[q] * h[q], R(3));
    } else {
#pragma unroll
      for (int q = 0; q < NT; q++) ri[q] = rsq(r2[q]);
    }
  };
  auto finish_point = [&](const R *r2, const R *sig, int b, auto sym_tag, R *it, R sfb,
                          const R *sqs) {
    constexpr bool SYM = decltype(sym_tag)::value;
    R ri[NT];
    inv_r(r2, ri);
    if constexpr (ADJ) {
      const R sg = sig[0] * sfb;
#pragma unroll
      for (int q = 0; q < NT; q++) {
        const R r3 = ri[q] * ri[q] * ri[q];
        acc[q][0] = fma(sg, r3, acc[q][0]);
        if constexpr (SYM) S[b][0] = fma(sqs[q], r3, S[b][0]);
      }
    } else if constexpr (DL) {
      R is = R(0);
#pragma unroll
      for (int q = 0; q < NT; q++) {
        const R r3 = ri[q] * ri[q] * ri[q];
        it[q] = fma(sig[0], r3, it[q]);
        if constexpr (SYM) is = fma(sq[q][0], r3, is);
      }
      if constexpr (SYM) S[b][0] = fma(sfb, is, S[b][0]);
    } else {
#pragma unroll
      for (int q = 0; q < NT; q++)
#pragma unroll
        for (int r = 0; r < NR; r++) {
          acc[q][r] = fma(sig[r], ri[q], acc[q][r]);
          if constexpr (SYM) S[b][r] = fma(sq[q][r], ri[q], S[b][r]);
        }
    }
  };

  // Kelvin, both directions of target point q and source point b with
  // difference d (doubled in FP64) and r2 = |d|^2 (+ the exclusion): the
  // kernel kappa I / r + d d^T / r^3 is even in d, so the source side takes
  // the target's strengths through the same 1/r and d; each side is
  // ri (kappa sigma + d (d.sigma) / r^2), the chunked kernel's form
  auto kelvin_pair = [&](R dx, R dy, R dz, R r2, const R *sb, const R *kb, int q, int b,
                         auto sym_tag) {
    constexpr bool SYM = decltype(sym_tag)::value;
    R ri;
    if constexpr (F64) {
      const R h = __builtin_amdgcn_rsq(r2);
      ri = h * fma(-r2, h * h, R(3));
    } else {
      ri = rsq(r2);
    }
    // 1/r^2 divided by the position scale squared
    const R ri2 = F64 ? R(0.25) * (ri * ri) : ri * ri;
    const R ds = fma(dx, sb[0], fma(dy, sb[1], dz * sb[2])) * ri2;
    acc[q][0] = fma(ri, fma(dx, ds, kb[0]), acc[q][0]);
    acc[q][1] = fma(ri, fma(dy, ds, kb[1]), acc[q][1]);
    acc[q][2] = fma(ri, fma(dz, ds, kb[2]), acc[q][2]);
    if constexpr (SYM) {
      const R us = fma(dx, sq[q][0], fma(dy, sq[q][1], dz * sq[q][2])) * ri2;
      S[b][0] = fma(ri, fma(dx, us, ksq[q][0]), S[b][0]);
      S[b][1] = fma(ri, fma(dy, us, ksq[q][1]), S[b][1]);
      S[b][2] = fma(ri, fma(dz, us, ksq[q][2]), S[b][2]);
    }
  };

  auto steps = [&](auto sym_tag) {
    constexpr bool SYM = decltype(sym_tag)::value;
    const int from = ((L + 1) & 63) << 2;  // ds_bpermute byte address
    RecT cur;
    if constexpr (ROT) cur = reinterpret_cast<const RecT *>(lds)[(L + k0) & 63];
    for (int k = k0; k < k1; k++) {
      const int jl = (L + k) & 63;
      // a touching source triangle (handled by the near field) gets r^2 +
      // 2^100 (FP32) or + 1e200 (FP64): 1/r below 1e-15 resp. 1e-100 in
      // both directions, so no strength needs masking
      const R excl = ((mask >> jl) & 1ull) ? Excl<R>::v : R(0);
      if constexpr (ROT) {
        const R Dx = Cx - cur.v[0], Dy = Cy - cur.v[1], Dz = Cz - cur.v[2];
        R r2[1];
        r2[0] = fma(Dx, Dx, fma(Dy, Dy, fma(Dz, Dz, excl)));
        if constexpr (KEL) {
          const R kb[3] = {a.kappa * cur.v[3], a.kappa * cur.v[4], a.kappa * cur.v[5]};
          kelvin_pair(Dx, Dy, Dz, r2[0], cur.v + 3, kb, 0, 0, sym_tag);
        } else {
          finish_point(r2, cur.v + 3, 0, sym_tag, nullptr, R(0), nullptr);
        }
#pragma unroll
        for (int c = 0; c < 3 + NR; c++) cur.v[c] = lane_rol1(cur.v[c]);
      } else {
        const R *g = lds + jl * GW;
        R Dx = Cx - g[0], Dy = Cy - g[1], Dz = Cz - g[2];
        if constexpr (HL) {
          Dx += Clx - g[Lay::LO];
          Dy += Cly - g[Lay::LO + 1];
          Dz += Clz - g[Lay::LO + 2];
        }
        if constexpr (KEL) {
          // d = D - o'_b + o_q with o'_b = (c_b / 2) F'_l of the source
#pragma unroll
          for (int b = 0; b < NT; b++) {
            R Dbx = Dx, Dby = Dy, Dbz = Dz;
            if (b >= C0) {
              const int lb = (b - C0) % 3;
              const R hb = R(-0.5) * rc[b];
              Dbx = fma(hb, lb == 0 ? g[3] : lb == 1 ? g[6] : -(g[3] + g[6]), Dx);
              Dby = fma(hb, lb == 0 ? g[4] : lb == 1 ? g[7] : -(g[4] + g[7]), Dy);
              Dbz = fma(hb, lb == 0 ? g[5] : lb == 1 ? g[8] : -(g[5] + g[8]), Dz);
            }
            const R *sb = g + Lay::SIG + b * NR;
            const R kb[3] = {a.kappa * sb[0], a.kappa * sb[1], a.kappa * sb[2]};
#pragma unroll
            for (int qi = 0; qi < NQ; qi++) {
              if (Q0 + qi >= NT) continue;  // (compile-time: the second half's end)
              const R dx = Dbx + oq[qi][0], dy = Dby + oq[qi][1], dz = Dbz + oq[qi][2];
              kelvin_pair(dx, dy, dz, fma(dx, dx, fma(dy, dy, fma(dz, dz, excl))), sb, kb,
                          qi, b, sym_tag);
            }
          }
        } else {
        const R A0 = fma(Dx, Dx, fma(Dy, Dy, fma(Dz, Dz, excl)));
        // D.F_k (target directions), D.F'_l (source directions), F_k.F'_l
        R U[3], P[3], W[3][3];
        U[0] = fma(Dx, F0x, fma(Dy, F0y, Dz * F0z));
        U[1] = fma(Dx, F1x, fma(Dy, F1y, Dz * F1z));
        U[2] = -(U[0] + U[1]);
        P[0] = fma(Dx, g[3], fma(Dy, g[4], Dz * g[5]));
        P[1] = fma(Dx, g[6], fma(Dy, g[7], Dz * g[8]));
        P[2] = -(P[0] + P[1]);
        if constexpr (NO > 0) {
          W[0][0] = fma(F0x, g[3], fma(F0y, g[4], F0z * g[5]));
          W[0][1] = fma(F0x, g[6], fma(F0y, g[7], F0z * g[8]));
          W[1][0] = fma(F1x, g[3], fma(F1y, g[4], F1z * g[5]));
          W[1][1] = fma(F1x, g[6], fma(F1y, g[7], F1z * g[8]));
          W[0][2] = -(W[0][0] + W[0][1]);
          W[1][2] = -(W[1][0] + W[1][1]);
          W[2][0] = -(W[0][0] + W[1][0]);
          W[2][1] = -(W[0][1] + W[1][1]);
          W[2][2] = -(W[0][2] + W[1][2]);
        }
        R B[NT];
#pragma unroll
        for (int q = 0; q < NT; q++) B[q] = q < C0 ? R(0) : fma(rc[q], U[(q - C0) % 3], E[q]);
        // double layer: tf[q] = (x_q - y).n_j and sf[b] = (y_b - x).n_i;
        // adjoint: own strengths times tf
        R it[NT], tf[NT], sf[NT], sqs[NT];
        if constexpr (DL) {
          const R *nj = g + Lay::NRM;
          const R dt = hps * fma(Dx, nj[0], fma(Dy, nj[1], Dz * nj[2]));
          const R ds = -hps * fma(Dx, Nix, fma(Dy, Niy, Dz * Niz));
          R Gt[3], Gs[3];
          Gt[0] = fma(F0x, nj[0], fma(F0y, nj[1], F0z * nj[2]));
          Gt[1] = fma(F1x, nj[0], fma(F1y, nj[1], F1z * nj[2]));
          Gt[2] = -(Gt[0] + Gt[1]);
          Gs[0] = fma(g[3], Nix, fma(g[4], Niy, g[5] * Niz));
          Gs[1] = fma(g[6], Nix, fma(g[7], Niy, g[8] * Niz));
          Gs[2] = -(Gs[0] + Gs[1]);
#pragma unroll
          for (int q = 0; q < NT; q++) {
            tf[q] = q < C0 ? dt : fma(hc[q], Gt[(q - C0) % 3], dt);
            sf[q] = q < C0 ? ds : fma(hc[q], Gs[(q - C0) % 3], ds);
            it[q] = R(0);
            if constexpr (ADJ) sqs[q] = sq[q][0] * tf[q];
          }
        }
        // P0-factored form: G = sum_q w_q J sum_b w'_b J' / r_qb of the
        // triangle pair (NT independent sums over b, then one over q), then
        // one FMA per density and direction
        R hq[PF ? NT : 1];
#pragma unroll
        for (int q = 0; q < (PF ? NT : 1); q++) hq[q] = R(0);
#pragma unroll
        for (int b = 0; b < NT; b++) {
          const int lb = (b - C0) % 3, ob = (b - C0) / 3;
          const R Ab = b < C0 ? A0 : fma(-rc[b], P[lb], A0 + g[9 + b - C0]);
          R r2[NT];
#pragma unroll
          for (int q = 0; q < NT; q++) {
            const int kq = (q - C0) % 3, oq = (q - C0) / 3;
            r2[q] = q < C0 ? Ab
                    : b < C0 ? Ab + B[q]
                             : fma(a.kap[2 * oq + ob], W[kq][lb], Ab + B[q]);
          }
          if constexpr (PF) {
            R ri[NT];
            inv_r(r2, ri);
            const R wb = g[Lay::SIG + b];
#pragma unroll
            for (int q = 0; q < NT; q++) hq[q] = fma(wb, ri[q], hq[q]);
          } else {
            finish_point(r2, g + Lay::SIG + b * NR, b, sym_tag, it, DL ? sf[b] : R(0), sqs);
          }
        }
        if constexpr (PF) {
          constexpr bool SYM = decltype(sym_tag)::value;
          R G = R(0);
#pragma unroll
          for (int q = 0; q < NT; q++) G = fma(wt[q], hq[q], G);
          const R *sj = g + Lay::SIG + NT;
#pragma unroll
          for (int r = 0; r < NR; r++) {
            acc[0][r] = fma(G, sj[r], acc[0][r]);
            if constexpr (SYM) S[0][r] = fma(G, st[r], S[0][r]);
          }
        }
        if constexpr (DL && !ADJ) {
#pragma unroll
          for (int q = 0; q < NT; q++) acc[q][0] = fma(tf[q], it[q], acc[q][0]);
        }
        }  // not KEL
      }
      if constexpr (SYM) {
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
          for (int r = 0; r < NR; r++) {
            if constexpr (ROT) {
              S[b][r] = lane_rol1(S[b][r]);
            } else if constexpr (sizeof(R) == 4) {
              S[b][r] = __int_as_float(
                  __builtin_amdgcn_ds_bpermute(from, __float_as_int(S[b][r])));
            } else {
              const long long v = __double_as_longlong(S[b][r]);
              const int lo = __builtin_amdgcn_ds_bpermute(from, (int)v);
              const int hi = __builtin_amdgcn_ds_bpermute(from, (int)(v >> 32));
              S[b][r] = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
            }
          }
      }
      if constexpr (FL)
        if ((k + 1 - k0) % KF == 0) flush(KF, sym_tag);
    }
    if constexpr (FL)
      if ((k1 - k0) % KF) flush((k1 - k0) % KF, sym_tag);
  };
  if (I < J) steps(std::true_type{});
  else steps(std::false_type{});
  if constexpr (!FL) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
#pragma unroll
      for (int q = 0; q < NA; q++) acc64[q][r] = (double)acc[q][r];
#pragma unroll
      for (int b = 0; b < NB; b++) S64[b][r] = (double)S[b][r];
    }
  }

  // point sums of triangle tri to partial row c: the target side's
  // (partial[c - c0][r][.][tri - t0]) or the source side's (spart[...])
  // (v: NV point sums of points q0 + qi -- the target side's NA from Q0,
  // the source side's NB from 0)
  auto store = [&](const auto &v, auto q0_tag, int tri, int c, bool src_side) {
    constexpr int NV = std::extent<std::remove_reference_t<decltype(v)>>::value;
    constexpr int q0 = decltype(q0_tag)::value;
    const size_t stride = (size_t)ne;
    double *base = src_side ? a.spart : a.partial;
    const int rowv = src_side ? a.sprow : a.prow, cols = src_side ? a.scols : a.tcols;
    const int tl = tri - a.t0;
    double wq[NV];
#pragma unroll
    for (int qi = 0; qi < NV; qi++)
      wq[qi] = q0 + qi < NT ? a.tw[(q0 + qi) * stride + tri] : 0.0;
    double *row = base + (size_t)(c - a.c0) * NR * rowv;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if constexpr (PF) {
        // the weights are inside G
        row[(size_t)r * rowv + tl] = v[0][r];
      } else if (a.nout == 1) {
        double s = 0.0;
#pragma unroll
        for (int qi = 0; qi < NV; qi++) s = fma(wq[qi], v[qi][r], s);
        row[(size_t)r * rowv + tl] = s;
      } else {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          double s = 0.0;
#pragma unroll
          for (int qi = 0; qi < NV; qi++)
            if (q0 + qi < NT) s = fma(wq[qi] * a.phi[q0 + qi][k], v[qi][r], s);
          // planes: right-hand side r, then output k; Kelvin: output k, then
          // component r (the chunked kernel's order)
          if constexpr (KEL) row[((size_t)k * NR + r) * cols + tl] = s;
          else row[(size_t)r * rowv + (a.pos ? a.pos[k * ne + tri] : k * cols + tl)] = s;
        }
      }
    }
  };
  if (t < ne) store(acc64, std::integral_constant<int, Q0>{}, t, J * ks + rng, false);
  if (I < J) {
    // after the steps k0 .. k1-1 lane L holds source (L + k1) mod 64
    const int j = J * 64 + ((L + k1) & 63);
    if (j < ne) store(S64, std::integral_constant<int, 0>{}, j, I * ks + rng, true);
  }
}

// at least 2 waves per SIMD: the FP64 7-point single layer sits at the
// 256-register edge, and one register more would halve its occupancy
// (the FP32 7-point double layer takes 129 registers, 3 waves per SIMD;
// asking for 4 spills two values and measures the same: 3.01 against 3.00
// ms at L5, 41.4 against 42.2 ms at L6)
template <typename R, int NT, int NR, int DK = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_far_sym(
    FarArgs<R> a)
{
  __shared__ __align__(16) R lds[SymLay<R, NT, NR, DK>::LDS];
  sym_block<R, NT, NR, DK>(a, (int)blockIdx.x, lds);
}

// The Kelvin single layer (three components) on the symmetric schedule.
// FP64 holds 7 x 3 target sums, strengths, kappa-scaled strengths and point
// offsets plus the source side's sums in registers; with 6 or 7 points that
// is more than 256 per lane (391 at 7: one wave per SIMD and spills), so
// FP64 takes those rules in two target-point halves (sym_block QS = 2), each
// workgroup one half: slot rng = step range * 2 + half.
template <typename R, int NT>
struct KelQS { static constexpr int v = sizeof(R) == 8 && NT >= 6 ? 2 : 1; };
template <typename R, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_far_sym_kel(
    FarArgs<R> a)
{
  __shared__ __align__(16) R lds[SymLay<R, NT, 3, 3>::LDS];
  constexpr int QS = KelQS<R, NT>::v;
  if constexpr (QS == 1) {
    sym_block<R, NT, 3, 3>(a, (int)blockIdx.x, lds);
  } else {
    if ((int)blockIdx.x % a.ksplit % QS == 0)
      sym_block<R, NT, 3, 3, QS, 0>(a, (int)blockIdx.x, lds);
    else
      sym_block<R, NT, 3, 3, QS, 1>(a, (int)blockIdx.x, lds);
  }
}

// Persistent multi-mesh form of the symmetric schedule: the workgroups of
// several operators' launches are items of one queue (operator k owns items
// off[k] .. off[k+1]-1, item off[k] + b being its workgroup b), taken from
// an atomic counter until the queue is empty.  At least 2 waves per SIMD, as
// k_far_sym: FP64 at 7 points would otherwise take 261 registers and run
// one wave per SIMD.
template <typename R, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_far_sym_multi(
    const FarArgs<R> *__restrict__ descs,
                                                       const int *__restrict__ off,
                                                       int nops, int *__restrict__ counter)
{
  constexpr int L0 = SymLay<R, NT, 1, 0>::LDS, L4 = SymLay<R, NT, 1, 4>::LDS;
  __shared__ __align__(16) R lds[L0 > L4 ? L0 : L4];
  __shared__ int sh_item;
  const int total = off[nops];
  while (true) {
    __syncthreads();  // sh_item and the LDS records of the last item are free
    if (threadIdx.x == 0) sh_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = sh_item;
    if (item >= total) break;
    int k = 0;
    while (item >= off[k + 1]) k++;
    if constexpr (NT > 1)
      if (descs[k].pf) {  // P0 members, as their own applies
        sym_block<R, NT, 1, 4>(descs[k], item - off[k], lds);
        continue;
      }
    sym_block<R, NT, 1>(descs[k], item - off[k], lds);
  }
}

// Multi-RHS single layer.  One lane per target quadrature point, RB
// right-hand sides per lane; a block holds TB = blockDim/NT whole triangles.
// Source records are read with wave-uniform addresses (scalar loads: the
// strengths arrive as SGPR operands of v_fmac), so there is no LDS staging
// and no block barrier in the main loop.  Per interaction: 3 sub, 3 FMA,
// rsqrt, one multiply by the 0/1 exclusion factor, RB FMAs.
template <typename R>
struct FarMArgs {
  int ne, chunks, chunk, nout, groups, nrhs;
  const R *tq;                  // [nt][3][ne]
  const R *geom;                // [ne*nt][4]
  const R *sig;                 // [groups][ne*nt][RB]
  const double *sigd;           // [groups][ne][RB] (k_far_mrhs_tri)
  const R *tsl;                 // median rules: target frame, as FarArgs::tsl
  R rc[16];                     // median rules: c_q of point q
  R kap[4];                     // median rules: -c c' / 2 for orbits (o, o')
  const int *touch_ptr, *touch;
  const double *tw;             // [nt][ne]
  double *partial;              // [chunks][nrhs][nout][ne]
  double phi[16][3];
};

template <typename R, int NT, int RB>
__global__ __launch_bounds__(256) void k_far_mrhs(FarMArgs<R> a)
{
  constexpr int TB = 256 / NT;  // whole triangles per block
  __shared__ double red[256];
  const int ne = a.ne;
  int b = blockIdx.x;
  const int s = b % a.chunks;
  b /= a.chunks;
  const int g = b % a.groups;
  const int tb = b / a.groups;
  const int lt = threadIdx.x / NT, q = threadIdx.x - lt * NT;
  const bool lane_ok = lt < TB;
  const int t = tb * TB + lt;
  const int tt = (lane_ok && t < ne) ? t : ne - 1;
  const int qq = lane_ok ? q : 0;
  const int j0 = s * a.chunk, j1 = min(ne, j0 + a.chunk);
  const R X = a.tq[(3 * qq + 0) * (size_t)ne + tt];
  const R Y = a.tq[(3 * qq + 1) * (size_t)ne + tt];
  const R Z = a.tq[(3 * qq + 2) * (size_t)ne + tt];
  double acc64[RB];
#pragma unroll
  for (int r = 0; r < RB; r++) acc64[r] = 0.0;
  int cur = a.touch_ptr[tt];
  const int cend = a.touch_ptr[tt + 1];
  while (cur < cend && a.touch[cur] < j0) cur++;
  int nextx = cur < cend ? a.touch[cur] : INT_MAX;
  const R eps = Eps<R>::v;
  const R *__restrict__ geom = a.geom;
  const R *__restrict__ sig = a.sig + (size_t)g * ne * NT * RB;
  for (int js = j0; js < j1; js += 64) {
    const int nt = min(64, j1 - js);
    uint64_t mask = 0;
    while (nextx < js + nt) {
      mask |= 1ull << (nextx - js);
      ++cur;
      nextx = cur < cend ? a.touch[cur] : INT_MAX;
    }
    R acc[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) acc[r] = R(0);
    for (int jj = 0; jj < nt; jj++) {
      const R keep = (R)(1u - (unsigned)((mask >> jj) & 1ull));
      const int p0 = (js + jj) * NT;
#pragma unroll
      for (int bq = 0; bq < NT; bq++) {
        const R *gp = geom + (size_t)(p0 + bq) * 4;
        const R dx = X - gp[0], dy = Y - gp[1], dz = Z - gp[2];
        const R r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
        const R ri = rsq(r2) * keep;
        const R *sp = sig + (size_t)(p0 + bq) * RB;
#pragma unroll
        for (int r = 0; r < RB; r++) acc[r] = fma(sp[r], ri, acc[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < RB; r++) acc64[r] += (double)acc[r];
  }
  // reduce the NT points of each triangle through LDS, one RHS at a time
  const double wq = lane_ok ? a.tw[(size_t)qq * ne + tt] : 0.0;
  const int nrhs_here = min(RB, a.nrhs - g * RB);
  const int nred = TB * a.nout;
  for (int r = 0; r < nrhs_here; r++) {
    __syncthreads();
    red[threadIdx.x] = acc64[r] * wq;
    __syncthreads();
    if (threadIdx.x < nred) {
      const int k = threadIdx.x / TB, lt2 = threadIdx.x - k * TB;
      const int t2 = tb * TB + lt2;
      if (t2 < ne) {
        double v = 0.0;
#pragma unroll
        for (int q2 = 0; q2 < NT; q2++)
          v += red[lt2 * NT + q2] * (a.nout == 1 ? 1.0 : a.phi[q2][k]);
        a.partial[(((size_t)s * a.nrhs + g * RB + r) * a.nout + k) * ne + t2] = v;
      }
    }
  }
}

// Multi-RHS single layer with piecewise-constant test and trial functions.
// A source triangle's strength is one value per RHS over all its points, and
// a target triangle's output is its points' weighted sum, so each triangle
// pair first sums G = sum_q w_q sum_b w_b / |x_q - y_b| over its NT x NT point
// pairs (in R) and then spends RB FP64 FMAs on G * sigma_j: RB FMAs per
// triangle pair instead of per point pair.  One lane per target triangle;
// source records as in k_far_mrhs (wave-uniform addresses, scalar loads),
// with the source weight in the fourth slot of each point and FP64 strengths
// per triangle.  FP64 and non-packed FP32 FMAs issue at one rate, so the FP64
// sums cost no more than FP32 ones and need no FP32 tile in registers.
//
// MED (rules of 6 or 7 points on the medians, point q = C + c_q (P_k - C)):
// r^2 of each point pair from triangle-pair terms as in the pair-symmetric
// kernel, |D|^2 + E_q + E'_b + c_q D.F_k - c'_b D.F'_l - (c_q c'_b / 2) F_k.F'_l
// with D = C - C' and F = 2 (P - C): one add and one FMA per point pair.
// The source record of triangle j is then C', F'_0, F'_1, E'_b, w_b and, in
// FP32, the rounding remainder C' - (float)C' ([ne][MedRec]) instead of its
// points; FP32 forms D as (C_hi - C'_hi) + (C_lo - C'_lo), so that D carries
// no rounding of the positions (systematic for each triangle pair).  The
// record fits the 4 values per point of the point form.
template <int NT> struct MedRec {
  static constexpr int LO = 9 + 2 * NT;
  static constexpr int W = (LO + 3 + 3) / 4 * 4;
  static_assert(LO + 3 <= W && W <= 4 * NT, "median record width");
};

// A target triangle of the triangle-pair kernels: its points (MED: its
// frame and the E_q, c_q of its points) and weights in registers.  sum(a, j,
// keep) is G of the pair with source triangle j, 0 when keep is 0.
template <typename R, int NT, bool MED>
struct TriTarget {
  static constexpr bool HL = MED && sizeof(R) == 4;
  R X[NT], Y[NT], Z[NT], W[NT];
  R Cx = 0, Cy = 0, Cz = 0, F0x = 0, F0y = 0, F0z = 0, F1x = 0, F1y = 0, F1z = 0;
  R Lx = 0, Ly = 0, Lz = 0;  // FP32: C - (float)C

  __device__ __forceinline__ void load(const FarMArgs<R> &a, int tt)
  {
    const int ne = a.ne;
#pragma unroll
    for (int q = 0; q < NT; q++) {
      if constexpr (MED) {
        X[q] = a.tsl[(9 + q) * (size_t)ne + tt];  // E_q
        Y[q] = a.rc[q];
        Z[q] = R(0);
      } else {
        X[q] = a.tq[(3 * q + 0) * (size_t)ne + tt];
        Y[q] = a.tq[(3 * q + 1) * (size_t)ne + tt];
        Z[q] = a.tq[(3 * q + 2) * (size_t)ne + tt];
      }
      W[q] = (R)a.tw[(size_t)q * ne + tt];
    }
    if constexpr (MED) {
      Cx = a.tsl[0 * (size_t)ne + tt];
      Cy = a.tsl[1 * (size_t)ne + tt];
      Cz = a.tsl[2 * (size_t)ne + tt];
      F0x = a.tsl[3 * (size_t)ne + tt];
      F0y = a.tsl[4 * (size_t)ne + tt];
      F0z = a.tsl[5 * (size_t)ne + tt];
      F1x = a.tsl[6 * (size_t)ne + tt];
      F1y = a.tsl[7 * (size_t)ne + tt];
      F1z = a.tsl[8 * (size_t)ne + tt];
    }
    if constexpr (HL) {
      Lx = a.tsl[(9 + NT) * (size_t)ne + tt];
      Ly = a.tsl[(10 + NT) * (size_t)ne + tt];
      Lz = a.tsl[(11 + NT) * (size_t)ne + tt];
    }
  }

  __device__ __forceinline__ R sum(const FarMArgs<R> &a, int j, R keep) const
  {
    const R eps = Eps<R>::v;
    if constexpr (MED) {
      constexpr int C0 = NT % 3;
      const R *gp = a.geom + (size_t)j * MedRec<NT>::W;
      R Dx = Cx - gp[0], Dy = Cy - gp[1], Dz = Cz - gp[2];
      if constexpr (HL) {
        constexpr int LO = MedRec<NT>::LO;
        Dx += Lx - gp[LO];
        Dy += Ly - gp[LO + 1];
        Dz += Lz - gp[LO + 2];
      }
      // a touching pair's expanded r^2 can round below zero at coincident
      // points: lift it far away before the rsqrt (and multiply by 0 after)
      const R excl = keep == R(0) ? R(1e30) : R(0);
      const R A0 = fma(Dx, Dx, fma(Dy, Dy, fma(Dz, Dz, eps + excl)));
      R U[3], P[3], Wk[3][3];
      U[0] = fma(Dx, F0x, fma(Dy, F0y, Dz * F0z));
      U[1] = fma(Dx, F1x, fma(Dy, F1y, Dz * F1z));
      U[2] = -(U[0] + U[1]);
      P[0] = fma(Dx, gp[3], fma(Dy, gp[4], Dz * gp[5]));
      P[1] = fma(Dx, gp[6], fma(Dy, gp[7], Dz * gp[8]));
      P[2] = -(P[0] + P[1]);
      Wk[0][0] = fma(F0x, gp[3], fma(F0y, gp[4], F0z * gp[5]));
      Wk[0][1] = fma(F0x, gp[6], fma(F0y, gp[7], F0z * gp[8]));
      Wk[1][0] = fma(F1x, gp[3], fma(F1y, gp[4], F1z * gp[5]));
      Wk[1][1] = fma(F1x, gp[6], fma(F1y, gp[7], F1z * gp[8]));
      Wk[0][2] = -(Wk[0][0] + Wk[0][1]);
      Wk[1][2] = -(Wk[1][0] + Wk[1][1]);
#pragma unroll
      for (int l = 0; l < 3; l++) Wk[2][l] = -(Wk[0][l] + Wk[1][l]);
      R B[NT], Gq[NT];
#pragma unroll
      for (int q = 0; q < NT; q++) {
        B[q] = q < C0 ? R(0) : fma(Y[q], U[(q - C0) % 3], X[q]);
        Gq[q] = R(0);
      }
#pragma unroll
      for (int b = 0; b < NT; b++) {
        const int lb = (b - C0) % 3, ob = (b - C0) / 3;
        const R Ab = b < C0 ? A0 : fma(-a.rc[b], P[lb], A0 + gp[9 + b]);
        const R wb = gp[9 + NT + b];
#pragma unroll
        for (int q = 0; q < NT; q++) {
          const int kq = (q - C0) % 3, oq = (q - C0) / 3;
          const R r2 = q < C0   ? Ab
                       : b < C0 ? Ab + B[q]
                                : fma(a.kap[2 * oq + ob], Wk[kq][lb], Ab + B[q]);
          Gq[q] = fma(wb, rsq(r2), Gq[q]);
        }
      }
      R G = R(0);
#pragma unroll
      for (int q = 0; q < NT; q++) G = fma(W[q], Gq[q], G);
      return G * keep;
    }
    const R *gp = a.geom + (size_t)j * NT * 4;
    R G = R(0);
#pragma unroll
    for (int q = 0; q < NT; q++) {
      R Gq = R(0);
#pragma unroll
      for (int bq = 0; bq < NT; bq++) {
        const R dx = X[q] - gp[4 * bq], dy = Y[q] - gp[4 * bq + 1],
                dz = Z[q] - gp[4 * bq + 2];
        const R r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
        Gq = fma(gp[4 * bq + 3], rsq(r2), Gq);
      }
      G = fma(W[q], Gq, G);
    }
    return G * keep;
  }
};

template <typename R, int NT, int RB, bool MED>
__global__ __launch_bounds__(256) void k_far_mrhs_tri(FarMArgs<R> a)
{
  const int ne = a.ne;
  int b = blockIdx.x;
  const int s = b % a.chunks;
  b /= a.chunks;
  const int g = b % a.groups;
  const int tb = b / a.groups;
  const int t = tb * 256 + (int)threadIdx.x;
  const int tt = t < ne ? t : ne - 1;
  const int j0 = s * a.chunk, j1 = min(ne, j0 + a.chunk);
  TriTarget<R, NT, MED> tgt;
  tgt.load(a, tt);
  double acc64[RB];
#pragma unroll
  for (int r = 0; r < RB; r++) acc64[r] = 0.0;
  int cur = a.touch_ptr[tt];
  const int cend = a.touch_ptr[tt + 1];
  while (cur < cend && a.touch[cur] < j0) cur++;
  int nextx = cur < cend ? a.touch[cur] : INT_MAX;
  const double *__restrict__ sig = a.sigd + (size_t)g * ne * RB;
  for (int js = j0; js < j1; js += 64) {
    const int nt = min(64, j1 - js);
    uint64_t mask = 0;
    while (nextx < js + nt) {
      mask |= 1ull << (nextx - js);
      ++cur;
      nextx = cur < cend ? a.touch[cur] : INT_MAX;
    }
    for (int jj = 0; jj < nt; jj++) {
      // G of the pair (t, js + jj), 0 for an excluded (touching) pair
      const R keep = (R)(1u - (unsigned)((mask >> jj) & 1ull));
      const double G = (double)tgt.sum(a, js + jj, keep);
      const double *sp = sig + (size_t)(js + jj) * RB;
#pragma unroll
      for (int r = 0; r < RB; r++) acc64[r] = fma(sp[r], G, acc64[r]);
    }
  }
  if (t < ne) {
    const int nrhs_here = min(RB, a.nrhs - g * RB);
    for (int r = 0; r < nrhs_here; r++)
      a.partial[((size_t)s * a.nrhs + g * RB + r) * ne + t] = acc64[r];
  }
}

// Pair-symmetric multi-RHS schedule for P0 test and trial functions.  G of
// a triangle pair (k_far_mrhs_tri's triangle-pair sum; the kernel and the
// point sets are symmetric) serves both orders of the pair, so it is formed
// once per unordered pair of 64-triangle groups and applied in both
// directions by a second kernel (k_mrhs_tile_mfma): half the point-pair
// work of k_far_mrhs_tri, and the RB FP64 products per ordered pair on the
// matrix cores as 64 x 64 x RB tile products.
//
// Tiles of a pass: the group pairs (I, J), I0 <= I < I1, J >= I, row by
// row (tile (I, J) has global index rowoff[I] + J - I); tile (I, J) holds
// G(64 I + i, 64 J + j) at [j][i], 0 for touching pairs and for triangles
// past the mesh.  Items of a pass (K, kind, M0, M1), each one wave and at
// most a fixed number of tiles, give output group K the contributions of
// the pass's tiles: kind 1, tiles (M, K) transposed for M0 <= M < M1 (M <
// K); kind 0, tiles (K, M) for M0 <= M < M1 (M >= K, the rows of the
// pass).  The items of K are listed in increasing M, and
// k_mrhs_tile_reduce adds them in that order to K's sums of the earlier
// passes: a fixed summation order, no atomics.  The buffers hold one
// pass's tiles and item sums, bounded by the operator's partial limit.
struct MtArgs {
  int ng, I0, I1, t0, ntiles, items, groups, first;
  const int *rowoff;  // [ng + 1] first tile of each row
  const int4 *item;   // (K, kind, M0, M1) of the pass's items
  const int *order;   // the items in launch order (longest first)
  const int *kptr;    // [ng - I0 + 1]: items of output group I0 + k
  double *part;       // [items][groups][RB][64] item sums
};

// One wave a tile: lane i pairs its target triangle 64 I + i with the 64
// source triangles of group J in order (wave-uniform source records).
template <typename R, int NT, bool MED>
__global__ __launch_bounds__(256) void k_mrhs_tiles(FarMArgs<R> a, MtArgs m,
                                                    R *__restrict__ tiles)
{
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (w >= m.ntiles) return;  // wave-uniform; no barrier in this kernel
  const int L = threadIdx.x & 63;
  const int tg = m.t0 + w;
  int lo = m.I0, hi = m.I1 - 1;  // the row of tile tg
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.rowoff[mid] <= tg) lo = mid;
    else hi = mid - 1;
  }
  const int I = lo, J = I + (tg - m.rowoff[I]);
  const int ne = a.ne, t = I * 64 + L, tt = t < ne ? t : ne - 1;
  TriTarget<R, NT, MED> tgt;
  tgt.load(a, tt);
  const int j0 = J * 64, nj = min(64, ne - j0);
  // touching triangles of group J (a lane past the mesh excludes them all)
  uint64_t mask = ~0ull;
  if (t < ne) {
    mask = 0;
    for (int c = a.touch_ptr[tt]; c < a.touch_ptr[tt + 1]; c++) {
      const unsigned j = (unsigned)(a.touch[c] - j0);
      if (j < 64u) mask |= 1ull << j;
    }
  }
  R *__restrict__ out = tiles + (size_t)w * 4096 + L;
  for (int jj = 0; jj < 64; jj++) {
    R G = R(0);
    if (jj < nj) {
      const R keep = (R)(1u - (unsigned)((mask >> jj) & 1ull));
      G = tgt.sum(a, j0 + jj, keep);
    }
    out[jj * 64] = G;
  }
}

// One wave an item, on the matrix cores: per tile the item's 64 outputs
// take Out(64 x RB) += G(64 x 64) X(64 x RB) in FP64 MFMA 16x16x4 blocks:
// four 16-output blocks (ib) times RB / 16 right-hand-side blocks (rb), 4
// sources a step; A = G, lane i + 16 k holding output i, source k of a 16
// x 4 block; B = X, lane r + 16 k holding source k, right-hand side r;
// D: output i, right-hand side r at lane r + 16 (i % 4), item i / 4
// (measured on gfx950).  The operands of 4 steps (16 sources) are loaded
// together, then their 64 (RB = 64) MFMAs run.  RB = 8 fills half of each
// B block, the other lanes 0.
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename R, int RB, bool ROW>
__device__ __forceinline__ void mfma_item(const FarMArgs<R> &a, const MtArgs &m,
                                          const R *__restrict__ tiles, int w, int4 it)
{
  constexpr int NRB = RB >= 16 ? RB / 16 : 1, SB = 4;  // steps a block
  const int g = blockIdx.y, l = threadIdx.x & 63, ne = a.ne;
  const int li = l & 15, lk = l >> 4, K = it.x;
  const double *__restrict__ sig = a.sigd + (size_t)g * ne * RB;
  v4d acc[4][NRB];
#pragma unroll
  for (int ib = 0; ib < 4; ib++)
#pragma unroll
    for (int rb = 0; rb < NRB; rb++) acc[ib][rb] = v4d{0.0, 0.0, 0.0, 0.0};
  // block b: tile M = M0 + b / 4, sources 16 (b % 4) .. + 15 of it; row
  // items read tile (K, M) at [source][output], column items tile (M, K),
  // M < K, at [output][source]
  const int nblk = (it.w - it.z) * 4;
  for (int b = 0; b < nblk; b++) {
    const int M = it.z + (b >> 2), s0 = (b & 3) * 16 + lk;
    const R *T = ROW ? tiles + (size_t)(m.rowoff[K] + (M - K) - m.t0) * 4096 + li
                     : tiles + (size_t)(m.rowoff[M] + (K - M) - m.t0) * 4096 + li * 64;
    R ga[SB][4];
    double xb[SB][NRB];
#pragma unroll
    for (int st = 0; st < SB; st++) {
      const int src = s0 + 4 * st;
#pragma unroll
      for (int ib = 0; ib < 4; ib++) ga[st][ib] = ROW ? T[src * 64 + ib * 16] : T[ib * 16 * 64 + src];
      // a source past the mesh (last group) has G = 0; its index is clamped
      const double *xp = sig + (size_t)min(M * 64 + src, ne - 1) * RB;
#pragma unroll
      for (int rb = 0; rb < NRB; rb++)
        xb[st][rb] = (RB >= 16 || li < RB) ? xp[rb * 16 + li] : 0.0;
    }
#pragma unroll
    for (int st = 0; st < SB; st++)
#pragma unroll
      for (int ib = 0; ib < 4; ib++) {
        const double av = (double)ga[st][ib];
#pragma unroll
        for (int rb = 0; rb < NRB; rb++)
          acc[ib][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xb[st][rb], acc[ib][rb], 0, 0, 0);
      }
  }
  const int nrhs_here = min(RB, a.nrhs - g * RB);
  double *__restrict__ out = m.part + ((size_t)w * m.groups + g) * RB * 64;
#pragma unroll
  for (int ib = 0; ib < 4; ib++)
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = rb * 16 + li;
        if (r < nrhs_here) out[r * 64 + ib * 16 + 4 * v + lk] = acc[ib][rb][v];
      }
}

template <typename R, int RB>
__global__ __launch_bounds__(256) void k_mrhs_tile_mfma(FarMArgs<R> a, MtArgs m,
                                                        const R *__restrict__ tiles)
{
  const int wl = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (wl >= m.items) return;  // wave-uniform; no barrier in this kernel
  const int w = m.order[wl];
  const int4 it = m.item[w];
  if (it.y == 0) mfma_item<R, RB, true>(a, m, tiles, w, it);
  else mfma_item<R, RB, false>(a, m, tiles, w, it);
}

// Output group K = I0 + blockIdx.x / (RB / 4), right-hand side r of group
// blockIdx.y (4 a workgroup, one wave each), triangle lane: the sums of the
// earlier passes (none in the first) plus the pass's items of K in order.
template <int RB>
__global__ __launch_bounds__(256) void k_mrhs_tile_reduce(MtArgs m, int ne, int nrhs,
                                                          double *__restrict__ out)
{
  const int g = blockIdx.y, L = threadIdx.x & 63;
  const int kr = blockIdx.x / (RB / 4);
  const int r = (blockIdx.x - kr * (RB / 4)) * 4 + (int)(threadIdx.x >> 6);
  const int t = (m.I0 + kr) * 64 + L, rr = g * RB + r;
  if (t >= ne || rr >= nrhs) return;
  double s = m.first ? 0.0 : out[(size_t)rr * ne + t];
  for (int i = m.kptr[kr]; i < m.kptr[kr + 1]; i++)
    s += m.part[(((size_t)i * m.groups + g) * RB + r) * 64 + L];
  out[(size_t)rr * ne + t] = s;
}

// Source records of the multi-RHS kernels.  Per point strengths
// w_b * scale * x(y_b) (k_far_mrhs), or with TRI (P0 trial, k_far_mrhs_tri)
// the weight w_b in the point record and scale * x_j once per triangle.
template <typename R, int RB, bool TRI>
__global__ void k_fold_mrhs(int ne, int nt, int nrhs, int groups,
                            const double *__restrict__ sg,
                            const double *__restrict__ swt,
                            const int *__restrict__ tri, int trial_p1,
                            RuleArgs rule, const double *__restrict__ x,
                            int ncols, double scale, R *__restrict__ geom,
                            R *__restrict__ sig, double *__restrict__ sigd,
                            const R *__restrict__ tsl, int med)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int npts = ne * nt;
  if (idx >= npts * groups) return;
  const int g = idx / npts, p = idx - g * npts;
  const int j = p / nt, b = p - j * nt;
  if (TRI && med) {
    // median record of source triangle j (k_far_mrhs_tri, MED)
    if (g == 0) {
      R *o = geom + (size_t)j * med;
      if (b == 0) {
        for (int f = 0; f < 9; f++) o[f] = tsl[(size_t)f * ne + j];
        // FP32: the centroid's rounding remainders (target-frame rows
        // 9 + nt .. 11 + nt) after the weights
        if constexpr (sizeof(R) == 4)
          for (int k = 0; k < 3; k++) o[9 + 2 * nt + k] = tsl[(size_t)(9 + nt + k) * ne + j];
      }
      o[9 + b] = tsl[(size_t)(9 + b) * ne + j];
      o[9 + nt + b] = (R)swt[p];
    }
  } else if (g == 0) {
    R *o = geom + (size_t)p * 4;
    o[0] = (R)sg[3 * p];
    o[1] = (R)sg[3 * p + 1];
    o[2] = (R)sg[3 * p + 2];
    o[3] = TRI ? (R)swt[p] : (R)0;
  }
  if constexpr (TRI) {
    if (b != 0) return;
    double *o = sigd + ((size_t)g * ne + j) * RB;
    for (int r = 0; r < RB; r++) {
      const int rr = g * RB + r;
      o[r] = rr < nrhs ? scale * x[(size_t)rr * ncols + j] : 0.0;
    }
    return;
  }
  const double f = swt[p] * scale;
  const double e1 = rule.x[b], e2 = rule.y[b];
  R *o = sig + ((size_t)g * npts + p) * RB;
  for (int r = 0; r < RB; r++) {
    const int rr = g * RB + r;
    double d = 0.0;
    if (rr < nrhs) {
      const double *xc = x + (size_t)rr * ncols;
      d = trial_p1 ? (1.0 - e1 - e2) * xc[tri[3 * j]] + e1 * xc[tri[3 * j + 1]] +
                         e2 * xc[tri[3 * j + 2]]
                   : xc[j];
    }
    o[r] = (R)(f * d);
  }
}

// Source strengths.  One lane per source quadrature point (j, b).
template <typename R, int KIND>
__device__ __forceinline__ void fold_point(int idx, int ne, int nt,
                                           const double *__restrict__ sg,
                                           const double *__restrict__ swt,
                                           const double *__restrict__ normals,
                                           const int *__restrict__ tri, int trial_p1,
                                           const RuleArgs &rule,
                                           const double *__restrict__ x, int ncols,
                                           double scale, R *__restrict__ src)
{
  if (idx >= ne * nt) return;
  const int j = idx / nt, b = idx - j * nt;
  constexpr int NC = KIND == 2 ? 3 : 1;
  double dens[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const double *xc = x + (size_t)c * ncols;
    if (trial_p1) {
      const double e1 = rule.x[b], e2 = rule.y[b];
      dens[c] = (1.0 - e1 - e2) * xc[tri[3 * j]] + e1 * xc[tri[3 * j + 1]] +
                e2 * xc[tri[3 * j + 2]];
    } else {
      dens[c] = xc[j];
    }
  }
  const double f = swt[idx] * scale;
  constexpr int REC = Rec<R, KIND>::n;
  R *o = src + (size_t)idx * REC;
  o[0] = (R)sg[3 * idx];
  o[1] = (R)sg[3 * idx + 1];
  o[2] = (R)sg[3 * idx + 2];
  if (KIND == 0) {
    o[3] = (R)(f * dens[0]);
  } else if (KIND == 1 && REC == 12) {
    const double m = f * dens[0];
    for (int k = 0; k < 3; k++) {
      o[3 + k] = (R)(sg[3 * idx + k] - (double)o[k]);
      o[6 + k] = (R)(m * normals[3 * j + k]);
      o[9 + k] = (R)0;
    }
  } else if (KIND == 1) {
    const double m = f * dens[0];
    o[3] = (R)(m * normals[3 * j]);
    o[4] = (R)(m * normals[3 * j + 1]);
    o[5] = (R)(m * normals[3 * j + 2]);
    o[6] = o[7] = (R)0;
  } else {
    o[3] = (R)(f * dens[0]);
    o[4] = (R)(f * dens[NC > 1 ? 1 : 0]);
    o[5] = (R)(f * dens[NC > 2 ? 2 : 0]);
    o[6] = o[7] = (R)0;
  }
}

template <typename R, int KIND>
__global__ void k_fold(int ne, int nt, const double *__restrict__ sg,
                       const double *__restrict__ swt,
                       const double *__restrict__ normals,
                       const int *__restrict__ tri, int trial_p1,
                       RuleArgs rule, const double *__restrict__ x, int ncols,
                       double scale, R *__restrict__ src)
{
  fold_point<R, KIND>(blockIdx.x * blockDim.x + threadIdx.x, ne, nt, sg, swt, normals,
                      tri, trial_p1, rule, x, ncols, scale, src);
}

__device__ __forceinline__ void reduce_one(int k, int n, int chunks,
                                           const double *__restrict__ partial,
                                           double *__restrict__ out)
{
  if (k >= n) return;
  double s = 0.0;
  for (int c = 0; c < chunks; c++) s += partial[(size_t)c * n + k];
  out[k] = s;
}

__global__ void k_reduce(int n, int chunks, const double *__restrict__ partial,
                         double *__restrict__ out)
{
  reduce_one(blockIdx.x * blockDim.x + threadIdx.x, n, chunks, partial, out);
}

// Chunk reduction for many chunks (the pair-symmetric schedule): 32 outputs
// per workgroup, 8 threads per output each summing every 8th chunk, then the
// 8 sums in a fixed order.
__global__ __launch_bounds__(256) void k_reduce_wide(int n, int chunks,
                                                     const double *__restrict__ partial,
                                                     double *__restrict__ out)
{
  __shared__ double red[8][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int k = blockIdx.x * 32 + tx;
  double s = 0.0;
  if (k < n)
    for (int c = ty; c < chunks; c += 8) s += partial[(size_t)c * n + k];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && k < n) {
    double v = red[0][tx];
    for (int y = 1; y < 8; y++) v += red[y][tx];
    out[k] = v;
  }
}

// One pass of the pair-symmetric schedule in passes (block rows m0 ..
// m1-1, partial rows c0 .. of the pass): for every entry e = (plane r,
// output k, triangle t) with t >= t0, slice ty (of 8) adds the pass's
// partial rows c = ty mod 8 in increasing order to acc[ty][e] -- the
// source-side rows c0 .. c1-1 (t beyond the pass's triangles) or c0 ..
// group(t) ks - 1 then the target-side rows group(t) ks .. ng ks - 1 (t of
// the pass's triangles t0 .. t1-1).  Over all passes every entry sees rows
// 0 .. ng ks - 1 in increasing order per slice, as k_reduce_wide sums them,
// so the last pass, which writes the slice totals (in slice order) to out
// for every entry, gives k_reduce_wide's sums bitwise.
struct PassArgs {
  int ne, nplanes, nout, ks, ng, c0, c1, t0, t1, tcols, scols, first, last;
  const double *tb, *sb;
  double *acc, *out;
};

__global__ __launch_bounds__(256) void k_pass_reduce(PassArgs p)
{
  __shared__ double red[8][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int lo = p.last ? 0 : p.t0;  // entries of triangles lo .. ne-1
  const long w = p.ne - lo;
  const long n = (long)p.nplanes * p.nout * w;
  const size_t nall = (size_t)p.nplanes * p.nout * p.ne;
  const long i = (long)blockIdx.x * 32 + tx;
  double s = 0.0;
  size_t e = 0;
  if (i < n) {
    const int rk = (int)(i / w), t = lo + (int)(i - rk * w);
    e = (size_t)rk * p.ne + t;
    if (!p.first) s = p.acc[(size_t)ty * nall + e];
    if (t >= p.t0) {
      const int tl = t - p.t0;
      const bool own = t < p.t1;
      const size_t srow = (size_t)p.nplanes * p.nout * p.scols;
      const double *S = p.sb + (size_t)rk * p.scols + tl;
      const int send = own ? (t >> 6) * p.ks : p.c1;
      int c = p.c0 + (((ty - p.c0) % 8) + 8) % 8;
      for (; c < send; c += 8) s += S[(size_t)(c - p.c0) * srow];
      if (own) {
        const size_t trow = (size_t)p.nplanes * p.nout * p.tcols;
        const double *Tq = p.tb + (size_t)rk * p.tcols + tl;
        for (; c < p.ng * p.ks; c += 8) s += Tq[(size_t)(c - p.c0) * trow];
      }
    }
  }
  if (p.last) {  // kernel-uniform
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && i < n) {
      double v = red[0][tx];
      for (int y = 1; y < 8; y++) v += red[y][tx];
      p.out[e] = v;
    }
  } else if (i < n) {
    p.acc[(size_t)ty * nall + e] = s;
  }
}

// Lanes per test DOF in the near-field pass: every lane of a row takes
// every 8th CSR entry and every 8th (triangle, local node) entry, and a
// fixed butterfly adds the 8 partial sums, so the loads of a wave cover 8
// consecutive entries of 8 rows rather than one entry of 64 rows.
constexpr int kRowLanes = 8;

__device__ __forceinline__ double row_sum(double v)
{
#pragma unroll
  for (int m = 1; m < kRowLanes; m <<= 1) v += __shfl_xor(v, m, kRowLanes);
  return v;
}

// One test DOF of right-hand side rhs (thread gid: row gid / 8, lane
// gid % 8): near CSR row times x plus the far sums gathered from the test
// triangles.
__device__ __forceinline__ void finish_row(int gid, int rhs, int nrows, int ncols, int nc,
                                           int test_p1, int ne,
                                           const int *__restrict__ ptr,
                                           const int *__restrict__ col,
                                           const double *__restrict__ val,
                                           const double *__restrict__ x,
                                           const double *__restrict__ far,
                                           const int *__restrict__ ntp,
                                           const int *__restrict__ nt,
                                           const int *__restrict__ ntk,
                                           double *__restrict__ y)
{
  const int r = gid / kRowLanes, sub = gid % kRowLanes;
  if (r >= nrows) return;  // uniform over the row's lanes
  x += (size_t)rhs * ncols * nc;
  y += (size_t)rhs * nrows * nc;
  far += (size_t)rhs * (test_p1 ? 3 : 1) * nc * ne;
  if (nc == 1) {
    double s = 0.0;
    for (int e = ptr[r] + sub; e < ptr[r + 1]; e += kRowLanes) s = fma(val[e], x[col[e]], s);
    if (test_p1) {
      for (int p = ntp[r] + sub; p < ntp[r + 1]; p += kRowLanes)
        s += far[(size_t)ntk[p] * ne + nt[p]];
    } else if (sub == 0) {
      s += far[r];
    }
    s = row_sum(s);
    if (sub == 0) y[r] = s;
    return;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int e = ptr[r] + sub; e < ptr[r + 1]; e += kRowLanes) {
    const double *B = val + (size_t)e * 6;  // xx yy zz xy xz yz
    const int c = col[e];
    const double x0 = x[c], x1 = x[(size_t)ncols + c], x2 = x[2 * (size_t)ncols + c];
    s0 += B[0] * x0 + B[3] * x1 + B[4] * x2;
    s1 += B[3] * x0 + B[1] * x1 + B[5] * x2;
    s2 += B[4] * x0 + B[5] * x1 + B[2] * x2;
  }
  if (test_p1) {
    for (int p = ntp[r] + sub; p < ntp[r + 1]; p += kRowLanes) {
      const int k = ntk[p], t = nt[p];
      s0 += far[(size_t)(3 * k + 0) * ne + t];
      s1 += far[(size_t)(3 * k + 1) * ne + t];
      s2 += far[(size_t)(3 * k + 2) * ne + t];
    }
  } else if (sub == 0) {
    s0 += far[r];
    s1 += far[(size_t)ne + r];
    s2 += far[2 * (size_t)ne + r];
  }
  s0 = row_sum(s0);
  s1 = row_sum(s1);
  s2 = row_sum(s2);
  if (sub == 0) {
    y[r] = s0;
    y[(size_t)nrows + r] = s1;
    y[2 * (size_t)nrows + r] = s2;
  }
}

// y = near CSR * x + far; 8 lanes per test DOF (launch 8 nrows threads).
// blockIdx.y selects the right-hand side (vectors back to back).
__global__ __launch_bounds__(256) void k_finish(int nrows, int ncols, int nc, int test_p1,
                                                int ne, const int *__restrict__ ptr,
                                                const int *__restrict__ col,
                                                const double *__restrict__ val,
                                                const double *__restrict__ x,
                                                const double *__restrict__ far,
                                                const int *__restrict__ ntp,
                                                const int *__restrict__ nt,
                                                const int *__restrict__ ntk,
                                                double *__restrict__ y)
{
  finish_row(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, nrows, ncols, nc,
             test_p1, ne, ptr, col, val, x, far, ntp, nt, ntk, y);
}

// workgroups of 256 threads for the near-field pass over nrows rows
inline unsigned finish_blocks(int nrows)
{
  return (unsigned)(((size_t)nrows * kRowLanes + 255) / 256);
}

static_assert(kFinRows * kRowLanes == 256, "one fused-pass row block per workgroup");

// The pair-symmetric schedule's partial sums and the near-field pass in one
// launch, for right-hand side plane rhs of nplanes (prow values per row and
// plane).  Workgroup b owns test DOF rows r0 .. r1-1: with P1 test
// functions node row block b of the mesh, whose incidences' far sums the far
// kernel stored from padded position fpos[b] (a multiple of kFinPos) of
// every partial row; with P0 test functions the kFinRows triangles from
// position r0 = b kFinRows.  Stage 1 sums each position over the partial
// rows in k_reduce_wide's order (8 slices of every 8th row, then the slices
// in order) into LDS, kFinPos aligned positions a pass; stage 2 is
// finish_row with those sums, so the result is bitwise that of
// k_reduce_wide + k_finish.
__device__ __forceinline__ void reduce_finish(int b, int rhs, int nplanes, int prow,
                                              int nrows, int ncols, int test_p1, int chunks,
                                              const int *__restrict__ blk,
                                              const int *__restrict__ fpos,
                                              const int *__restrict__ ptr,
                                              const int *__restrict__ col,
                                              const double *__restrict__ val,
                                              const double *__restrict__ x,
                                              const double *__restrict__ partial,
                                              const int *__restrict__ ntp,
                                              double *__restrict__ y,
                                              double (*red)[kFinPos + 1], double *farv)
{
  const int r0 = test_p1 ? blk[b] : b * kFinRows;
  const int r1 = test_p1 ? blk[b + 1] : min(r0 + kFinRows, nrows);
  const int p0 = test_p1 ? ntp[r0] : r0, p1 = test_p1 ? ntp[r1] : r1;
  const int base = test_p1 ? fpos[b] : r0;  // padded position of p0
  const size_t stride = (size_t)nplanes * prow;  // one partial row
  const double *src = partial + (size_t)rhs * prow + (base - p0);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q0 = p0; q0 < p1; q0 += kFinPos) {  // workgroup-uniform
    const int p = q0 + tx;
    double s = 0.0;
    if (p < p1) {
      const double *v = src + p;
      int c = ty;
      // 8 loads in flight, added in row order
      for (; c + 56 < chunks; c += 64) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = v[(size_t)(c + 8 * u) * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) s += t[u];
      }
      for (; c < chunks; c += 8) s += v[(size_t)c * stride];
    }
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && p < p1) {
      double w = red[0][tx];
#pragma unroll
      for (int k = 1; k < 8; k++) w += red[k][tx];
      farv[p - p0] = w;
    }
    __syncthreads();
  }
  const int r = r0 + (int)threadIdx.x / kRowLanes, sub = (int)threadIdx.x % kRowLanes;
  if (r >= r1) return;  // uniform over the row's lanes
  x += (size_t)rhs * ncols;
  y += (size_t)rhs * nrows;
  double s = 0.0;
  for (int e = ptr[r] + sub; e < ptr[r + 1]; e += kRowLanes) s = fma(val[e], x[col[e]], s);
  if (test_p1) {
    for (int p = ntp[r] + sub; p < ntp[r + 1]; p += kRowLanes) s += farv[p - p0];
  } else if (sub == 0) {
    s += farv[r - p0];
  }
  s = row_sum(s);
  if (sub == 0) y[r] = s;
}

// y = near CSR * x + the summed far partials (one component); blockIdx.y
// selects the right-hand side plane of nplanes
__global__ __launch_bounds__(256) void k_reduce_finish(int nplanes, int prow, int nrows,
                                                       int ncols, int test_p1, int chunks,
                                                       const int *__restrict__ blk,
                                                       const int *__restrict__ fpos,
                                                       const int *__restrict__ ptr,
                                                       const int *__restrict__ col,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ partial,
                                                       const int *__restrict__ ntp,
                                                       double *__restrict__ y)
{
  __shared__ double red[8][kFinPos + 1];
  __shared__ double farv[kFinMax];
  reduce_finish((int)blockIdx.x, (int)blockIdx.y, nplanes, prow, nrows, ncols, test_p1,
                chunks, blk, fpos, ptr, col, val, x, partial, ntp, y, red, farv);
}

// values per partial row and right-hand side of the fused pass
inline int fused_prow(const ebh_op *op)
{
  return op->d.test_basis ? op->mesh->nfpos : op->mesh->T.ne;
}

// workgroups of the fused pass
inline unsigned fused_blocks(const ebh_op *op)
{
  return op->d.test_basis ? (unsigned)op->mesh->nfin
                          : (unsigned)((op->nrows + kFinRows - 1) / kFinRows);
}

// the fused pass of an operator (nplanes right-hand sides back to back)
int launch_reduce_finish(const ebh_op *op, const double *x, double *y, int nplanes,
                         hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce_finish, dim3(fused_blocks(op), nplanes), dim3(256), 0, s,
                     nplanes, fused_prow(op), op->nrows, op->ncols, op->d.test_basis,
                     op->chunks, op->mesh->fin_blk.p, op->mesh->fin_pos.p, op->near_ptr.p,
                     op->near_col.p, op->near_val.p, x, op->partial.p,
                     op->mesh->node_tri_ptr.p, y);
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

template <typename R, int KIND>
int launch_far(int nt, dim3 grid, dim3 blk, size_t lds, hipStream_t s,
               const FarArgs<R> &a)
{
  switch (nt) {
  case 1: hipLaunchKernelGGL((k_far<R, KIND, 1>), grid, blk, lds, s, a); break;
  case 3: hipLaunchKernelGGL((k_far<R, KIND, 3>), grid, blk, lds, s, a); break;
  case 4: hipLaunchKernelGGL((k_far<R, KIND, 4>), grid, blk, lds, s, a); break;
  case 6: hipLaunchKernelGGL((k_far<R, KIND, 6>), grid, blk, lds, s, a); break;
  case 7: hipLaunchKernelGGL((k_far<R, KIND, 7>), grid, blk, lds, s, a); break;
  case 12: hipLaunchKernelGGL((k_far<R, KIND, 12>), grid, blk, lds, s, a); break;
  case 13: hipLaunchKernelGGL((k_far<R, KIND, 13>), grid, blk, lds, s, a); break;
  default:
    set_error("unsupported far-field rule size " + std::to_string(nt));
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

template <typename R>
void fill_far_args(const ebh_op *op, FarArgs<R> &a)
{
  const int ne = op->mesh->T.ne, nt = op->rule.nt;
  a.ne = ne;
  a.chunks = op->chunks;
  a.chunk = op->chunk;
  a.tile = op->tile;
  a.ksplit = op->sym ? sym_slots(op) : 1;
  a.nout = op->d.test_basis ? 3 : 1;
  // P0 test and trial functions of the single layer on the symmetric
  // schedule: one kernel sum per triangle pair (rules of several points)
  a.pf = op->sym && op->d.kernel == EBH_LAPLACE_SL && !op->d.test_basis &&
         !op->d.trial_basis && nt > 1 ? 1 : 0;
  a.kappa = (R)(3.0 - 4.0 * op->d.nu);
  if constexpr (sizeof(R) == 4) {
    a.tq = op->tq32.p;
    a.tc = op->tc32.p;
    a.tsl = op->tsl32.p;
    a.src = op->src32.p;
  } else {
    a.tq = op->tq64.p;
    a.tc = nullptr;
    a.tsl = op->tsl64.p;
    a.src = op->src64.p;
  }
  a.touch_ptr = op->mesh->touch_ptr.p;
  a.touch = op->mesh->touch.p;
  a.tw = op->tw.p;
  a.tn = op->normals.p;
  a.partial = op->partial.p;
  const bool fpos = fused_pass(op) && op->d.test_basis;
  a.pos = fpos ? op->mesh->tri_pos.p : nullptr;
  a.prow = fpos ? op->mesh->nfpos : a.nout * ne;
  a.m0 = a.t0 = a.c0 = 0;
  a.tcols = a.scols = ne;
  a.spart = a.partial;
  a.sprow = a.prow;
  for (int q = 0; q < 16; q++) {
    const bool in = q < nt;
    a.rc[q] = (R)op->ray_c[q];
    a.phi[q][0] = in ? 1.0 - op->rule.x[q] - op->rule.y[q] : 0.0;
    a.phi[q][1] = in ? op->rule.x[q] : 0.0;
    a.phi[q][2] = in ? op->rule.y[q] : 0.0;
  }
  // kappa of the symmetric kernel for median orbits o, o' (nt <= 7: at most
  // two orbits of three points after the centroid, if any)
  const int c0 = nt % 3, norb = (nt - c0) / 3;
  for (int o = 0; o < 2; o++)
    for (int p = 0; p < 2; p++) {
      const double co = o < norb ? op->ray_c[c0 + 3 * o] : 0.0;
      const double cp = p < norb ? op->ray_c[c0 + 3 * p] : 0.0;
      a.kap[2 * o + p] = (R)(-0.5 * co * cp * (sizeof(R) == 8 ? 4.0 : 1.0));
    }
}

// LDS bytes of one source tile
size_t far_lds(const ebh_op *op, int R_bytes)
{
  const int k = op->d.kernel;
  const int rec = k == EBH_LAPLACE_SL || k == EBH_LAPLACE_ADL
                      ? 4 : (k == EBH_LAPLACE_DL && R_bytes == 4 ? 12 : 8);
  return (size_t)op->tile * op->rule.nt * rec * R_bytes;
}

template <typename R, int NR, int DK = 0>
int launch_sym(int nt, dim3 grid, hipStream_t s, const FarArgs<R> &a)
{
  const dim3 blk(256);
  switch (nt) {
  case 1:
    // (the P0-factored form is for rules of several points: fill_far_args
    // sets pf only there)
    hipLaunchKernelGGL((k_far_sym<R, 1, NR, DK == 4 ? 0 : DK>), grid, blk, 0, s, a);
    break;
  case 3: hipLaunchKernelGGL((k_far_sym<R, 3, NR, DK>), grid, blk, 0, s, a); break;
  case 4: hipLaunchKernelGGL((k_far_sym<R, 4, NR, DK>), grid, blk, 0, s, a); break;
  case 6: hipLaunchKernelGGL((k_far_sym<R, 6, NR, DK>), grid, blk, 0, s, a); break;
  case 7: hipLaunchKernelGGL((k_far_sym<R, 7, NR, DK>), grid, blk, 0, s, a); break;
  default:
    set_error("no pair-symmetric kernel for rule size " + std::to_string(nt));
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

// nr: right-hand sides per launch (1, or 3 with the 8-wide source records).
// In passes (op->pass_m) each pass's launch is followed by its reduction,
// the last writing the group sums to tri_out.
template <typename R>
int run_far_sym(ebh_op *op, hipStream_t s, const FarArgs<R> &a, int nr = 1)
{
  const int ks = sym_slots(op), ng = op->chunks / ks;
  if (op->d.kernel == EBH_KELVIN_SL) nr = 3;  // the three components
  auto launch = [&](dim3 grid, const FarArgs<R> &b) -> int {
    if (op->d.kernel == EBH_KELVIN_SL) {
      switch (op->rule.nt) {
      case 1: hipLaunchKernelGGL((k_far_sym_kel<R, 1>), grid, dim3(256), 0, s, b); break;
      case 3: hipLaunchKernelGGL((k_far_sym_kel<R, 3>), grid, dim3(256), 0, s, b); break;
      case 4: hipLaunchKernelGGL((k_far_sym_kel<R, 4>), grid, dim3(256), 0, s, b); break;
      case 6: hipLaunchKernelGGL((k_far_sym_kel<R, 6>), grid, dim3(256), 0, s, b); break;
      case 7: hipLaunchKernelGGL((k_far_sym_kel<R, 7>), grid, dim3(256), 0, s, b); break;
      default:
        set_error("no pair-symmetric Kelvin kernel for rule size " +
                  std::to_string(op->rule.nt));
        return EBH_ERR_UNSUPPORTED;
      }
      EBH_HIP(hipGetLastError());
      return EBH_OK;
    }
    if (op->d.kernel == EBH_LAPLACE_DL || op->d.kernel == EBH_LAPLACE_ADL)
      return op->d.kernel == EBH_LAPLACE_DL ? launch_sym<R, 1, 1>(op->rule.nt, grid, s, b)
                                            : launch_sym<R, 1, 2>(op->rule.nt, grid, s, b);
    if (b.pf)
      return nr == 3 ? launch_sym<R, 3, 4>(op->rule.nt, grid, s, b)
                     : launch_sym<R, 1, 4>(op->rule.nt, grid, s, b);
    return nr == 3 ? launch_sym<R, 3>(op->rule.nt, grid, s, b)
                   : launch_sym<R, 1>(op->rule.nt, grid, s, b);
  };
  if (op->pass_m.size() <= 2) {
    int nblk = 0;
    for (int M = 0; 4 * M < ng; M++) nblk += ng - 4 * M;
    return launch(dim3((unsigned)(nblk * ks)), a);
  }
  const int ne = op->mesh->T.ne, nout = a.nout, npass = (int)op->pass_m.size() - 1;
  double *tb = op->partial.p, *sb = tb + op->pass_tsz * op->nr_max,
         *acc = sb + op->pass_ssz * op->nr_max;
  for (int p = 0; p < npass; p++) {
    const int m0 = op->pass_m[p], m1 = op->pass_m[p + 1];
    FarArgs<R> b = a;
    b.m0 = m0;
    b.t0 = 256 * m0;
    b.c0 = 4 * m0 * ks;
    b.tcols = std::min(256 * (m1 - m0), ne - b.t0);
    b.scols = ne - b.t0;
    b.partial = tb;
    b.spart = sb;
    b.prow = nout * b.tcols;
    b.sprow = nout * b.scols;
    b.pos = nullptr;
    int nblk = 0;
    for (int M = m0; M < m1; M++) nblk += ng - 4 * M;
    EBH_TRY(launch(dim3((unsigned)(nblk * ks)), b));
    PassArgs q{};
    q.ne = ne;
    q.nplanes = nr;
    q.nout = nout;
    q.ks = ks;
    q.ng = ng;
    q.c0 = b.c0;
    q.c1 = std::min(4 * m1, ng) * ks;
    q.t0 = b.t0;
    q.t1 = std::min(256 * m1, ne);
    q.tcols = b.tcols;
    q.scols = b.scols;
    q.first = p == 0;
    q.last = p == npass - 1;
    q.tb = tb;
    q.sb = sb;
    q.acc = acc;
    q.out = op->tri_out.p;
    const long n = (long)nr * nout * (q.last ? ne : ne - b.t0);
    hipLaunchKernelGGL(k_pass_reduce, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, s, q);
    EBH_HIP(hipGetLastError());
  }
  return EBH_OK;
}

// source triangles per LDS tile of k_far_pt (at most 64, the mask width, and
// 32 KiB of records)
int pt_tile(const ebh_op *op, int R_bytes)
{
  const int k = op->d.kernel;
  const int rec = k == EBH_LAPLACE_SL || k == EBH_LAPLACE_ADL
                      ? 4 : (k == EBH_LAPLACE_DL && R_bytes == 4 ? 12 : 8);
  const int per = op->rule.nt * rec * R_bytes;
  return std::max(1, std::min(64, 32768 / per));
}

template <typename R>
int run_far_pt(ebh_op *op, hipStream_t s)
{
  const int ne = op->mesh->T.ne, nt = op->rule.nt;
  FarPtArgs<R> a;
  a.ne = ne;
  a.nt = nt;
  a.chunks = op->chunks;
  a.chunk = op->chunk;
  a.ts = pt_tile(op, (int)sizeof(R));
  a.nout = op->d.test_basis ? 3 : 1;
  a.kappa = (R)(3.0 - 4.0 * op->d.nu);
  if constexpr (sizeof(R) == 4) {
    a.tq = op->tq32.p;
    a.tc = op->tc32.p;
    a.src = op->src32.p;
  } else {
    a.tq = op->tq64.p;
    a.tc = nullptr;
    a.src = op->src64.p;
  }
  a.tn = op->normals.p;
  a.touch_ptr = op->mesh->touch_ptr.p;
  a.touch = op->mesh->touch.p;
  a.tw = op->tw.p;
  a.phi = op->rphi.p;
  a.partial = op->partial.p;
  const int k = op->d.kernel;
  const int rec = k == EBH_LAPLACE_SL || k == EBH_LAPLACE_ADL
                      ? 4 : (k == EBH_LAPLACE_DL && sizeof(R) == 4 ? 12 : 8);
  const size_t lds = std::max((size_t)a.ts * nt * rec * sizeof(R), (size_t)256 * 8);
  const int tbn = (ne + 256 / nt - 1) / (256 / nt);
  const dim3 grid((unsigned)(tbn * op->chunks)), blk(256);
  switch (k) {
  case EBH_LAPLACE_SL: hipLaunchKernelGGL((k_far_pt<R, 0>), grid, blk, lds, s, a); break;
  case EBH_LAPLACE_DL: hipLaunchKernelGGL((k_far_pt<R, 1>), grid, blk, lds, s, a); break;
  case EBH_LAPLACE_ADL: hipLaunchKernelGGL((k_far_pt<R, 3>), grid, blk, lds, s, a); break;
  default: hipLaunchKernelGGL((k_far_pt<R, 2>), grid, blk, lds, s, a); break;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

template <typename R>
int run_far(ebh_op *op, hipStream_t s)
{
  if (!reg_rule(op->rule.nt)) return run_far_pt<R>(op, s);
  FarArgs<R> a;
  fill_far_args(op, a);
  if (op->sym) return run_far_sym<R>(op, s, a);
  const int ne = op->mesh->T.ne, nt = op->rule.nt;
  const int tb = (ne + op->block - 1) / op->block;
  dim3 grid((unsigned)(tb * op->chunks)), blk(op->block);
  const int k = op->d.kernel;
  const size_t lds = far_lds(op, (int)sizeof(R));
  switch (k) {
  case EBH_LAPLACE_SL: return launch_far<R, 0>(nt, grid, blk, lds, s, a);
  case EBH_LAPLACE_DL: return launch_far<R, 1>(nt, grid, blk, lds, s, a);
  case EBH_LAPLACE_ADL: return launch_far<R, 3>(nt, grid, blk, lds, s, a);
  default: return launch_far<R, 2>(nt, grid, blk, lds, s, a);
  }
}

template <typename R>
int run_fold(ebh_op *op, const double *x, hipStream_t s)
{
  const int ne = op->mesh->T.ne, nt = op->rule.nt;
  const int n = ne * nt;
  R *src;
  if constexpr (sizeof(R) == 4) src = op->src32.p; else src = op->src64.p;
  const double scale = op->d.kernel == EBH_KELVIN_SL
                           ? (1.0 + op->d.nu) / (8.0 * M_PI * (1.0 - op->d.nu))
                           : kInv4Pi;
  dim3 grid((n + 255) / 256), blk(256);
  const int p1 = op->d.trial_basis;
  // the adjoint double layer and the symmetric double layer take point
  // strengths, as the single layer
  const int kind = op->sym && op->d.kernel == EBH_LAPLACE_DL ? EBH_LAPLACE_SL : op->d.kernel;
  switch (kind) {
  case EBH_LAPLACE_SL:
  case EBH_LAPLACE_ADL:
    hipLaunchKernelGGL((k_fold<R, 0>), grid, blk, 0, s, ne, nt, op->sg64.p,
                       op->swt.p, op->normals.p, op->mesh->tri.p, p1, op->rule,
                       x, op->ncols, scale, src);
    break;
  case EBH_LAPLACE_DL:
    hipLaunchKernelGGL((k_fold<R, 1>), grid, blk, 0, s, ne, nt, op->sg64.p,
                       op->swt.p, op->normals.p, op->mesh->tri.p, p1, op->rule,
                       x, op->ncols, scale, src);
    break;
  default:
    hipLaunchKernelGGL((k_fold<R, 2>), grid, blk, 0, s, ne, nt, op->sg64.p,
                       op->swt.p, op->normals.p, op->mesh->tri.p, p1, op->rule,
                       x, op->ncols, scale, src);
    break;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

// the triangle-pair multi-RHS kernel serves P0 test and trial functions with
// rules of at most 7 points (its target points live in registers)
bool mrhs_tri(const ebh_op *op)
{
  return !op->d.test_basis && !op->d.trial_basis && op->rule.nt <= 7;
}

// record width of the median form of the triangle-pair sums (rules of 6 or
// 7 points on the medians; the record fits the 4 values per point of
// mgeom), else 0.  FP64 k_far_mrhs_tri keeps the point form: there the
// median form's registers, beside its 64 accumulators, cost occupancy
// (measured 15% slower at L6); k_mrhs_tiles holds no accumulators.
int mrhs_med(const ebh_op *op)
{
  const int nt = op->rule.nt;
  if (!mrhs_tri(op) || !op->ray || nt < 6) return 0;
  if (op->d.far_precision != EBH_FP32 && !op->mts) return 0;
  return nt == 6 ? MedRec<6>::W : MedRec<7>::W;
}

template <typename R, int RB>
int launch_mrhs_tri(int nt, bool med, dim3 grid, hipStream_t s, const FarMArgs<R> &a)
{
  switch (nt) {
  case 1: hipLaunchKernelGGL((k_far_mrhs_tri<R, 1, RB, false>), grid, dim3(256), 0, s, a); break;
  case 3: hipLaunchKernelGGL((k_far_mrhs_tri<R, 3, RB, false>), grid, dim3(256), 0, s, a); break;
  case 4: hipLaunchKernelGGL((k_far_mrhs_tri<R, 4, RB, false>), grid, dim3(256), 0, s, a); break;
  case 6:
    if constexpr (sizeof(R) == 4)
      if (med) {
        hipLaunchKernelGGL((k_far_mrhs_tri<R, 6, RB, true>), grid, dim3(256), 0, s, a);
        break;
      }
    hipLaunchKernelGGL((k_far_mrhs_tri<R, 6, RB, false>), grid, dim3(256), 0, s, a);
    break;
  case 7:
    if constexpr (sizeof(R) == 4)
      if (med) {
        hipLaunchKernelGGL((k_far_mrhs_tri<R, 7, RB, true>), grid, dim3(256), 0, s, a);
        break;
      }
    hipLaunchKernelGGL((k_far_mrhs_tri<R, 7, RB, false>), grid, dim3(256), 0, s, a);
    break;
  default:
    set_error("unsupported far-field rule size " + std::to_string(nt) +
              " for the P0 multi-RHS kernel");
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

// kernels of the pair-symmetric multi-RHS schedule
template <typename R>
int launch_mt_tiles(int nt, bool med, dim3 grid, hipStream_t s, const FarMArgs<R> &a,
                    const MtArgs &m, R *tiles)
{
  switch (nt) {
  case 1: hipLaunchKernelGGL((k_mrhs_tiles<R, 1, false>), grid, dim3(256), 0, s, a, m, tiles); break;
  case 3: hipLaunchKernelGGL((k_mrhs_tiles<R, 3, false>), grid, dim3(256), 0, s, a, m, tiles); break;
  case 4: hipLaunchKernelGGL((k_mrhs_tiles<R, 4, false>), grid, dim3(256), 0, s, a, m, tiles); break;
  case 6:
    if (med) hipLaunchKernelGGL((k_mrhs_tiles<R, 6, true>), grid, dim3(256), 0, s, a, m, tiles);
    else hipLaunchKernelGGL((k_mrhs_tiles<R, 6, false>), grid, dim3(256), 0, s, a, m, tiles);
    break;
  case 7:
    if (med) hipLaunchKernelGGL((k_mrhs_tiles<R, 7, true>), grid, dim3(256), 0, s, a, m, tiles);
    else hipLaunchKernelGGL((k_mrhs_tiles<R, 7, false>), grid, dim3(256), 0, s, a, m, tiles);
    break;
  default:
    set_error("unsupported far-field rule size " + std::to_string(nt) +
              " for the pair-symmetric multi-RHS schedule");
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

template <typename R, int RB>
int launch_mt_apply(dim3 ga, dim3 gr, hipStream_t s, const FarMArgs<R> &a, const MtArgs &m,
                    const R *tiles, double *out)
{
  hipLaunchKernelGGL((k_mrhs_tile_mfma<R, RB>), ga, dim3(256), 0, s, a, m, tiles);
  EBH_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_mrhs_tile_reduce<RB>), gr, dim3(256), 0, s, m, a.ne, a.nrhs, out);
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

// Arguments of pass p of the pair-symmetric multi-RHS schedule.
MtArgs mt_args(const ebh_op *op, int p, int groups)
{
  MtArgs m;
  m.ng = (op->mesh->T.ne + 63) / 64;
  m.I0 = op->mt_rows[p];
  m.I1 = op->mt_rows[p + 1];
  m.rowoff = op->mt_rowoff.p;
  m.t0 = op->mt_t0[p];
  m.ntiles = op->mt_t0[p + 1] - m.t0;
  m.items = op->mt_ioff[p + 1] - op->mt_ioff[p];
  m.groups = groups;
  m.first = p == 0;
  m.item = reinterpret_cast<const int4 *>(op->mt_item.p) + op->mt_ioff[p];
  m.order = op->mt_order.p + op->mt_ioff[p];
  m.kptr = op->mt_kptr.p + op->mt_koff[p];
  m.part = op->mt_part.p;
  return m;
}

// The passes of the pair-symmetric multi-RHS schedule: per pass the tiles,
// the items and their reduction into tri_out ([nrhs][ne] far sums).  With
// resident tiles (one pass, formed when the operator was configured) only
// the products and the reduction run.
template <typename R>
int run_mrhs_tiles(ebh_op *op, const FarMArgs<R> &a, bool med, hipStream_t s)
{
  const int ne = a.ne, ng = (ne + 63) / 64, rb = op->rb;
  R *tiles;
  if constexpr (sizeof(R) == 4) tiles = op->mt_tile32.p;
  else tiles = op->mt_tile64.p;
  const int np = (int)op->mt_rows.size() - 1;
  for (int p = 0; p < np; p++) {
    const MtArgs m = mt_args(op, p, a.groups);
    // far-kernel profiling records the tile launches (the VALU-bound kernel)
    // and the matrix-core products and reductions follow outside the pairs;
    // with resident tiles it records the products and reductions
    const bool prof = op->prof_on && 2 * (op->prof_count + 1) <= (int)op->prof_ev.size();
    if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count], s));
    if (!op->mt_res)
      EBH_TRY(launch_mt_tiles<R>(op->rule.nt, med, dim3((m.ntiles + 3) / 4), s, a, m, tiles));
    if (prof && !op->mt_res) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count++ + 1], s));
    const dim3 ga((m.items + 3) / 4, a.groups), gr((ng - m.I0) * (rb / 4), a.groups);
    switch (rb) {
    case 64: EBH_TRY((launch_mt_apply<R, 64>(ga, gr, s, a, m, tiles, op->tri_out.p))); break;
    case 8: EBH_TRY((launch_mt_apply<R, 8>(ga, gr, s, a, m, tiles, op->tri_out.p))); break;
    case 16: EBH_TRY((launch_mt_apply<R, 16>(ga, gr, s, a, m, tiles, op->tri_out.p))); break;
    default: EBH_TRY((launch_mt_apply<R, 32>(ga, gr, s, a, m, tiles, op->tri_out.p))); break;
    }
    if (prof && op->mt_res) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count++ + 1], s));
  }
  return EBH_OK;
}

template <typename R, int RB>
int launch_mrhs_nt(int nt, dim3 grid, hipStream_t s, const FarMArgs<R> &a)
{
  switch (nt) {
  case 1: hipLaunchKernelGGL((k_far_mrhs<R, 1, RB>), grid, dim3(256), 0, s, a); break;
  case 3: hipLaunchKernelGGL((k_far_mrhs<R, 3, RB>), grid, dim3(256), 0, s, a); break;
  case 4: hipLaunchKernelGGL((k_far_mrhs<R, 4, RB>), grid, dim3(256), 0, s, a); break;
  case 6: hipLaunchKernelGGL((k_far_mrhs<R, 6, RB>), grid, dim3(256), 0, s, a); break;
  case 7: hipLaunchKernelGGL((k_far_mrhs<R, 7, RB>), grid, dim3(256), 0, s, a); break;
  case 12: hipLaunchKernelGGL((k_far_mrhs<R, 12, RB>), grid, dim3(256), 0, s, a); break;
  case 13: hipLaunchKernelGGL((k_far_mrhs<R, 13, RB>), grid, dim3(256), 0, s, a); break;
  default:
    set_error("unsupported far-field rule size " + std::to_string(nt));
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

// The fold of the multi-RHS kernels (source records, strengths of the nrhs
// right-hand sides; nrhs 0 writes the records and zero strengths and reads
// no x) and their arguments.
template <typename R>
int mrhs_setup(ebh_op *op, const double *x, int nrhs, hipStream_t s, FarMArgs<R> &a)
{
  const int ne = op->mesh->T.ne, nt = op->rule.nt, rb = op->rb;
  const int groups = std::max(1, (nrhs + rb - 1) / rb);
  R *geom, *sig;
  if constexpr (sizeof(R) == 4) { geom = op->mgeom32.p; sig = op->msig32.p; }
  else { geom = op->mgeom64.p; sig = op->msig64.p; }
  const int n = ne * nt * groups;
  const int p1 = op->d.trial_basis;
  const bool tri = mrhs_tri(op);
  const int med = mrhs_med(op);
  const R *tsl = sizeof(R) == 4 ? (const R *)op->tsl32.p : (const R *)op->tsl64.p;
  const dim3 fg((n + 255) / 256);
#define EBH_FOLD(RB_, TRI_)                                                               \
  hipLaunchKernelGGL((k_fold_mrhs<R, RB_, TRI_>), fg, dim3(256), 0, s, ne, nt, nrhs,    \
                     groups, op->sg64.p, op->swt.p, op->mesh->tri.p, p1, op->rule, x,   \
                     op->ncols, kInv4Pi, geom, sig, op->msig64.p, tsl, med)
  switch (rb) {
  case 64: EBH_FOLD(64, true); break;
  case 8: if (tri) EBH_FOLD(8, true); else EBH_FOLD(8, false); break;
  case 16: if (tri) EBH_FOLD(16, true); else EBH_FOLD(16, false); break;
  default: if (tri) EBH_FOLD(32, true); else EBH_FOLD(32, false); break;
  }
#undef EBH_FOLD
  EBH_HIP(hipGetLastError());
  a.ne = ne;
  a.chunks = op->mchunks;
  a.chunk = op->mchunk;
  a.nout = op->d.test_basis ? 3 : 1;
  a.groups = groups;
  a.nrhs = nrhs;
  a.tq = sizeof(R) == 4 ? (const R *)op->tq32.p : (const R *)op->tq64.p;
  a.geom = geom;
  a.sig = sig;
  a.sigd = op->msig64.p;
  a.tsl = tsl;
  {
    const int c0 = nt % 3, norb = (nt - c0) / 3;
    for (int q = 0; q < 16; q++) a.rc[q] = (R)op->ray_c[q];
    for (int o = 0; o < 2; o++)
      for (int p = 0; p < 2; p++) {
        const double co = o < norb ? op->ray_c[c0 + 3 * o] : 0.0;
        const double cp = p < norb ? op->ray_c[c0 + 3 * p] : 0.0;
        a.kap[2 * o + p] = (R)(-0.5 * co * cp);
      }
  }
  a.touch_ptr = op->mesh->touch_ptr.p;
  a.touch = op->mesh->touch.p;
  a.tw = op->tw.p;
  a.partial = op->partial.p;
  for (int q = 0; q < 16; q++) {
    const bool in = q < nt;
    a.phi[q][0] = in ? 1.0 - op->rule.x[q] - op->rule.y[q] : 0.0;
    a.phi[q][1] = in ? op->rule.x[q] : 0.0;
    a.phi[q][2] = in ? op->rule.y[q] : 0.0;
  }
  return EBH_OK;
}

// Resident tiles: the one pass's tiles formed once, on the operator's
// stream, for every later apply to reuse.
template <typename R>
int build_resident_tiles(ebh_op *op)
{
  FarMArgs<R> a;
  EBH_TRY(mrhs_setup<R>(op, nullptr, 0, op->stream, a));
  R *tiles;
  if constexpr (sizeof(R) == 4) tiles = op->mt_tile32.p;
  else tiles = op->mt_tile64.p;
  const MtArgs m = mt_args(op, 0, a.groups);
  EBH_TRY(launch_mt_tiles<R>(op->rule.nt, mrhs_med(op) > 0, dim3((m.ntiles + 3) / 4),
                             op->stream, a, m, tiles));
  EBH_HIP(hipStreamSynchronize(op->stream));
  op->mt_res = true;
  return EBH_OK;
}

template <typename R>
int run_mrhs(ebh_op *op, const double *x, int nrhs, hipStream_t s)
{
  const int ne = op->mesh->T.ne, nt = op->rule.nt, rb = op->rb;
  const int groups = (nrhs + rb - 1) / rb;
  const int med = mrhs_med(op);
  FarMArgs<R> a;
  EBH_TRY(mrhs_setup<R>(op, x, nrhs, s, a));
  const int tpb = mrhs_tri(op) ? 256 : 256 / nt;
  const int tbn = (ne + tpb - 1) / tpb;
  dim3 grid((unsigned)(tbn * groups * op->mchunks));
  if (op->mts) return run_mrhs_tiles<R>(op, a, med > 0, s);
  const bool prof = op->prof_on && 2 * (op->prof_count + 1) <= (int)op->prof_ev.size();
  if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count], s));
  int rc;
  if (mrhs_tri(op)) {
    switch (rb) {
    case 64: rc = launch_mrhs_tri<R, 64>(nt, med > 0, grid, s, a); break;
    case 8: rc = launch_mrhs_tri<R, 8>(nt, med > 0, grid, s, a); break;
    case 16: rc = launch_mrhs_tri<R, 16>(nt, med > 0, grid, s, a); break;
    default: rc = launch_mrhs_tri<R, 32>(nt, med > 0, grid, s, a); break;
    }
  } else {
    switch (rb) {
    case 8: rc = launch_mrhs_nt<R, 8>(nt, grid, s, a); break;
    case 16: rc = launch_mrhs_nt<R, 16>(nt, grid, s, a); break;
    default: rc = launch_mrhs_nt<R, 32>(nt, grid, s, a); break;
    }
  }
  if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count++ + 1], s));
  return rc;
}

}  // namespace

int chunks_for_grid(const ebh_op *op, int block, int grid)
{
  const int ne = op->mesh->T.ne;
  // k_far_pt holds 256 / nt target triangles a workgroup
  const int tpb = reg_rule(op->rule.nt) ? block : 256 / op->rule.nt;
  const int tb = (ne + tpb - 1) / tpb;
  int chunks = std::max(1, (grid + tb - 1) / tb);
  if (chunks >= 8) chunks = (chunks + 7) / 8 * 8;
  return std::min(chunks, ne);
}

// step ranges per group pair without a tuned value: about 640 workgroups
// for small meshes, at most 8 ranges (measured on MI355X: more ranges cost
// more in partial sums than they gain in occupancy)
int sym_default_split(int ng)
{
  int nblk = 0;
  for (int M = 0; 4 * M < ng; M++) nblk += ng - 4 * M;
  int ks = 1;
  while (ks < 8 && nblk * ks * 2 <= 640) ks *= 2;
  return ks;
}

// The symmetric schedule's partial sums take (ne / 64) x nout x ne FP64
// values per step range and right-hand side of a launch in one pass,
// quadratic in the mesh size; above the operator's limit (ebh_op's
// partial_limit, 4 GiB by default) it runs in passes of block rows, whose
// buffers stay within the limit, plus 8 slice sums per output.

bool sym_eligible(const ebh_op *op)
{
  // the single layer with equal bases, or the double layer or its adjoint
  // (their point sums are weighted by the test functions and their
  // strengths by the trial density, so the bases may differ)
  const bool sl = op->d.kernel == EBH_LAPLACE_SL && op->d.test_basis == op->d.trial_basis;
  const bool dl = (op->d.kernel == EBH_LAPLACE_DL || op->d.kernel == EBH_LAPLACE_ADL) &&
                  op->nr_max == 1;
  // the Kelvin single layer with equal bases (its three components are the
  // launch's three right-hand-side planes)
  const bool kel = op->d.kernel == EBH_KELVIN_SL && op->d.test_basis == op->d.trial_basis &&
                   op->nr_max == 1;
  return (sl || dl || kel) && op->ray && op->rule.nt <= 7;
}

bool chunked_ok(const ebh_op *op)
{
  return !(op->d.kernel == EBH_LAPLACE_ADL && op->d.far_precision == EBH_FP32);
}

// Passes of the pair-symmetric schedule for partial sums above the
// operator's limit: block rows (4 target groups each) are taken greedily
// while a pass's target-side buffer (its triangles, partial rows of its
// groups and later) and source-side buffer (its groups' rows, triangles
// from its first on) stay within the limit; one block row at least.
// pass_tsz / pass_ssz: the largest buffers, values per right-hand side.
static void plan_passes(ebh_op *op, size_t nout)
{
  const size_t ne = op->mesh->T.ne, ks = sym_slots(op);
  const size_t ng = op->chunks / ks, nrow = (ng + 3) / 4;
  auto sizes = [&](size_t m0, size_t m1, size_t &tsz, size_t &ssz) {
    const size_t t0 = 256 * m0;
    const size_t tc = std::min(256 * (m1 - m0), ne - t0), sc = ne - t0;
    tsz = (ng - 4 * m0) * ks * nout * tc;
    ssz = (std::min(4 * m1, ng) - 4 * m0) * ks * nout * sc;
  };
  const double per = 8.0 * op->nr_max;
  op->pass_m.assign(1, 0);
  for (size_t m0 = 0; m0 < nrow;) {
    size_t m1 = m0 + 1, t, sv;
    while (m1 < nrow) {
      sizes(m0, m1 + 1, t, sv);
      if ((double)(t + sv) * per > op->partial_limit) break;
      m1++;
    }
    sizes(m0, m1, t, sv);
    op->pass_tsz = std::max(op->pass_tsz, t);
    op->pass_ssz = std::max(op->pass_ssz, sv);
    op->pass_m.push_back((int)m1);
    m0 = m1;
  }
}

// The pair-symmetric multi-RHS schedule: items of at most kMtSeg tiles (a
// wave's share: kMtSeg x 4096 triangle pairs of RB FMAs each)
constexpr int kMtSeg = 32;

bool mts_eligible(const ebh_op *op)
{
  // tile indices (ng (ng + 1) / 2) fit an int
  return op->rb && mrhs_tri(op) && (op->mesh->T.ne + 63) / 64 <= 65000;
}

// Passes of the pair-symmetric multi-RHS schedule (mt_plan): blocks of
// kMtSeg tile rows within the operator's partial limit, column items on the
// same grid, each output's items in increasing source group, so every
// output sees the same sequence of additions whatever the passes and y does
// not depend on the limit.  *resident: every tile and item fits in the
// resident limit, planned as one pass.
static int plan_mrhs_tiles(ebh_op *op, bool *resident)
{
  const bool f32 = op->d.far_precision == EBH_FP32;
  const size_t groups = (op->d.max_rhs + op->rb - 1) / op->rb;
  const double tile_bytes = 4096.0 * (f32 ? 4 : 8);
  const double item_bytes = 8.0 * (double)(groups * op->rb * 64);
  MtPlan p;
  *resident = false;
  if (op->mt_res_limit > 0) {
    mt_plan(op->mesh->T.ne, tile_bytes, item_bytes, op->mt_res_limit, kMtSeg, p);
    const double all = (double)p.rowoff.back() * tile_bytes +
                       (double)(p.item.size() / 4) * item_bytes;
    *resident = p.rows.size() == 2 && all <= op->mt_res_limit;
  }
  if (!*resident)
    mt_plan(op->mesh->T.ne, tile_bytes, item_bytes, op->partial_limit, kMtSeg, p);
  op->mt_rows = p.rows;
  op->mt_t0 = p.t0;
  op->mt_ioff = p.ioff;
  op->mt_koff = p.koff;
  op->mt_tiles = p.max_tiles;
  op->mt_items = p.max_items;
  EBH_TRY(op->mt_rowoff.upload(p.rowoff));
  EBH_TRY(op->mt_item.upload(p.item));
  EBH_TRY(op->mt_kptr.upload(p.kptr));
  EBH_TRY(op->mt_order.upload(p.order));
  if (f32) {
    op->mt_tile64.release();
    EBH_TRY(op->mt_tile32.alloc(op->mt_tiles * 4096));
  } else {
    op->mt_tile32.release();
    EBH_TRY(op->mt_tile64.alloc(op->mt_tiles * 4096));
  }
  return op->mt_part.alloc(op->mt_items * groups * op->rb * 64);
}

int op_configure(ebh_op *op, int block, int tile, int chunks, bool sym)
{
  const int ne = op->mesh->T.ne;
  bool resident = false;
  if (!sym && !chunked_ok(op)) {
    // the chunked kernel forms (y - x).n_x from rounded FP32 positions
    set_error("the FP32 adjoint double layer runs on the pair-symmetric schedule only");
    return EBH_ERR_UNSUPPORTED;
  }
  if (sym) {
    if (!sym_eligible(op)) {
      set_error("the pair-symmetric schedule needs a Laplace or Kelvin single layer with "
                "equal test and trial bases or a Laplace double layer (or its adjoint) "
                "with one right-hand side, and a median-symmetric far rule of at most 7 "
                "points");
      return EBH_ERR_UNSUPPORTED;
    }
    // groups of 64 triangles; chunks (when given) is the number of step
    // ranges per group pair, 1 .. 64, a power of 2
    const int ng = (ne + 63) / 64;
    int ks = chunks;
    if (ks <= 0) ks = sym_default_split(ng);
    if (ks > 64 || (ks & (ks - 1))) {
      set_error("pair-symmetric step split must be a power of 2 up to 64");
      return EBH_ERR_ARG;
    }
    block = 256;
    tile = 64;
    op->ksplit = ks;
    // FP64 Kelvin with 6 or 7 points: two target-point halves per pair
    op->qsplit = op->d.kernel == EBH_KELVIN_SL && op->d.far_precision == EBH_FP64 &&
                         op->rule.nt >= 6 ? 2 : 1;
    chunks = ng * ks * op->qsplit;
  } else {
    op->qsplit = 1;
  }
  op->sym = sym;
  if (block) {
    if (block != 64 && block != 128 && block != 256) {
      set_error("far-field block size must be 64, 128 or 256");
      return EBH_ERR_ARG;
    }
    op->block = block;
  }
  if (tile) {
    if (tile < 1 || tile > 64) {
      set_error("far-field tile must be 1..64 source triangles");
      return EBH_ERR_ARG;
    }
    op->tile = tile;
  }
  // default: ~4096 workgroups, 256 CUs x several resident blocks
  if (chunks <= 0) chunks = chunks_for_grid(op, op->block, 4096);
  if (sym) {
    op->chunk = 64;
    op->chunks = chunks;
  } else {
    chunks = std::min(chunks, ne);
    op->chunk = (ne + chunks - 1) / chunks;
    op->chunks = (ne + op->chunk - 1) / op->chunk;
  }
  const size_t nout = (size_t)(op->d.test_basis ? 3 : 1) * op->nc;
  // values per partial row and right-hand side: the fused pass's padded
  // row with P1 test functions (fused_prow), nout ne otherwise
  size_t prow = nout * ne;
  if (sym && op->d.test_basis) prow = std::max(prow, (size_t)op->mesh->nfpos);
  size_t need = (size_t)op->chunks * prow * (sym ? op->nr_max : 1);
  op->pass_m.clear();
  op->pass_tsz = op->pass_ssz = 0;
  if (sym && (double)need * 8.0 > op->partial_limit) {
    // passes: the largest pass buffers, then the 8 slice sums
    plan_passes(op, nout);
    need = (op->pass_tsz + op->pass_ssz) * op->nr_max + (size_t)8 * op->nr_max * nout * ne;
  }
  if (op->rb) {
    // multi-RHS launch: one lane per target point, 256/nt triangles a block
    // (P0 test and trial: one lane per target triangle, 256 a block)
    const int tpb = mrhs_tri(op) ? 256 : 256 / op->rule.nt;
    const int tbn = (ne + tpb - 1) / tpb;
    const int groups = (op->d.max_rhs + op->rb - 1) / op->rb;
    int mc = std::max(1, (op->mgrid + tbn * groups - 1) / (tbn * groups));
    if (mc >= 8) mc = (mc + 7) / 8 * 8;
    mc = std::min(mc, ne);
    op->mchunk = (ne + mc - 1) / mc;
    op->mchunks = (ne + op->mchunk - 1) / op->mchunk;
    op->mts = op->mrhs_sched != EBH_MRHS_TARGET && mts_eligible(op);
    op->mt_res = false;
    if (op->mts) {
      EBH_TRY(plan_mrhs_tiles(op, &resident));
    } else {
      op->mt_tile32.release();
      op->mt_tile64.release();
      op->mt_part.release();
      need = std::max(need, (size_t)op->mchunks * op->d.max_rhs * nout * ne);
    }
  }
  if (op->sym && op->nr_max == 3) {
    // three right-hand sides per symmetric launch: 8-wide source records,
    // partial sums and chunk sums of 3 RHS
    const size_t npts = (size_t)ne * op->rule.nt;
    if (op->d.far_precision == EBH_FP32) EBH_TRY(op->src3_32.alloc(npts * 8));
    else EBH_TRY(op->src3_64.alloc(npts * 8));
    if (op->tri_out.n < 3 * nout * ne) EBH_TRY(op->tri_out.alloc(3 * nout * ne));
  }
  EBH_TRY(op->partial.alloc(need));
  op->gen++;
  if (resident) {
    if (op->d.far_precision == EBH_FP32) EBH_TRY(build_resident_tiles<float>(op));
    else EBH_TRY(build_resident_tiles<double>(op));
  }
  return EBH_OK;
}

int op_build_far(ebh_op *op)
{
  const Topology &T = op->mesh->T;
  const int ne = T.ne, nt = op->rule.nt;
  std::vector<double> tq((size_t)nt * 3 * ne), sg((size_t)ne * nt * 3),
      tw((size_t)nt * ne), swt((size_t)ne * nt), nrm((size_t)3 * ne);
  for (int t = 0; t < ne; t++) {
    const double *P0 = &T.nodes[3 * T.tri[3 * t]];
    const double *P1 = &T.nodes[3 * T.tri[3 * t + 1]];
    const double *P2 = &T.nodes[3 * T.tri[3 * t + 2]];
    double u[3], v[3];
    for (int k = 0; k < 3; k++) { u[k] = P1[k] - P0[k]; v[k] = P2[k] - P0[k]; }
    const double c0 = u[1] * v[2] - u[2] * v[1], c1 = u[2] * v[0] - u[0] * v[2],
                 c2 = u[0] * v[1] - u[1] * v[0];
    const double J = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    nrm[3 * t] = c0 / J; nrm[3 * t + 1] = c1 / J; nrm[3 * t + 2] = c2 / J;
    for (int q = 0; q < nt; q++) {
      const double e1 = op->rule.x[q], e2 = op->rule.y[q];
      for (int k = 0; k < 3; k++) {
        // x = A + ksi1 u + ksi2 v, src/BEM3dLaplace.cpp:171-173
        const double xk = P0[k] + e1 * u[k] + e2 * v[k];
        tq[(size_t)(3 * q + k) * ne + t] = xk;
        sg[((size_t)t * nt + q) * 3 + k] = xk;
      }
      tw[(size_t)q * ne + t] = op->rule.w[q] * J;
      swt[(size_t)t * nt + q] = op->rule.w[q] * J;
    }
  }
  if (op->d.far_precision == EBH_FP32) {
    std::vector<float> t32(tq.begin(), tq.end());
    if (op->d.kernel == EBH_LAPLACE_DL) {
      // target points as offsets from the centroid; centroid as hi + lo
      std::vector<float> tc((size_t)6 * ne);
      for (int t = 0; t < ne; t++) {
        double c[3] = {0, 0, 0};
        for (int k = 0; k < 3; k++)
          c[k] = (T.nodes[3 * T.tri[3 * t] + k] + T.nodes[3 * T.tri[3 * t + 1] + k] +
                  T.nodes[3 * T.tri[3 * t + 2] + k]) / 3.0;
        for (int k = 0; k < 3; k++) {
          const float hi = (float)c[k];
          tc[(size_t)k * ne + t] = hi;
          tc[(size_t)(3 + k) * ne + t] = (float)(c[k] - (double)hi);
          for (int q = 0; q < nt; q++)
            t32[(size_t)(3 * q + k) * ne + t] = (float)(tq[(size_t)(3 * q + k) * ne + t] - c[k]);
        }
      }
      EBH_TRY(op->tc32.upload(tc));
    }
    EBH_TRY(op->tq32.upload(t32));
    EBH_TRY(op->src32.alloc((size_t)ne * nt *
                            (op->d.kernel == EBH_LAPLACE_SL || op->d.kernel == EBH_LAPLACE_ADL
                                 ? 4 : op->d.kernel == EBH_LAPLACE_DL ? 12 : 8)));
  } else {
    EBH_TRY(op->tq64.upload(tq));
    EBH_TRY(op->src64.alloc((size_t)ne * nt *
                            (op->d.kernel == EBH_LAPLACE_SL || op->d.kernel == EBH_LAPLACE_ADL
                                 ? 4 : 8)));
  }
  if (reg_rule(nt) &&
      (op->d.kernel == EBH_LAPLACE_SL || op->d.kernel == EBH_LAPLACE_ADL ||
       (op->d.kernel == EBH_LAPLACE_DL || op->d.kernel == EBH_KELVIN_SL) && op->ray &&
           nt <= 7)) {
    // single-layer target frame: centroid, then the median form (ray rules)
    // or the offsets of every point from the centroid; the double layer and
    // the Kelvin single layer use it on the symmetric schedule
    const bool ray = op->ray && nt <= 7;
    // FP32 median rules: three more rows, the rounding remainders of the
    // centroid, C - (float)C (the symmetric kernel's hi/lo differences)
    const bool f32 = op->d.far_precision == EBH_FP32;
    const size_t rows = ray ? 9 + nt + (f32 ? 3 : 0) : 3 + 4 * nt;
    std::vector<double> tsl(rows * ne);
    for (int t = 0; t < ne; t++) {
      double c[3];
      for (int k = 0; k < 3; k++)
        c[k] = (T.nodes[3 * T.tri[3 * t] + k] + T.nodes[3 * T.tri[3 * t + 1] + k] +
                T.nodes[3 * T.tri[3 * t + 2] + k]) / 3.0;
      for (int k = 0; k < 3; k++) tsl[(size_t)k * ne + t] = c[k];
      if (ray) {
        double e[3][3], e2[3];
        for (int v = 0; v < 3; v++) {
          e2[v] = 0.0;
          for (int k = 0; k < 3; k++) {
            e[v][k] = T.nodes[3 * T.tri[3 * t + v] + k] - c[k];
            e2[v] += e[v][k] * e[v][k];
          }
        }
        for (int k = 0; k < 3; k++) {
          tsl[(size_t)(3 + k) * ne + t] = 2.0 * e[0][k];
          tsl[(size_t)(6 + k) * ne + t] = 2.0 * e[1][k];
        }
        const int c0 = nt % 3;
        for (int q = 0; q < nt; q++) {
          const double cq = op->ray_c[q];
          tsl[(size_t)(9 + q) * ne + t] = q < c0 ? 0.0 : cq * cq * e2[(q - c0) % 3];
        }
        if (f32)
          for (int k = 0; k < 3; k++)
            tsl[(size_t)(9 + nt + k) * ne + t] = c[k] - (double)(float)c[k];
        continue;
      }
      for (int q = 0; q < nt; q++) {
        double e = 0.0;
        for (int k = 0; k < 3; k++) {
          const double d = tq[(size_t)(3 * q + k) * ne + t] - c[k];
          tsl[(size_t)(3 + 4 * q + k) * ne + t] = 2.0 * d;
          e += d * d;
        }
        tsl[(size_t)(6 + 4 * q) * ne + t] = e;
      }
    }
    if (op->d.far_precision == EBH_FP32) {
      std::vector<float> t32(tsl.begin(), tsl.end());
      EBH_TRY(op->tsl32.upload(t32));
    } else {
      EBH_TRY(op->tsl64.upload(tsl));
    }
  }
  EBH_TRY(op->sg64.upload(sg));
  EBH_TRY(op->tw.upload(tw));
  EBH_TRY(op->swt.upload(swt));
  EBH_TRY(op->normals.upload(nrm));
  const size_t nout = (size_t)(op->d.test_basis ? 3 : 1) * op->nc;
  EBH_TRY(op->tri_out.alloc(nout * ne * (size_t)op->d.max_rhs));
  if (!reg_rule(nt)) {
    // k_far_pt reads the test shape functions at the points from the device
    std::vector<double> phi(3 * (size_t)nt);
    for (int q = 0; q < nt; q++) {
      phi[3 * q] = 1.0 - op->rule.x[q] - op->rule.y[q];
      phi[3 * q + 1] = op->rule.x[q];
      phi[3 * q + 2] = op->rule.y[q];
    }
    EBH_TRY(op->rphi.upload(phi));
  }
  // multi-RHS kernels for the rules a lane holds in registers; other rules
  // apply their right-hand sides one at a time
  if (op->d.kernel == EBH_LAPLACE_SL && op->d.max_rhs > 1 && reg_rule(nt)) {
    // 64 RHS a group for the triangle-pair kernel (each pair's sum once)
    op->rb = op->d.max_rhs >= 64 && mrhs_tri(op) ? 64
             : op->d.max_rhs >= 32                ? 32
             : op->d.max_rhs >= 16                ? 16
                                                  : 8;
    const size_t groups = (op->d.max_rhs + op->rb - 1) / op->rb;
    const size_t npts = (size_t)ne * nt;
    if (op->d.far_precision == EBH_FP32) {
      EBH_TRY(op->mgeom32.alloc(npts * 4));
      // FP64 strengths per triangle for the triangle-pair kernel
      if (mrhs_tri(op)) EBH_TRY(op->msig64.alloc((size_t)ne * groups * op->rb));
      else EBH_TRY(op->msig32.alloc(npts * groups * op->rb));
    } else {
      EBH_TRY(op->mgeom64.alloc(npts * 4));
      EBH_TRY(op->msig64.alloc(npts * groups * op->rb));
    }
  }
  // default: the pair-symmetric schedule where it applies (half the
  // rsqrts), else chunked; a tuned table entry may replace either
  if (sym_eligible(op)) return op_configure(op, 0, 0, 0, true);
  const int block = ne >= 16384 ? 256 : ne >= 4096 ? 128 : 64;
  return op_configure(op, block, 64, 0);
}

// chunk reduction of the far partial sums, then near-field SpMV plus the
// far sums gathered to the test DOFs
static int op_finish(ebh_op *op, const double *x, double *y, hipStream_t s)
{
  const int ne = op->mesh->T.ne;
  const int nred = (op->d.test_basis ? 3 : 1) * op->nc * ne;
  if (fused_pass(op)) {
    const bool nprof = op->prof_on && 2 * (op->prof_ncount + 1) <= (int)op->prof_nev.size();
    if (nprof) EBH_HIP(hipEventRecord(op->prof_nev[2 * op->prof_ncount], s));
    EBH_TRY(launch_reduce_finish(op, x, y, 1, s));
    if (nprof) EBH_HIP(hipEventRecord(op->prof_nev[2 * op->prof_ncount++ + 1], s));
    return EBH_OK;
  }
  if (op->sym && op->pass_m.size() > 2) {
    // the passes' last reduction wrote tri_out
  } else if (op->sym)
    hipLaunchKernelGGL(k_reduce_wide, dim3((nred + 31) / 32), dim3(256), 0, s, nred,
                       op->chunks, op->partial.p, op->tri_out.p);
  else
    hipLaunchKernelGGL(k_reduce, dim3((nred + 255) / 256), dim3(256), 0, s, nred,
                       op->chunks, op->partial.p, op->tri_out.p);
  EBH_HIP(hipGetLastError());
  const bool nprof = op->prof_on && 2 * (op->prof_ncount + 1) <= (int)op->prof_nev.size();
  if (nprof) EBH_HIP(hipEventRecord(op->prof_nev[2 * op->prof_ncount], s));
  hipLaunchKernelGGL(k_finish, dim3(finish_blocks(op->nrows)), dim3(256), 0, s,
                     op->nrows, op->ncols, op->nc, op->d.test_basis, ne,
                     op->near_ptr.p, op->near_col.p, op->near_val.p, x,
                     op->tri_out.p, op->mesh->node_tri_ptr.p,
                     op->mesh->node_tri.p, op->mesh->node_tri_k.p, y);
  EBH_HIP(hipGetLastError());
  if (nprof) EBH_HIP(hipEventRecord(op->prof_nev[2 * op->prof_ncount++ + 1], s));
  return EBH_OK;
}

int op_apply_one(ebh_op *op, const double *x, double *y, hipStream_t s,
                 bool timed)
{
  if (op->d.kernel == EBH_SPARSE_INVERSE) return spai_apply(op, x, y, 1, s, timed);
  if (op->inner) return composite_apply(op, x, y, 1, s, timed);
  if (op->pscale) {
    EBH_TRY(mass_inv_apply(op, x, op->hsw.p, 1, s));
    x = op->hsw.p;
  }
  const int ne = op->mesh->T.ne;
  const bool f32 = op->d.far_precision == EBH_FP32;
  if (timed) EBH_HIP(hipEventRecord(op->ev[0], s));
  EBH_TRY(f32 ? run_fold<float>(op, x, s) : run_fold<double>(op, x, s));
  if (timed) EBH_HIP(hipEventRecord(op->ev[1], s));
  const bool prof = op->prof_on && 2 * (op->prof_count + 1) <= (int)op->prof_ev.size();
  if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count], s));
  EBH_TRY(f32 ? run_far<float>(op, s) : run_far<double>(op, s));
  if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count++ + 1], s));
  if (timed) EBH_HIP(hipEventRecord(op->ev[2], s));
  EBH_TRY(op_finish(op, x, y, s));
  if (op->pscale) EBH_TRY(mass_inv_apply(op, y, y, 1, s));
  if (timed) EBH_HIP(hipEventRecord(op->ev[3], s));
  return EBH_OK;
}

// Three right-hand sides (back to back) in one pair-symmetric launch: the
// three strengths share a source record (the Kelvin fold's layout) and each
// rsqrt serves 3 x 2 accumulations.
template <typename R>
int run_three(ebh_op *op, const double *x, hipStream_t s)
{
  const int ne = op->mesh->T.ne, nt = op->rule.nt;
  const int n = ne * nt;
  R *src;
  if constexpr (sizeof(R) == 4) src = op->src3_32.p; else src = op->src3_64.p;
  hipLaunchKernelGGL((k_fold<R, 2>), dim3((n + 255) / 256), dim3(256), 0, s, ne, nt,
                     op->sg64.p, op->swt.p, op->normals.p, op->mesh->tri.p,
                     op->d.trial_basis, op->rule, x, op->ncols, kInv4Pi, src);
  EBH_HIP(hipGetLastError());
  FarArgs<R> a;
  fill_far_args(op, a);
  a.src = src;
  const bool prof = op->prof_on && 2 * (op->prof_count + 1) <= (int)op->prof_ev.size();
  if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count], s));
  EBH_TRY(run_far_sym<R>(op, s, a, 3));
  if (prof) EBH_HIP(hipEventRecord(op->prof_ev[2 * op->prof_count++ + 1], s));
  return EBH_OK;
}

static int op_apply_three(ebh_op *op, const double *x, double *y, hipStream_t s)
{
  const int ne = op->mesh->T.ne;
  EBH_TRY(op->d.far_precision == EBH_FP32 ? run_three<float>(op, x, s)
                                          : run_three<double>(op, x, s));
  if (fused_pass(op)) return launch_reduce_finish(op, x, y, 3, s);
  const int nred = 3 * (op->d.test_basis ? 3 : 1) * ne;
  if (op->pass_m.size() <= 2) {
    hipLaunchKernelGGL(k_reduce_wide, dim3((nred + 31) / 32), dim3(256), 0, s, nred,
                       op->chunks, op->partial.p, op->tri_out.p);
    EBH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finish, dim3(finish_blocks(op->nrows), 3), dim3(256), 0, s,
                     op->nrows, op->ncols, op->nc, op->d.test_basis, ne,
                     op->near_ptr.p, op->near_col.p, op->near_val.p, x,
                     op->tri_out.p, op->mesh->node_tri_ptr.p, op->mesh->node_tri.p,
                     op->mesh->node_tri_k.p, y);
  EBH_HIP(hipGetLastError());
  return EBH_OK;
}

int op_apply_multi(ebh_op *op, const double *x, double *y, int nrhs,
                   hipStream_t s, bool timed)
{
  if (op->d.kernel == EBH_SPARSE_INVERSE) return spai_apply(op, x, y, nrhs, s, timed);
  if (op->inner) return composite_apply(op, x, y, nrhs, s, timed);
  if (nrhs == 1 || !op->rb || op->pscale) {
    const size_t cx = (size_t)op->ncols * op->nc, cy = (size_t)op->nrows * op->nc;
    int r = 0;
    if (op->sym && op->nr_max == 3 && !op->pscale && !timed)
      for (; r + 3 <= nrhs; r += 3)
        EBH_TRY(op_apply_three(op, x + r * cx, y + r * cy, s));
    for (; r < nrhs; r++)
      EBH_TRY(op_apply_one(op, x + r * cx, y + r * cy, s, timed));
    return EBH_OK;
  }
  if (nrhs > op->d.max_rhs) {
    set_error("nrhs " + std::to_string(nrhs) + " exceeds the operator's max_rhs " +
              std::to_string(op->d.max_rhs));
    return EBH_ERR_ARG;
  }
  const int ne = op->mesh->T.ne;
  const bool f32 = op->d.far_precision == EBH_FP32;
  if (timed) EBH_HIP(hipEventRecord(op->ev[0], s));
  if (timed) EBH_HIP(hipEventRecord(op->ev[1], s));
  EBH_TRY(f32 ? run_mrhs<float>(op, x, nrhs, s) : run_mrhs<double>(op, x, nrhs, s));
  if (timed) EBH_HIP(hipEventRecord(op->ev[2], s));
  if (!op->mts) {
    // (the pair-symmetric schedule's passes wrote the far sums to tri_out)
    const int nred = (op->d.test_basis ? 3 : 1) * ne * nrhs;
    hipLaunchKernelGGL(k_reduce, dim3((nred + 255) / 256), dim3(256), 0, s, nred,
                       op->mchunks, op->partial.p, op->tri_out.p);
    EBH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finish, dim3(finish_blocks(op->nrows), nrhs), dim3(256), 0, s,
                     op->nrows, op->ncols, op->nc, op->d.test_basis, ne,
                     op->near_ptr.p, op->near_col.p, op->near_val.p, x,
                     op->tri_out.p, op->mesh->node_tri_ptr.p,
                     op->mesh->node_tri.p, op->mesh->node_tri_k.p, y);
  EBH_HIP(hipGetLastError());
  if (timed) EBH_HIP(hipEventRecord(op->ev[3], s));
  return EBH_OK;
}

// ---------------------------------------------------- multi-mesh batches

constexpr int kMaxBatch = 64;

// Per-member data of the batched fold / reduce / finish kernels.
struct MemberDesc {
  int ne, nt, trial_p1, ncols, nrows, nc, test_p1, chunks, nred;
  int sym;  // pair-symmetric schedule (group partials, wide reduction)
  int fused, nfin, prow;  // the fused pass (fused_pass), its workgroups and
                          // partial row length
  const int *fblk, *fpos;  // its node row blocks and their padded positions
  double scale;
  const double *sg, *swt, *normals;
  const int *tri;
  void *src;
  const double *partial;
  double *tri_out;
  const int *near_ptr, *near_col;
  const double *near_val;
  const int *ntp, *ntri, *ntk;
};

// Input and output vectors of one batch apply, by value.
struct BatchPtrs {
  const double *x[kMaxBatch];
  double *y[kMaxBatch];
};

template <typename R, int KIND>
__global__ void k_fold_multi(const MemberDesc *__restrict__ d, RuleArgs rule, BatchPtrs p)
{
  const MemberDesc &o = d[blockIdx.y];
  fold_point<R, KIND>(blockIdx.x * blockDim.x + threadIdx.x, o.ne, o.nt, o.sg, o.swt,
                      o.normals, o.tri, o.trial_p1, rule, p.x[blockIdx.y], o.ncols,
                      o.scale, reinterpret_cast<R *>(o.src));
}

// chunk partial sums of every chunked member
__global__ void k_reduce_multi(const MemberDesc *__restrict__ d)
{
  const MemberDesc &o = d[blockIdx.y];
  if (o.sym) return;
  reduce_one(blockIdx.x * blockDim.x + threadIdx.x, o.nred, o.chunks, o.partial,
             o.tri_out);
}

// group partial sums of every symmetric member, in k_reduce_wide's order
__global__ __launch_bounds__(256) void k_reduce_wide_multi(const MemberDesc *__restrict__ d)
{
  const MemberDesc &o = d[blockIdx.y];
  if (!o.sym || o.fused || (int)blockIdx.x * 32 >= o.nred) return;  // workgroup-uniform
  __shared__ double red[8][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int k = blockIdx.x * 32 + tx;
  const int n = o.nred;
  double s = 0.0;
  if (k < n)
    for (int c = ty; c < o.chunks; c += 8) s += o.partial[(size_t)c * n + k];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && k < n) {
    double v = red[0][tx];
    for (int y = 1; y < 8; y++) v += red[y][tx];
    o.tri_out[k] = v;
  }
}

__global__ void k_finish_multi(const MemberDesc *__restrict__ d, BatchPtrs p)
{
  const MemberDesc &o = d[blockIdx.y];
  if (o.fused) return;  // workgroup-uniform
  finish_row(blockIdx.x * blockDim.x + threadIdx.x, 0, o.nrows, o.ncols, o.nc,
             o.test_p1, o.ne, o.near_ptr, o.near_col, o.near_val, p.x[blockIdx.y],
             o.tri_out, o.ntp, o.ntri, o.ntk, p.y[blockIdx.y]);
}

// the fused pass of every member that has it
__global__ __launch_bounds__(256) void k_reduce_finish_multi(const MemberDesc *__restrict__ d,
                                                             BatchPtrs p)
{
  const MemberDesc &o = d[blockIdx.y];
  if (!o.fused || (int)blockIdx.x >= o.nfin) return;  // workgroup-uniform
  __shared__ double red[8][kFinPos + 1];
  __shared__ double farv[kFinMax];
  reduce_finish((int)blockIdx.x, 0, 1, o.prow, o.nrows, o.ncols, o.test_p1, o.chunks, o.fblk,
                o.fpos, o.near_ptr, o.near_col, o.near_val, p.x[blockIdx.y], o.partial,
                o.ntp, p.y[blockIdx.y], red, farv);
}

__global__ void k_zero_int(int n, int *p)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename R, int KIND>
const void *multi_kernel(int nt)
{
  switch (nt) {
  case 1: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 1>);
  case 3: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 3>);
  case 4: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 4>);
  case 6: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 6>);
  case 7: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 7>);
  case 12: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 12>);
  case 13: return reinterpret_cast<const void *>(&k_far_multi<R, KIND, 13>);
  default: return nullptr;
  }
}

template <typename R>
const void *multi_sym_kernel(int nt)
{
  switch (nt) {
  case 1: return reinterpret_cast<const void *>(&k_far_sym_multi<R, 1>);
  case 3: return reinterpret_cast<const void *>(&k_far_sym_multi<R, 3>);
  case 4: return reinterpret_cast<const void *>(&k_far_sym_multi<R, 4>);
  case 6: return reinterpret_cast<const void *>(&k_far_sym_multi<R, 6>);
  case 7: return reinterpret_cast<const void *>(&k_far_sym_multi<R, 7>);
  default: return nullptr;
  }
}

template <typename R>
const void *multi_kernel_for(int kernel, int nt)
{
  switch (kernel) {
  case EBH_LAPLACE_SL: return multi_kernel<R, 0>(nt);
  case EBH_LAPLACE_DL: return multi_kernel<R, 1>(nt);
  case EBH_KELVIN_SL: return multi_kernel<R, 2>(nt);
  case EBH_LAPLACE_ADL: return multi_kernel<R, 3>(nt);
  default: return nullptr;
  }
}

}  // namespace ebh

// Operators of one kind on several meshes, applied with one persistent
// far-field launch per schedule.  Chunked members: target blocks of 256
// triangles, each operator with its own source chunks and LDS tile.
// Symmetric members: each operator's workgroups as queue items.  Every
// result equals the operator's own apply.
struct ebh_batch {
  std::vector<ebh_op *> ops;
  int kernel = 0, prec = 0, nt = 0, nops = 0;
  ebh::DevBuf<unsigned char> desc;  // FarArgs<R>[nops]
  ebh::DevBuf<unsigned char> mdesc; // MemberDesc[nops]
  int max_fold = 0, max_red = 0, max_rows = 0, max_fin = 0;
  ebh::DevBuf<int> off;             // [8][nops + 1] first item per queue
  ebh::DevBuf<int> soff;            // [nops + 1] first symmetric item
  ebh::DevBuf<int> counters;        // [9]: 8 chunked queues, 1 symmetric
  std::vector<int> gens;
  size_t lds = 0;
  int grid = 0, items = 0, sgrid = 0, sitems = 0;
  const void *kern = nullptr, *skern = nullptr;
};

namespace ebh {
namespace {

// the persistent symmetric launch holds single-layer members
bool batch_sym(const ebh_op *op)
{
  return op->d.kernel == EBH_LAPLACE_SL && sym_eligible(op);
}

// Members' far-field schedules for a batch: the symmetric schedule with one
// step range where it applies (the batch supplies the parallelism), else
// one chunk size for all, chosen so that the launch holds about 8 work
// items per resident workgroup.
int batch_shapes(ebh_batch *b, int grid)
{
  double work = 0.0;
  for (const ebh_op *op : b->ops)
    if (!batch_sym(op)) work += (double)((op->mesh->T.ne + 255) / 256) * op->mesh->T.ne;
  int cs = (int)std::ceil(work / (8.0 * std::max(grid, 1)));
  cs = std::max(64, (cs + 63) / 64 * 64);
  for (ebh_op *op : b->ops) {
    const int ne = op->mesh->T.ne;
    // (the persistent launch runs one pass per member)
    if (batch_sym(op) && op_configure(op, 0, 0, 1, true) == EBH_OK && op->pass_m.size() <= 2)
      continue;
    EBH_TRY(op_configure(op, 256, 64, std::max(1, (ne + cs - 1) / cs)));
  }
  return EBH_OK;
}

template <typename R>
int batch_prepare(ebh_batch *b)
{
  const int nops = b->nops;
  std::vector<MemberDesc> md(nops);
  b->max_fold = b->max_red = b->max_rows = b->max_fin = 0;
  for (int k = 0; k < nops; k++) {
    const ebh_op *op = b->ops[k];
    MemberDesc &m = md[k];
    m.ne = op->mesh->T.ne;
    m.nt = op->rule.nt;
    m.trial_p1 = op->d.trial_basis;
    m.ncols = op->ncols;
    m.nrows = op->nrows;
    m.nc = op->nc;
    m.test_p1 = op->d.test_basis;
    m.chunks = op->chunks;
    m.sym = op->sym ? 1 : 0;
    m.fused = fused_pass(op) ? 1 : 0;
    m.nfin = m.fused ? (int)fused_blocks(op) : 0;
    m.fblk = op->mesh->fin_blk.p;
    m.fpos = op->mesh->fin_pos.p;
    m.prow = m.fused ? fused_prow(op) : 0;
    m.nred = (op->d.test_basis ? 3 : 1) * op->nc * m.ne;
    m.scale = op->d.kernel == EBH_KELVIN_SL
                  ? (1.0 + op->d.nu) / (8.0 * M_PI * (1.0 - op->d.nu))
                  : kInv4Pi;
    m.sg = op->sg64.p;
    m.swt = op->swt.p;
    m.normals = op->normals.p;
    m.tri = op->mesh->tri.p;
    if constexpr (sizeof(R) == 4) m.src = op->src32.p; else m.src = op->src64.p;
    m.partial = op->partial.p;
    m.tri_out = op->tri_out.p;
    m.near_ptr = op->near_ptr.p;
    m.near_col = op->near_col.p;
    m.near_val = op->near_val.p;
    m.ntp = op->mesh->node_tri_ptr.p;
    m.ntri = op->mesh->node_tri.p;
    m.ntk = op->mesh->node_tri_k.p;
    b->max_fold = std::max(b->max_fold, m.ne * m.nt);
    b->max_red = std::max(b->max_red, m.nred);
    b->max_rows = std::max(b->max_rows, m.nrows);
    b->max_fin = std::max(b->max_fin, m.nfin);
  }
  EBH_TRY(b->mdesc.upload(reinterpret_cast<const unsigned char *>(md.data()),
                          sizeof(MemberDesc) * nops));
  std::vector<FarArgs<R>> d(nops);
  std::vector<int> off(8 * (size_t)(nops + 1), 0);
  size_t lds = 0;
  for (int k = 0; k < nops; k++) {
    fill_far_args(b->ops[k], d[k]);
    lds = std::max(lds, far_lds(b->ops[k], (int)sizeof(R)));
  }
  int items = 0;
  for (int x = 0; x < 8; x++) {
    int *o = &off[(size_t)x * (nops + 1)];
    for (int k = 0; k < nops; k++) {
      const ebh_op *op = b->ops[k];
      const int tb = (op->mesh->T.ne + 255) / 256;
      const int cx = op->sym ? 0 : std::max(0, (op->chunks - x + 7) / 8);
      o[k + 1] = o[k] + tb * cx;
    }
    items += o[nops];
  }
  // symmetric members: their workgroups, operator by operator
  std::vector<int> soff((size_t)nops + 1, 0);
  for (int k = 0; k < nops; k++) {
    const ebh_op *op = b->ops[k];
    int nblk = 0;
    if (op->sym) {
      const int ng = op->chunks / sym_slots(op);
      for (int M = 0; 4 * M < ng; M++) nblk += ng - 4 * M;
      nblk *= sym_slots(op);
    }
    soff[k + 1] = soff[k] + nblk;
  }
  EBH_TRY(b->desc.upload(reinterpret_cast<const unsigned char *>(d.data()),
                         sizeof(FarArgs<R>) * nops));
  EBH_TRY(b->off.upload(off));
  EBH_TRY(b->soff.upload(soff));
  EBH_TRY(b->counters.alloc(9));
  b->kern = multi_kernel_for<R>(b->kernel, b->nt);
  b->skern = multi_sym_kernel<R>(b->nt);
  if (!b->kern || (soff[nops] > 0 && !b->skern)) {
    set_error("unsupported kernel or rule size for a batch");
    return EBH_ERR_UNSUPPORTED;
  }
  int per_cu = 0, cus = 0;
  EBH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->kern, 256, lds));
  EBH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                b->ops[0]->mesh->device));
  b->grid = std::max(1, std::min(items, std::max(1, per_cu) * cus));
  b->items = items;
  b->sitems = soff[nops];
  b->sgrid = 0;
  if (b->sitems > 0) {
    int sper = 0;
    EBH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sper, b->skern, 256, 0));
    b->sgrid = std::max(1, std::min(b->sitems, std::max(1, sper) * cus));
  }
  b->lds = lds;
  for (int k = 0; k < nops; k++) b->gens[k] = b->ops[k]->gen;
  return EBH_OK;
}

template <typename R>
int batch_apply(ebh_batch *b, const double *const *x, double *const *y, hipStream_t s)
{
  bool stale = b->grid == 0;
  for (int k = 0; k < b->nops; k++) stale |= b->gens[k] != b->ops[k]->gen;
  if (stale) EBH_TRY(batch_prepare<R>(b));
  BatchPtrs p{};
  for (int k = 0; k < b->nops; k++) {
    p.x[k] = x[k];
    p.y[k] = y[k];
  }
  const MemberDesc *md = reinterpret_cast<const MemberDesc *>(b->mdesc.p);
  const dim3 fgrid((b->max_fold + 255) / 256, b->nops);
  switch (b->kernel) {
  case EBH_LAPLACE_DL:
    hipLaunchKernelGGL((k_fold_multi<R, 1>), fgrid, dim3(256), 0, s, md,
                       b->ops[0]->rule, p);
    break;
  case EBH_KELVIN_SL:
    hipLaunchKernelGGL((k_fold_multi<R, 2>), fgrid, dim3(256), 0, s, md,
                       b->ops[0]->rule, p);
    break;
  default:  // single layer and adjoint double layer: point strengths
    hipLaunchKernelGGL((k_fold_multi<R, 0>), fgrid, dim3(256), 0, s, md,
                       b->ops[0]->rule, p);
    break;
  }
  EBH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, 9, b->counters.p);
  EBH_HIP(hipGetLastError());
  const FarArgs<R> *descs = reinterpret_cast<const FarArgs<R> *>(b->desc.p);
  int nops = b->nops;
  if (b->items > 0) {
    const int *offp = b->off.p;
    int *ctr = b->counters.p;
    void *args[] = {(void *)&descs, (void *)&offp, (void *)&nops, (void *)&ctr};
    EBH_HIP(hipLaunchKernel(b->kern, dim3(b->grid), dim3(256), args, b->lds, s));
    hipLaunchKernelGGL(k_reduce_multi, dim3((b->max_red + 255) / 256, b->nops),
                       dim3(256), 0, s, md);
    EBH_HIP(hipGetLastError());
  }
  if (b->sitems > 0) {
    const int *soffp = b->soff.p;
    int *sctr = b->counters.p + 8;
    void *args[] = {(void *)&descs, (void *)&soffp, (void *)&nops, (void *)&sctr};
    EBH_HIP(hipLaunchKernel(b->skern, dim3(b->sgrid), dim3(256), args, 0, s));
    hipLaunchKernelGGL(k_reduce_wide_multi, dim3((b->max_red + 31) / 32, b->nops),
                       dim3(256), 0, s, md);
    EBH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finish_multi, dim3(finish_blocks(b->max_rows), b->nops),
                     dim3(256), 0, s, md, p);
  EBH_HIP(hipGetLastError());
  if (b->max_fin > 0) {
    hipLaunchKernelGGL(k_reduce_finish_multi, dim3(b->max_fin, b->nops), dim3(256), 0, s,
                       md, p);
    EBH_HIP(hipGetLastError());
  }
  return EBH_OK;
}

}  // namespace
}  // namespace ebh

using namespace ebh;

extern "C" {

int ebh_batch_create(ebh_op *const *ops, int nops, ebh_batch **out)
{
  if (!ops || !out || nops < 1 || nops > kMaxBatch) {
    set_error("a batch holds 1.." + std::to_string(kMaxBatch) + " operators");
    return EBH_ERR_ARG;
  }
  *out = nullptr;
  for (int k = 0; k < nops; k++) {
    if (!ops[k]) { set_error("null operator in batch"); return EBH_ERR_ARG; }
    for (int j = 0; j < k; j++)
      if (ops[j] == ops[k]) {
        set_error("an operator appears twice in the batch");
        return EBH_ERR_ARG;
      }
    if (ops[k]->inner || ops[k]->pscale) {
      set_error("batch the operators a composite operator wraps instead");
      return EBH_ERR_UNSUPPORTED;
    }
    if (!reg_rule(ops[k]->rule.nt)) {
      set_error("batches take far rules of at most 13 points (triangle rule "
                "indices 0..6)");
      return EBH_ERR_UNSUPPORTED;
    }
    if (!chunked_ok(ops[k])) {
      // a batch runs double layers on the chunked schedule
      set_error("the FP32 adjoint double layer runs on the pair-symmetric schedule "
                "only, which batches take for single layers");
      return EBH_ERR_UNSUPPORTED;
    }
    if (ops[k]->d.kernel != ops[0]->d.kernel ||
        ops[k]->d.far_precision != ops[0]->d.far_precision ||
        ops[k]->rule.nt != ops[0]->rule.nt ||
        ops[k]->mesh->device != ops[0]->mesh->device) {
      set_error("batched operators need one kernel, far-field precision, rule "
                "size and device");
      return EBH_ERR_ARG;
    }
  }
  ebh_batch *b = new ebh_batch();
  b->ops.assign(ops, ops + nops);
  b->nops = nops;
  b->kernel = ops[0]->d.kernel;
  b->prec = ops[0]->d.far_precision;
  b->nt = ops[0]->rule.nt;
  b->gens.assign(nops, -1);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            ops[0]->mesh->device) != hipSuccess)
    cus = 256;
  int rc = batch_shapes(b, 4 * cus);
  if (rc == EBH_OK)
    rc = b->prec == EBH_FP32 ? batch_prepare<float>(b) : batch_prepare<double>(b);
  if (rc != EBH_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return EBH_OK;
}

int ebh_batch_apply(ebh_batch *b, const double *const *x, double *const *y,
                    void *stream)
{
  if (!b || !x || !y) { set_error("bad batch apply arguments"); return EBH_ERR_ARG; }
  for (int k = 0; k < b->nops; k++)
    if (!x[k] || !y[k]) { set_error("null vector in batch apply"); return EBH_ERR_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : b->ops[0]->stream;
  return b->prec == EBH_FP32 ? batch_apply<float>(b, x, y, s)
                             : batch_apply<double>(b, x, y, s);
}

int ebh_batch_info(const ebh_batch *b, int64_t *info)
{
  if (!b || !info) { set_error("null argument"); return EBH_ERR_ARG; }
  info[0] = b->nops;
  info[1] = b->items + b->sitems;
  info[2] = b->grid + b->sgrid;
  info[3] = (int64_t)b->lds;
  return EBH_OK;
}

void ebh_batch_destroy(ebh_batch *b) { delete b; }

}  // extern "C"

