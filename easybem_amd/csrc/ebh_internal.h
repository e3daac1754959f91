// Internal definitions shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <utility>
#include <vector>

#include "easybem_hip.h"
#include "host_geometry.h"

namespace ebh {

void set_error(const std::string &msg);
const char *get_error();

#define EBH_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      ::ebh::set_error(std::string(#call) + " failed: " +                 \
                       hipGetErrorString(e_));                            \
      return EBH_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

#define EBH_TRY(expr)                                                     \
  do {                                                                    \
    int rc_ = (expr);                                                     \
    if (rc_ != EBH_OK) return rc_;                                        \
  } while (0)

// Device-byte accounting.  Every library allocation (DevBuf) is charged to
// the account current on the allocating thread and to that account's
// parents, and returned to the same accounts when it is freed, on whatever
// thread.  The solve server gives each request an account whose parent is
// the server's, so it can compare measured bytes with its estimates.
struct MemAccount {
  std::atomic<int64_t> live{0}, peak{0};
  std::shared_ptr<MemAccount> parent;
  void add(int64_t b)
  {
    for (MemAccount *a = this; a; a = a->parent.get()) {
      const int64_t v = a->live.fetch_add(b) + b;
      int64_t p = a->peak.load();
      while (v > p && !a->peak.compare_exchange_weak(p, v)) {}
    }
  }
};
// the calling thread's current account (empty: allocations are not charged)
std::shared_ptr<MemAccount> &current_account();
// makes `a` the thread's current account until the scope ends
struct AccountScope {
  std::shared_ptr<MemAccount> prev;
  explicit AccountScope(std::shared_ptr<MemAccount> a) : prev(current_account())
  {
    current_account() = std::move(a);
  }
  ~AccountScope() { current_account() = std::move(prev); }
  AccountScope(const AccountScope &) = delete;
  AccountScope &operator=(const AccountScope &) = delete;
};

// Bound on the pair-symmetric schedule's partial sums (bytes) that
// operators created on the calling thread start with (ebh_op::partial_limit;
// 4 GiB unless a PartialLimitScope is open)
double &thread_partial_limit();
struct PartialLimitScope {
  double prev;
  explicit PartialLimitScope(double bytes) : prev(thread_partial_limit())
  {
    thread_partial_limit() = bytes;
  }
  ~PartialLimitScope() { thread_partial_limit() = prev; }
  PartialLimitScope(const PartialLimitScope &) = delete;
  PartialLimitScope &operator=(const PartialLimitScope &) = delete;
};

// Capture gate.  A stream capture (a solver recording an Arnoldi step as a
// hipGraph) holds it exclusively; the library's device allocations, frees
// and uploads hold it shared, so that no thread allocates while another
// captures (the solve server runs several solvers on host threads).
std::shared_mutex &capture_gate();
// Runs a call on the legacy stream (a synchronous copy or memset, a device
// synchronisation, a host or device allocation) under the capture gate: the
// runtime refuses such calls while another thread captures.
template <typename F>
hipError_t gated(F call)
{
  std::shared_lock<std::shared_mutex> g(capture_gate());
  return call();
}

// Owning device buffer (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  std::shared_ptr<MemAccount> acct;  // charged with the bytes while allocated
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (p) {
      std::shared_lock<std::shared_mutex> g(capture_gate());
      (void)hipFree(p);
    }
    if (acct) acct->add(-(int64_t)(n * sizeof(T)));
    acct.reset();
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count)
  {
    if (count == n && p) return EBH_OK;
    release();
    if (count == 0) return EBH_OK;
    hipError_t e;
    {
      std::shared_lock<std::shared_mutex> g(capture_gate());
      e = hipMalloc(&p, count * sizeof(T));
    }
    if (e != hipSuccess) {
      p = nullptr;
      set_error("device allocation of " + std::to_string(count * sizeof(T)) +
                " bytes failed");
      return EBH_ERR_NOMEM;
    }
    n = count;
    acct = current_account();
    if (acct) acct->add((int64_t)(n * sizeof(T)));
    return EBH_OK;
  }
  int upload(const T *h, size_t count)
  {
    EBH_TRY(alloc(count));
    if (count) {
      std::shared_lock<std::shared_mutex> g(capture_gate());
      EBH_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
    return EBH_OK;
  }
  int upload(const std::vector<T> &v) { return upload(v.data(), v.size()); }
  size_t bytes() const { return n * sizeof(T); }
};

}  // namespace ebh

// Device-resident mesh: topology on the host, the arrays kernels read on the
// device.
struct ebh_mesh {
  int device = 0;
  ebh::Topology T;
  ebh::DevBuf<double> nodes;                    // np*3
  ebh::DevBuf<int> tri;                         // ne*3 (0-based)
  ebh::DevBuf<int> touch_ptr, touch;            // exclusion lists
  ebh::DevBuf<int> node_tri_ptr, node_tri, node_tri_k;
  // P1 test rows of the fused far reduction and near-field pass: padded
  // incidence positions of (t, k), node row blocks and their first padded
  // positions (Topology::tri_pos, fin_blk, fin_pos); nfpos the padded row
  // length; fin_ok: every block's far sums fit the pass's LDS stage
  ebh::DevBuf<int> tri_pos, fin_blk, fin_pos;
  int nfin = 0, nfpos = 0;
  bool fin_ok = false;
  // near-field CSR patterns with their contribution lists, by (test basis,
  // trial basis): every operator of the mesh with those bases shares one
  std::mutex csr_mu;
  std::map<std::pair<int, int>, std::shared_ptr<const ebh::NearCSR>> csr_cache;
  size_t bytes() const
  {
    return nodes.bytes() + tri.bytes() + touch_ptr.bytes() + touch.bytes() +
           node_tri_ptr.bytes() + node_tri.bytes() + node_tri_k.bytes() + tri_pos.bytes() +
           fin_blk.bytes() + fin_pos.bytes();
  }
};

// Far-field rule data passed by value to kernels.
struct RuleArgs {
  int nt;                 // points per triangle
  double x[ebh::kMaxTriPoints], y[ebh::kMaxTriPoints];  // parameter coordinates
  double w[ebh::kMaxTriPoints];                         // weights (sum 1/2)
};

struct ebh_op {
  ebh_mesh *mesh = nullptr;
  ebh_op_desc d{};
  int nq = 0;                    // near-field Gauss points per dimension
  RuleArgs rule{};
  // rule in centroid-median order (ray_order): point q = C + ray_c[q] (P_k - C)
  bool ray = false;
  double ray_c[ebh::kMaxTriPoints] = {};
  int nc = 1;                    // components (3 for Kelvin)
  int nrows = 0, ncols = 0;      // test / trial DOFs (per component)
  int nout = 1;                  // far outputs per target triangle (x nc)
  // near field (FP64 CSR; Kelvin stores 6 symmetric entries per nnz)
  ebh::DevBuf<int> near_ptr, near_col;
  ebh::DevBuf<double> near_val;
  int64_t near_nnz = 0;
  // far-field geometry
  ebh::DevBuf<float> tq32, sg32;     // target [nt][3][ne], source [ne*nt][3]
  ebh::DevBuf<float> tc32;           // FP32 double layer: centroids hi/lo [6][ne]
  ebh::DevBuf<float> tsl32;          // single layer target frame (see op_build_far)
  ebh::DevBuf<double> tsl64;
  ebh::DevBuf<double> tq64, sg64;
  ebh::DevBuf<double> tw, swt;       // w_q J per target / source point
  ebh::DevBuf<double> rphi;          // rules without a register kernel: the
                                     // test shape functions [nt][3] at the points
  ebh::DevBuf<double> normals;       // ne*3 unit normals
  // per-apply workspace
  ebh::DevBuf<float> src32;
  ebh::DevBuf<double> src64;
  ebh::DevBuf<double> partial, tri_out;
  // multi-RHS single layer (rb = RHS per lane, 0 when unused)
  int rb = 0, mchunks = 1, mchunk = 0;
  int mgrid = 4096;  // multi-RHS launch: target workgroup count
  ebh::DevBuf<float> mgeom32, msig32;
  ebh::DevBuf<double> mgeom64, msig64;
  // pair-symmetric multi-RHS P0 schedule (k_mrhs_tiles, k_mrhs_tile_mfma):
  // requested (EBH_MRHS_*) and in effect; tile rows mt_rows[p] ..
  // mt_rows[p+1]-1 form pass p, whose items are mt_ioff[p] .. mt_ioff[p+1]-1
  // of mt_item, with per-group item pointers at mt_koff[p] of mt_kptr;
  // the first tile of pass p is mt_t0[p]; mt_tiles / mt_items: the largest
  // pass's tiles and items.  Resident tiles: when every tile and item fits
  // in mt_res_limit bytes (0: off), one pass whose tiles are formed once at
  // configuration (mt_res once they are) and reused by every apply
  int mrhs_sched = 0;
  bool mts = false;
  double mt_res_limit = 0.0;
  bool mt_res = false;
  std::vector<int> mt_rows, mt_t0, mt_ioff, mt_koff;
  size_t mt_tiles = 0, mt_items = 0;
  ebh::DevBuf<int> mt_rowoff, mt_item, mt_kptr, mt_order;
  ebh::DevBuf<float> mt_tile32;
  ebh::DevBuf<double> mt_tile64, mt_part;
  // hypersingular (EBH_LAPLACE_HS): the wrapped P0 single layer, the curls
  // [ne][3][3], lumped node areas, curl densities / single-layer outputs
  // [nrhs][3][ne] and the partial m.x sums
  ebh_op *inner = nullptr;
  ebh::DevBuf<double> curl, mvec, hsw, hsz, hsdot;
  // Calderon preconditioner form of the hypersingular operator:
  // Q (D + stab m m^T) Q, Q the Chebyshev mass inverse below (input copy in
  // chebx)
  bool hs_scale = false;
  // single layer applied between two Chebyshev mass inverses (the Calderon
  // preconditioner of the hypersingular operator); uses hsw
  bool pscale = false;
  // Chebyshev mass inverse of those two forms: pcheb steps, the inverse
  // lumped node areas dinv, the CSR of the consistent P1 mass and work
  // vectors [max_rhs][np] (residual; cheb_d and cheb_t the search
  // direction, alternately; chebx the hypersingular form's input copy)
  int pcheb = 0;
  ebh::DevBuf<double> dinv;
  ebh::DevBuf<int> mass_ptr, mass_col;
  ebh::DevBuf<double> mass_val, cheb_r, cheb_d, cheb_t, chebx;
  // elasticity K and D (EBH_KELVIN_DL / EBH_KELVIN_HS): the Laplace single
  // layer (inner), Laplace double layer and Kelvin single layer they wrap,
  // and two work buffers
  ebh_op *inner_dl = nullptr, *inner_kel = nullptr;
  ebh::DevBuf<double> kt1, kt2;
  // launch configuration
  int block = 256, tile = 64, chunks = 1, chunk = 0;
  bool sym = false;  // pair-symmetric far-field schedule (k_far_sym)
  int ksplit = 1;    // its step ranges per group pair
  int qsplit = 1;    // target-point halves per group pair (FP64 Kelvin with
                     // 6 or 7 points: 2); ksplit * qsplit slots per pair
  // its partial sums reduced inside the near-field pass (k_reduce_finish)
  // instead of by k_reduce_wide before k_finish (ebh_op_set_reduction)
  bool fused_req = false;
  // its partial sums per launch at most this many bytes (ebh_op_set_partial_limit):
  // above it the schedule runs in passes of block rows pass_m[p] ..
  // pass_m[p+1]-1, each followed by a reduction into 8 slice sums (pass_tsz
  // and pass_ssz: the largest target- and source-side buffers, values per
  // right-hand side)
  double partial_limit = ebh::thread_partial_limit();
  std::vector<int> pass_m;
  size_t pass_tsz = 0, pass_ssz = 0;
  // right-hand sides per symmetric launch the buffers hold (1, or 3 for the
  // single layer inside a composite operator, applied to 3 or 9 densities)
  int nr_max = 1;
  ebh::DevBuf<float> src3_32;
  ebh::DevBuf<double> src3_64;
  int gen = 0;  // bumped when the launch shape changes (captured graphs key on it)
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // far-kernel profiling over a timed region: event pairs around each launch
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;   // 2 per recorded launch
  int prof_count = 0;
  std::vector<hipEvent_t> prof_nev;  // the same around the near-field pass
  int prof_ncount = 0;
  size_t bytes() const;
};

namespace ebh {
int op_apply_one(ebh_op *op, const double *x, double *y, hipStream_t s,
                 bool timed);
int op_apply_multi(ebh_op *op, const double *x, double *y, int nrhs,
                   hipStream_t s, bool timed);
int op_build_near(ebh_op *op);
int op_build_far(ebh_op *op);
// sym: the pair-symmetric schedule (block, tile and chunks then follow from
// the mesh); only for operators that sym_eligible accepts
int op_configure(ebh_op *op, int block, int tile, int chunks, bool sym = false);
bool sym_eligible(const ebh_op *op);
// partial-sum slots per group pair of the pair-symmetric schedule
inline int sym_slots(const ebh_op *op) { return op->ksplit * op->qsplit; }
// false for operators that run on the pair-symmetric schedule only (the
// FP32 adjoint double layer)
bool chunked_ok(const ebh_op *op);
// the pair-symmetric multi-RHS schedule applies (a P0 single layer with
// max_rhs > 1 and a rule of at most 7 points)
bool mts_eligible(const ebh_op *op);
// y = M x for an EBH_SPARSE_INVERSE operator (nrhs vectors back to back)
int spai_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s, bool timed);
// z = Q v with Q the Chebyshev approximation of the P1 mass inverse of a
// hypersingular Calderon preconditioner (z may alias v)
int mass_inv_apply(ebh_op *op, const double *v, double *z, int nrhs, hipStream_t s);
int sym_default_split(int ng);
int hs_build(ebh_op *op);
int hs_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s,
             bool timed);
int kel_build(ebh_op *op);
// applies an operator built from other operators (op->inner set)
int composite_apply(ebh_op *op, const double *x, double *y, int nrhs, hipStream_t s,
                    bool timed);
// the operator whose far-field kernels run (the wrapped single layer for
// the hypersingular operator)
inline ebh_op *far_op(ebh_op *op) { return op->inner ? op->inner : op; }
inline const ebh_op *far_op(const ebh_op *op) { return op->inner ? op->inner : op; }
// the pair-symmetric schedule's partial sums reduced inside the near-field
// pass (k_reduce_finish)
bool fused_pass(const ebh_op *op);
// rule sizes with a register-resident far kernel (1, 3, 4, 6, 7, 12, 13)
bool reg_rule(int nt);
// source chunks that give about `grid` workgroups at this block size
int chunks_for_grid(const ebh_op *op, int block, int grid);
// applies a tuned launch shape recorded for this operator's key, if any
void tune_apply(ebh_op *op);
// looks up the device address of ebh_evaluate_laplace_noalloc's workspace
// on device dev (which must be current); once per device
int eval_scratch_resolve(int dev);
// the largest pair-symmetric step split the tuning table records for meshes
// of ne triangles' size class (0 if none)
int tune_max_split(int ne);
}  // namespace ebh
