// C-ABI entry points (include/easybem_hip.h): error reporting, device memory
// helpers, mesh and operator handles, mass matrices.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ebh_internal.h"

namespace ebh {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
const char *get_error() { return g_error.c_str(); }
std::shared_mutex &capture_gate()
{
  static std::shared_mutex g;
  return g;
}
double &thread_partial_limit()
{
  static thread_local double bytes = 4.0 * 1024 * 1024 * 1024;
  return bytes;
}

std::shared_ptr<MemAccount> &current_account()
{
  static thread_local std::shared_ptr<MemAccount> a;
  return a;
}
}  // namespace ebh

using namespace ebh;

namespace {
// the single layer inside a composite takes its densities three at a time
int inner_three(ebh_op *v)
{
  v->nr_max = 3;
  if (!v->sym) return op_configure(v, v->block, v->tile, v->chunks);
  // three right-hand sides per symmetric launch triple its partial sums:
  // keep the split if they fit, else one step range, else chunked
  if (op_configure(v, 0, 0, v->ksplit, true) == EBH_OK) return EBH_OK;
  if (op_configure(v, 0, 0, 1, true) == EBH_OK) return EBH_OK;
  return op_configure(v, 0, 0, 0);
}
}  // namespace

size_t ebh_op::bytes() const
{
  return (inner ? inner->bytes() : 0) + (inner_dl ? inner_dl->bytes() : 0) +
         (inner_kel ? inner_kel->bytes() : 0) + kt1.bytes() + kt2.bytes() +
         curl.bytes() + mvec.bytes() + hsw.bytes() +
         hsz.bytes() + hsdot.bytes() +
         near_ptr.bytes() + near_col.bytes() + near_val.bytes() + tq32.bytes() +
         sg32.bytes() + tq64.bytes() + sg64.bytes() + tw.bytes() + swt.bytes() +
         normals.bytes() + src32.bytes() + src64.bytes() + partial.bytes() +
         tri_out.bytes() + mgeom32.bytes() + msig32.bytes() + mgeom64.bytes() +
         msig64.bytes() + src3_32.bytes() + src3_64.bytes() + tsl32.bytes() +
         tsl64.bytes() + tc32.bytes() + dinv.bytes() + mass_ptr.bytes() +
         mass_col.bytes() + mass_val.bytes() + cheb_r.bytes() + cheb_d.bytes() +
         cheb_t.bytes() + chebx.bytes() + mt_rowoff.bytes() + mt_item.bytes() +
         mt_kptr.bytes() + mt_order.bytes() + mt_tile32.bytes() + mt_tile64.bytes() +
         mt_part.bytes();
}

extern "C" {

const char *ebh_last_error(void) { return get_error(); }

int ebh_version(int *major, int *minor)
{
  if (major) *major = 1;
  if (minor) *minor = 0;
  return EBH_OK;
}

int ebh_malloc(void **ptr, size_t bytes)
{
  if (!ptr) { set_error("null pointer"); return EBH_ERR_ARG; }
  *ptr = nullptr;
  if (gated([&] { return hipMalloc(ptr, bytes ? bytes : 1); }) != hipSuccess) {
    *ptr = nullptr;
    set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
    return EBH_ERR_NOMEM;
  }
  return EBH_OK;
}
int ebh_free(void *ptr)
{
  if (ptr) EBH_HIP(gated([&] { return hipFree(ptr); }));
  return EBH_OK;
}
// Operator and solver streams are non-blocking, so these helpers first wait
// for all device work: they are for callers that do not manage streams.
int ebh_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
  EBH_HIP(gated([&] { return hipDeviceSynchronize(); }));
  EBH_HIP(gated([&] { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); }));
  return EBH_OK;
}
int ebh_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
  EBH_HIP(gated([&] { return hipDeviceSynchronize(); }));
  EBH_HIP(gated([&] { return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost); }));
  return EBH_OK;
}
int ebh_memset(void *dst, int value, size_t bytes)
{
  EBH_HIP(gated([&] { return hipDeviceSynchronize(); }));
  EBH_HIP(gated([&] { return hipMemset(dst, value, bytes); }));
  EBH_HIP(gated([&] { return hipDeviceSynchronize(); }));
  return EBH_OK;
}
int ebh_synchronize(void)
{
  EBH_HIP(gated([&] { return hipDeviceSynchronize(); }));
  return EBH_OK;
}
int ebh_device_info(int *n_devices, int *n_cu, size_t *free_bytes,
                    size_t *total_bytes)
{
  int n = 0;
  EBH_HIP(hipGetDeviceCount(&n));
  if (n_devices) *n_devices = n;
  if (n == 0) return EBH_OK;
  hipDeviceProp_t prop;
  EBH_HIP(hipGetDeviceProperties(&prop, 0));
  if (n_cu) *n_cu = prop.multiProcessorCount;
  size_t fr = 0, tot = 0;
  EBH_HIP(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = fr;
  if (total_bytes) *total_bytes = tot;
  return EBH_OK;
}

// ------------------------------------------------------------------ mesh
int ebh_mesh_create(int np, const double *nodes, int ne, const int *elems,
                    ebh_mesh **out)
{
  if (!out || !nodes || !elems) { set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  ebh_mesh *m = new ebh_mesh();
  std::string err = build_topology(np, nodes, ne, elems, m->T);
  if (!err.empty()) {
    delete m;
    set_error("mesh rejected: " + err);
    return EBH_ERR_MESH;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    delete m;
    set_error("no HIP device");
    return EBH_ERR_HIP;
  }
  m->device = dev;
  const Topology &T = m->T;
  int rc = eval_scratch_resolve(dev);
  if (rc != EBH_OK) {
    delete m;
    return rc;
  }
  if ((rc = m->nodes.upload(T.nodes)) || (rc = m->tri.upload(T.tri)) ||
      (rc = m->touch_ptr.upload(T.touch_ptr)) || (rc = m->touch.upload(T.touch)) ||
      (rc = m->node_tri_ptr.upload(T.node_tri_ptr)) ||
      (rc = m->node_tri.upload(T.node_tri)) ||
      (rc = m->node_tri_k.upload(T.node_tri_k)) || (rc = m->tri_pos.upload(T.tri_pos)) ||
      (rc = m->fin_blk.upload(T.fin_blk)) || (rc = m->fin_pos.upload(T.fin_pos))) {
    delete m;
    return rc;
  }
  m->nfin = (int)T.fin_blk.size() - 1;
  m->nfpos = T.fin_pos.back();
  m->fin_ok = T.fin_max <= kFinMax;
  *out = m;
  return EBH_OK;
}

int ebh_mesh_info(const ebh_mesh *m, int64_t *info)
{
  if (!m || !info) { set_error("null argument"); return EBH_ERR_ARG; }
  info[0] = m->T.np;
  info[1] = m->T.ne;
  info[2] = m->T.n_class[kCoincident];
  info[3] = m->T.n_class[kEdge];
  info[4] = m->T.n_class[kVertex];
  info[5] = m->T.n_class[kDisjoint];
  info[6] = m->T.max_touch;
  info[7] = (int64_t)m->bytes();
  return EBH_OK;
}

void ebh_mesh_destroy(ebh_mesh *m) { delete m; }

int ebh_mesh_check(int np, const double *nodes, int ne, const int *elems,
                   int64_t *info, int64_t *nnz)
{
  if (!nodes || !elems) { set_error("null argument"); return EBH_ERR_ARG; }
  Topology T;
  std::string err = build_topology(np, nodes, ne, elems, T);
  if (!err.empty()) {
    set_error("mesh rejected: " + err);
    return EBH_ERR_MESH;
  }
  if (info) {
    info[0] = T.np;
    info[1] = T.ne;
    info[2] = T.n_class[kCoincident];
    info[3] = T.n_class[kEdge];
    info[4] = T.n_class[kVertex];
    info[5] = T.n_class[kDisjoint];
    info[6] = T.max_touch;
  }
  if (nnz)
    for (int k = 0; k < 4; k++) {
      NearCSR c;
      build_near_csr(T, k >> 1, k & 1, c);
      nnz[k] = (int64_t)c.col.size();
    }
  return EBH_OK;
}

// -------------------------------------------------------------- operator
int ebh_op_create(ebh_mesh *mesh, const ebh_op_desc *desc, ebh_op **out)
{
  if (!mesh || !desc || !out) { set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  const ebh_op_desc &d = *desc;
  if (d.kernel < 0 || d.kernel > 6) { set_error("unknown kernel"); return EBH_ERR_ARG; }
  const bool elastic = d.kernel == EBH_KELVIN_SL || d.kernel == EBH_KELVIN_DL ||
                       d.kernel == EBH_KELVIN_HS;
  if ((d.test_basis != 0 && d.test_basis != 1) ||
      (d.trial_basis != 0 && d.trial_basis != 1)) {
    set_error("basis must be EBH_P0 or EBH_P1");
    return EBH_ERR_ARG;
  }
  if (d.kernel == EBH_LAPLACE_DL && d.trial_basis != EBH_P1) {
    set_error("the double layer needs a P1 trial space (as the reference's K)");
    return EBH_ERR_UNSUPPORTED;
  }
  if (d.kernel == EBH_LAPLACE_HS &&
      (d.test_basis != EBH_P1 || d.trial_basis != EBH_P1)) {
    set_error("the hypersingular operator needs P1 test and trial spaces");
    return EBH_ERR_UNSUPPORTED;
  }
  if (d.kernel == EBH_LAPLACE_ADL && d.test_basis != EBH_P1) {
    set_error("the adjoint double layer needs a P1 test space");
    return EBH_ERR_UNSUPPORTED;
  }
  if (d.kernel == EBH_LAPLACE_HS ? !std::isfinite(d.stab) : d.stab != 0.0) {
    set_error("stab must be finite, and 0 for kernels other than EBH_LAPLACE_HS");
    return EBH_ERR_ARG;
  }
  if (d.kernel == EBH_KELVIN_SL && d.test_basis != d.trial_basis) {
    set_error("Kelvin single layer needs equal test and trial bases");
    return EBH_ERR_UNSUPPORTED;
  }
  if (d.far_precision != EBH_FP64 && d.far_precision != EBH_FP32) {
    set_error("far_precision must be EBH_FP64 or EBH_FP32");
    return EBH_ERR_ARG;
  }
  if (d.kernel == EBH_KELVIN_DL && (d.test_basis != EBH_P0 || d.trial_basis != EBH_P1)) {
    set_error("the elasticity double layer has a P0 test and a P1 trial space");
    return EBH_ERR_UNSUPPORTED;
  }
  if (d.kernel == EBH_KELVIN_HS && (d.test_basis != EBH_P1 || d.trial_basis != EBH_P1)) {
    set_error("the elasticity hypersingular operator needs P1 test and trial spaces");
    return EBH_ERR_UNSUPPORTED;
  }
  if (elastic && !(d.nu > -1.0 && d.nu < 0.5)) {
    set_error("Poisson ratio must lie in (-1, 0.5)");
    return EBH_ERR_ARG;
  }
  const int nq = line_rule_size(d.line_rule);
  if (nq <= 0) { set_error("line_rule index must be 0..18"); return EBH_ERR_ARG; }
  std::vector<double> rx, ry, rw;
  if (!tri_rule(d.tri_rule, rx, ry, rw)) {
    set_error("tri_rule index must be 0..19");
    return EBH_ERR_UNSUPPORTED;
  }
  ebh_op *op = new ebh_op();
  op->mesh = mesh;
  op->d = d;
  if (op->d.max_rhs < 1) op->d.max_rhs = 1;
  op->nq = nq;
  op->rule.nt = (int)rw.size();
  for (int q = 0; q < op->rule.nt; q++) {
    op->rule.x[q] = rx[q];
    op->rule.y[q] = ry[q];
    op->rule.w[q] = rw[q];
  }
  op->ray = ray_order(op->rule.nt, op->rule.x, op->rule.y, op->rule.w, op->ray_c);
  if (op->rule.nt <= 7 && !op->ray) {
    delete op;
    set_error("far-field rules of up to 7 points must be median-symmetric");
    return EBH_ERR_UNSUPPORTED;
  }
  if (d.kernel == EBH_LAPLACE_ADL && d.far_precision == EBH_FP32 &&
      (op->rule.nt > 7 || op->d.max_rhs > 1)) {
    delete op;
    set_error("the FP32 adjoint double layer runs on the pair-symmetric schedule: far "
              "rules of at most 7 points (tri_rule 0..4) and max_rhs 1");
    return EBH_ERR_UNSUPPORTED;
  }
  op->nc = elastic ? 3 : 1;
  op->nrows = d.test_basis ? mesh->T.np : mesh->T.ne;
  op->ncols = d.trial_basis ? mesh->T.np : mesh->T.ne;
  int rc = EBH_OK;
  // non-blocking: no implicit ordering with the legacy default stream, so
  // other host threads' default-stream calls never touch this stream (or a
  // graph being captured on it)
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
  if (d.kernel == EBH_KELVIN_DL || d.kernel == EBH_KELVIN_HS) {
    // wrapped operators: Laplace V on 3 (K) or 9 (D) P0 densities per RHS,
    // Laplace K (P0 x P1) on the 3 components, Kelvin V on w
    const int per = d.kernel == EBH_KELVIN_DL ? 3 : 9;
    ebh_op_desc vd = d;
    vd.kernel = EBH_LAPLACE_SL;
    vd.test_basis = vd.trial_basis = EBH_P0;
    vd.nu = 0.0;
    vd.max_rhs = op->d.max_rhs > 1 ? per * op->d.max_rhs : 1;
    ebh_op_desc kd = vd;
    kd.kernel = EBH_KELVIN_SL;
    kd.nu = d.nu;
    kd.max_rhs = op->d.max_rhs;
    ebh_op_desc dd = vd;
    dd.kernel = EBH_LAPLACE_DL;
    dd.trial_basis = EBH_P1;
    dd.max_rhs = 1;
    if ((rc = ebh_op_create(mesh, &vd, &op->inner)) ||
        (rc = ebh_op_create(mesh, &kd, &op->inner_kel)) ||
        (d.kernel == EBH_KELVIN_DL && (rc = ebh_op_create(mesh, &dd, &op->inner_dl))) ||
        (rc = kel_build(op)) || (rc = inner_three(op->inner))) {
      ebh_op_destroy(op);
      return rc;
    }
  } else if (d.kernel == EBH_LAPLACE_HS) {
    // the wrapped P0 single layer takes three curl components per RHS
    ebh_op_desc vd = d;
    vd.kernel = EBH_LAPLACE_SL;
    vd.test_basis = vd.trial_basis = EBH_P0;
    vd.stab = 0.0;
    vd.max_rhs = op->d.max_rhs > 1 ? 3 * op->d.max_rhs : 1;
    if ((rc = ebh_op_create(mesh, &vd, &op->inner)) || (rc = hs_build(op)) ||
        (rc = inner_three(op->inner))) {
      ebh_op_destroy(op);
      return rc;
    }
  } else {
    if ((rc = op_build_near(op)) || (rc = op_build_far(op))) {
      ebh_op_destroy(op);
      return rc;
    }
    tune_apply(op);
  }
  *out = op;
  return EBH_OK;
}

int ebh_op_dims(const ebh_op *op, int *nrows, int *ncols)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  if (nrows) *nrows = op->nrows * op->nc;
  if (ncols) *ncols = op->ncols * op->nc;
  return EBH_OK;
}

int ebh_op_apply(ebh_op *op, const double *x, double *y, int nrhs, void *stream)
{
  if (!op || !x || !y || nrhs < 1) { set_error("bad apply arguments"); return EBH_ERR_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : op->stream;
  return op_apply_multi(op, x, y, nrhs, s, false);
}

int ebh_op_apply_host(ebh_op *op, const double *x, double *y, int nrhs)
{
  if (!op || !x || !y || nrhs < 1) { set_error("bad apply arguments"); return EBH_ERR_ARG; }
  const size_t cx = (size_t)op->ncols * op->nc * nrhs, cy = (size_t)op->nrows * op->nc * nrhs;
  for (size_t k = 0; k < cx; k++)
    if (!std::isfinite(x[k])) { set_error("non-finite input vector"); return EBH_ERR_NONFINITE; }
  DevBuf<double> dx, dy;
  EBH_TRY(dx.upload(x, cx));
  EBH_TRY(dy.alloc(cy));
  EBH_TRY(ebh_op_apply(op, dx.p, dy.p, nrhs, nullptr));
  EBH_HIP(hipStreamSynchronize(op->stream));
  EBH_HIP(gated([&] { return hipMemcpy(y, dy.p, cy * sizeof(double), hipMemcpyDeviceToHost); }));
  return EBH_OK;
}

namespace {
bool no_far(const ebh_op *op)
{
  if (op->d.kernel != EBH_SPARSE_INVERSE) return false;
  set_error("a sparse-inverse operator has no far field");
  return true;
}
}  // namespace

int ebh_op_set_launch(ebh_op *op, int block, int tile, int chunks)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  if (no_far(op)) return EBH_ERR_UNSUPPORTED;
  EBH_HIP(hipStreamSynchronize(op->stream));
  return op_configure(far_op(op), block, tile, chunks);
}

int ebh_op_set_schedule(ebh_op *op, int schedule, int split)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  if (no_far(op)) return EBH_ERR_UNSUPPORTED;
  if (schedule != EBH_SCHED_CHUNKED && schedule != EBH_SCHED_SYMMETRIC) {
    set_error("schedule must be EBH_SCHED_CHUNKED or EBH_SCHED_SYMMETRIC");
    return EBH_ERR_ARG;
  }
  if (split < 0) {
    set_error("split must be 0 (default) or a power of 2 up to 64");
    return EBH_ERR_ARG;
  }
  EBH_HIP(hipStreamSynchronize(op->stream));
  ebh_op *f = far_op(op);
  if (schedule == EBH_SCHED_SYMMETRIC) return op_configure(f, 0, 0, split, true);
  return f->sym ? op_configure(f, 0, 0, 0) : EBH_OK;
}

int ebh_op_get_schedule(const ebh_op *op, int *schedule, int *split)
{
  if (!op || !schedule) { set_error("null argument"); return EBH_ERR_ARG; }
  const ebh_op *f = far_op(op);
  *schedule = f->sym ? EBH_SCHED_SYMMETRIC : EBH_SCHED_CHUNKED;
  if (split) *split = f->sym ? f->ksplit : 0;
  return EBH_OK;
}

int ebh_op_set_reduction(ebh_op *op, int fused)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  ebh_op *f = far_op(op);
  f->fused_req = fused != 0;
  f->gen++;
  if (f != op) op->gen++;
  return EBH_OK;
}

int ebh_op_set_partial_limit(ebh_op *op, int64_t bytes)
{
  if (!op || bytes < 0) { set_error("null op or negative limit"); return EBH_ERR_ARG; }
  ebh_op *f = far_op(op);
  f->partial_limit = bytes ? (double)bytes : 4.0 * 1024 * 1024 * 1024;
  if (f->sym) EBH_TRY(op_configure(f, 0, 0, f->ksplit, true));
  else if (f->mts) EBH_TRY(op_configure(f, f->block, f->tile, f->chunks));
  if (f != op) op->gen++;
  return EBH_OK;
}

int ebh_op_set_mrhs_schedule(ebh_op *op, int schedule)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  if (schedule != EBH_MRHS_AUTO && schedule != EBH_MRHS_TARGET && schedule != EBH_MRHS_PAIRS) {
    set_error("schedule must be EBH_MRHS_AUTO, EBH_MRHS_TARGET or EBH_MRHS_PAIRS");
    return EBH_ERR_ARG;
  }
  if (no_far(op)) return EBH_ERR_UNSUPPORTED;
  ebh_op *f = far_op(op);
  if (schedule == EBH_MRHS_PAIRS && !mts_eligible(f)) {
    set_error("the pair-symmetric multi-RHS schedule needs a P0 single layer with "
              "max_rhs > 1 and a far rule of at most 7 points");
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipStreamSynchronize(op->stream));
  f->mrhs_sched = schedule;
  if (f->rb) EBH_TRY(op_configure(f, f->block, f->tile, f->sym ? f->ksplit : f->chunks, f->sym));
  if (f != op) op->gen++;
  return EBH_OK;
}

int ebh_op_set_resident_tiles(ebh_op *op, int64_t max_bytes)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  if (max_bytes < 0) { set_error("resident tile limit must be >= 0"); return EBH_ERR_ARG; }
  if (no_far(op)) return EBH_ERR_UNSUPPORTED;
  ebh_op *f = far_op(op);
  if (max_bytes > 0 && !(f->rb && f->mts)) {
    set_error("resident tiles need the pair-symmetric multi-RHS schedule (a P0 single "
              "layer with max_rhs > 1 and a far rule of at most 7 points)");
    return EBH_ERR_UNSUPPORTED;
  }
  EBH_HIP(hipStreamSynchronize(op->stream));
  f->mt_res_limit = (double)max_bytes;
  if (f->rb) EBH_TRY(op_configure(f, f->block, f->tile, f->sym ? f->ksplit : f->chunks, f->sym));
  if (f != op) op->gen++;
  return EBH_OK;
}

int ebh_op_get_resident_tiles(const ebh_op *op, int *resident, int64_t *bytes)
{
  if (!op || !resident) { set_error("null argument"); return EBH_ERR_ARG; }
  const ebh_op *f = far_op(op);
  *resident = f->mt_res ? 1 : 0;
  if (bytes)
    *bytes = f->mt_res ? (int64_t)(f->mt_tile32.bytes() + f->mt_tile64.bytes()) : 0;
  return EBH_OK;
}

int ebh_op_get_mrhs_schedule(const ebh_op *op, int *schedule, int *passes)
{
  if (!op || !schedule) { set_error("null argument"); return EBH_ERR_ARG; }
  const ebh_op *f = far_op(op);
  *schedule = f->mts ? EBH_MRHS_PAIRS : EBH_MRHS_TARGET;
  if (passes) *passes = f->mts ? (int)f->mt_rows.size() - 1 : 0;
  return EBH_OK;
}

int ebh_op_get_partial(const ebh_op *op, int64_t *limit, int *passes, int64_t *bytes)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  const ebh_op *f = far_op(op);
  if (limit) *limit = (int64_t)f->partial_limit;
  if (passes) *passes = f->sym ? std::max(1, (int)f->pass_m.size() - 1) : 0;
  if (bytes) *bytes = (int64_t)f->partial.bytes();
  return EBH_OK;
}

int ebh_op_get_reduction(const ebh_op *op, int *fused)
{
  if (!op || !fused) { set_error("null argument"); return EBH_ERR_ARG; }
  *fused = fused_pass(far_op(op)) ? 1 : 0;
  return EBH_OK;
}

int ebh_op_get_launch(const ebh_op *op, int *block, int *tile, int *chunks)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  const ebh_op *f = far_op(op);
  if (block) *block = f->block;
  if (tile) *tile = f->tile;
  if (chunks) *chunks = f->chunks;
  return EBH_OK;
}

int ebh_op_time(ebh_op *op, const double *x, double *y, int nrhs, int reps,
                double *ms)
{
  if (!op || !x || !y || !ms || reps < 1 || nrhs < 1) {
    set_error("bad timing arguments");
    return EBH_ERR_ARG;
  }
  double acc[4] = {0, 0, 0, 0};
  const size_t cx = (size_t)op->ncols * op->nc, cy = (size_t)op->nrows * op->nc;
  const bool multi = op->rb && nrhs > 1;
  for (int k = 0; k < reps; k++)
    for (int r = 0; r < (multi ? 1 : nrhs); r++) {
      if (multi) EBH_TRY(op_apply_multi(op, x, y, nrhs, op->stream, true));
      else EBH_TRY(op_apply_one(op, x + r * cx, y + r * cy, op->stream, true));
      EBH_HIP(hipEventSynchronize(op->ev[3]));
      float t01, t12, t23, t03;
      EBH_HIP(hipEventElapsedTime(&t01, op->ev[0], op->ev[1]));
      EBH_HIP(hipEventElapsedTime(&t12, op->ev[1], op->ev[2]));
      EBH_HIP(hipEventElapsedTime(&t23, op->ev[2], op->ev[3]));
      EBH_HIP(hipEventElapsedTime(&t03, op->ev[0], op->ev[3]));
      acc[0] += t01; acc[1] += t12; acc[2] += t23; acc[3] += t03;
    }
  for (int k = 0; k < 4; k++) ms[k] = acc[k] / reps;
  return EBH_OK;
}

int ebh_op_work(const ebh_op *op, int64_t *work)
{
  if (!op || !work) { set_error("null argument"); return EBH_ERR_ARG; }
  const int64_t nt = op->rule.nt;
  // composite operators: far interactions summed over the wrapped applies
  // (single-layer equivalents: hypersingular 3; elasticity K 3 V + 3 K +
  // Kelvin; elasticity D 9 V + Kelvin)
  const int64_t mult = op->d.kernel == EBH_LAPLACE_HS ? 3
                       : op->d.kernel == EBH_KELVIN_DL ? 7
                       : op->d.kernel == EBH_KELVIN_HS ? 10
                       : op->d.kernel == EBH_SPARSE_INVERSE ? 0 : 1;
  work[0] = op->mesh->T.n_class[kDisjoint] * nt * nt * mult;
  work[1] = far_op(op)->near_nnz;
  work[2] = op->nrows;
  work[3] = op->ncols;
  return EBH_OK;
}

int ebh_op_profile(ebh_op *op, int enable, int max_launches, double *far_ms,
                   int64_t *launches)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  if (op->inner) {
    // the wrapped single layer's far launches run on this op's stream
    EBH_HIP(hipStreamSynchronize(op->stream));
    return ebh_op_profile(op->inner, enable, max_launches, far_ms, launches);
  }
  EBH_HIP(hipStreamSynchronize(op->stream));
  if (far_ms || launches) {
    double tot = 0.0;
    for (int k = 0; k < op->prof_count; k++) {
      float ms = 0.f;
      EBH_HIP(hipEventElapsedTime(&ms, op->prof_ev[2 * k], op->prof_ev[2 * k + 1]));
      tot += ms;
    }
    if (far_ms) *far_ms = tot;
    if (launches) *launches = op->prof_count;
  }
  if (enable) {
    if (max_launches < 1) max_launches = 4096;
    for (auto *v : {&op->prof_ev, &op->prof_nev})
      while ((int)v->size() < 2 * max_launches) {
        hipEvent_t e;
        EBH_HIP(hipEventCreate(&e));
        v->push_back(e);
      }
  }
  op->prof_on = enable != 0;
  op->prof_count = 0;
  op->prof_ncount = 0;
  return EBH_OK;
}

int ebh_op_profile_near(ebh_op *op, double *near_ms, int64_t *launches)
{
  if (!op) { set_error("null op"); return EBH_ERR_ARG; }
  ebh_op *f = far_op(op);
  EBH_HIP(hipStreamSynchronize(op->stream));
  double tot = 0.0;
  for (int k = 0; k < f->prof_ncount; k++) {
    float ms = 0.f;
    EBH_HIP(hipEventElapsedTime(&ms, f->prof_nev[2 * k], f->prof_nev[2 * k + 1]));
    tot += ms;
  }
  if (near_ms) *near_ms = tot;
  if (launches) *launches = f->prof_ncount;
  return EBH_OK;
}

void ebh_op_destroy(ebh_op *op)
{
  if (!op) return;
  if (op->stream) (void)hipStreamSynchronize(op->stream);
  ebh_op_destroy(op->inner);
  ebh_op_destroy(op->inner_dl);
  ebh_op_destroy(op->inner_kel);
  op->inner = op->inner_dl = op->inner_kel = nullptr;
  for (auto &e : op->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : op->prof_ev) (void)hipEventDestroy(e);
  for (auto &e : op->prof_nev) (void)hipEventDestroy(e);
  if (op->stream) (void)gated([&] { return hipStreamDestroy(op->stream); });
  delete op;
}

// -------------------------------------------------------- mass matrices
// P0 x P1: each triangle adds J/6 to its three nodes
// (src/BEM3dLaplace.cpp:464-477).  P1 x P1: the element matrix
// (J/24) [[2,1,1],[1,2,1],[1,1,2]].
int ebh_surface_curls(int np, const double *nodes, int ne, const int *elems,
                      double *curl)
{
  if (!nodes || !elems || !curl || np < 3 || ne < 1) {
    set_error("bad curl arguments");
    return EBH_ERR_ARG;
  }
  Topology T;
  T.np = np;
  T.ne = ne;
  T.nodes.assign(nodes, nodes + 3 * (size_t)np);
  T.tri.resize(3 * (size_t)ne);
  for (size_t k = 0; k < 3 * (size_t)ne; k++) {
    if (elems[k] < 1 || elems[k] > np) {
      set_error("element index out of range");
      return EBH_ERR_MESH;
    }
    T.tri[k] = elems[k] - 1;
  }
  for (int i = 0; i < ne; i++) {
    const double *a = &nodes[3 * T.tri[3 * i]], *b = &nodes[3 * T.tri[3 * i + 1]],
                 *c = &nodes[3 * T.tri[3 * i + 2]];
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
                 n2 = u[0] * v[1] - u[1] * v[0];
    if (!(n0 * n0 + n1 * n1 + n2 * n2 > 0.0)) {
      set_error("degenerate triangle " + std::to_string(i));
      return EBH_ERR_MESH;
    }
  }
  std::vector<double> cu;
  surface_curls(T, cu);
  std::copy(cu.begin(), cu.end(), curl);
  return EBH_OK;
}

int ebh_mass_apply(const ebh_mesh *m, int test_basis, const double *u, double *y)
{
  if (!m || !u || !y) { set_error("null argument"); return EBH_ERR_ARG; }
  const Topology &T = m->T;
  const int rows = test_basis ? T.np : T.ne;
  std::fill(y, y + rows, 0.0);
  for (int i = 0; i < T.ne; i++) {
    const int *f = &T.tri[3 * i];
    const double *a = &T.nodes[3 * f[0]], *b = &T.nodes[3 * f[1]], *c = &T.nodes[3 * f[2]];
    double p[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double q[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n0 = p[1] * q[2] - p[2] * q[1], n1 = p[2] * q[0] - p[0] * q[2],
           n2 = p[0] * q[1] - p[1] * q[0];
    double J = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    if (!test_basis) {
      y[i] += J / 6.0 * (u[f[0]] + u[f[1]] + u[f[2]]);
    } else {
      const double s = u[f[0]] + u[f[1]] + u[f[2]];
      for (int k = 0; k < 3; k++) y[f[k]] += J / 24.0 * (s + u[f[k]]);
    }
  }
  return EBH_OK;
}

}  // extern "C"
