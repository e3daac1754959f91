// Solve server: interior Dirichlet Laplace solves served as they arrive,
// with continuous batching on one GPU (see include/easybem_hip.h,
// ebh_server_*).  No reference counterpart: the reference runs one scipy
// solve at a time (python/examples/laplace3dDirichletGPU.py:60-81).
//
// Design:
//   * max_slots worker threads; each admitted request runs its set-up and its
//     solve on its own HIP streams (its operators' streams), so the device
//     interleaves the work of all admitted requests;
//   * admission is FIFO and by device footprint: the head of the queue is
//     admitted when a slot is free and its estimated bytes fit in the
//     budget; idle cached meshes are evicted (least recently used) to make
//     room;
//   * a request's result depends only on its own inputs: no state or
//     reduction is shared between requests, so a result is identical alone
//     and in a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "ebh_internal.h"

namespace {

using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b)
{
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// FNV-1a over the bytes of a mesh: a cached mesh is reused only for a
// request whose mesh is the same
uint64_t mesh_hash(const std::vector<double> &nodes, const std::vector<int> &elems)
{
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t k = 0; k < n; k++) h = (h ^ c[k]) * 1099511628211ull;
  };
  mix(nodes.data(), nodes.size() * sizeof(double));
  mix(elems.data(), elems.size() * sizeof(int));
  return h;
}

struct Request {
  int64_t id = 0;
  int problem = EBH_DIRICHLET;
  int np = 0, ne = 0;
  std::vector<double> nodes;
  std::vector<int> elems;
  uint64_t hash = 0;
  double x0[3] = {0, 0, 0};
  uint64_t key = 0;
  size_t footprint = 0;
  // device bytes the library allocates for this request (set-up and solve)
  std::shared_ptr<ebh::MemAccount> acct;
  Clock::time_point t_submit, t_admit, t_done;
  ebh_request_result res{};
  std::vector<double> t;
  std::string err;
  bool done = false;
};

// A resident mesh with the operators and solver of one problem type:
// Dirichlet: A = P0 single layer V, data operator the double layer K;
// Neumann: A = stabilised hypersingular D + m m^T on P1, data operator the
// adjoint double layer K'.
struct Entry {
  uint64_t key = 0;
  int np = 0, ne = 0;
  uint64_t hash = 0;  // mesh_hash of the mesh it was built for
  ebh_mesh *mesh = nullptr;
  ebh_op *A64 = nullptr, *A32 = nullptr, *K64 = nullptr;
  ebh_op *P64 = nullptr, *P32 = nullptr;  // right preconditioners
  ebh_solver *solver = nullptr;
  size_t bytes = 0;
  uint64_t last_use = 0;
  ~Entry()
  {
    ebh_solver_destroy(solver);
    ebh_op_destroy(A64);
    ebh_op_destroy(A32);
    ebh_op_destroy(K64);
    ebh_op_destroy(P64);
    ebh_op_destroy(P32);
    ebh_mesh_destroy(mesh);
  }
};

using CacheKey = std::pair<uint64_t, int>;  // (mesh key, problem)

// Far-field partial sums of one operator, row FP64 values per source slot
// (outputs per triangle times triangles times right-hand sides per launch;
// with P1 test functions the fused pass's padded row).  The pair-symmetric
// schedule keeps ceil(ne/64) x split slots in one pass while that fits in
// the limit, else it runs in passes (op_configure); the split is the larger
// of the default and any tuned split for this size class, as op_configure
// applies them.
double partial_bytes(int ne, double row, double limit)
{
  const int ng = (ne + 63) / 64;
  const int ks = std::max(ebh::sym_default_split(ng), ebh::tune_max_split(ne));
  const double one = (double)ng * ks * row * 8.0;
  if (one <= limit) return one;
  // passes: buffers within the limit (one block row at least: its 256
  // triangles' rows and its 4 groups' rows of every triangle), then 8 slice
  // sums per output
  const double block_row = (ng * 256.0 + 4.0 * ne) * ks * (row / ne) * 8.0;
  return std::max(limit, block_row) + 8.0 * row * 8.0;
}

// The fused pass's padded partial row length for P1 test functions on a
// mesh (1-based elements; indices out of range are skipped here and
// rejected when the mesh is built).
double padded_p1_row(int np, int ne, const int *elems)
{
  std::vector<int> ntp((size_t)np + 1, 0), blk, base;
  for (size_t i = 0; i < (size_t)3 * ne; i++)
    if (elems[i] >= 1 && elems[i] <= np) ntp[elems[i]]++;
  for (int n = 0; n < np; n++) ntp[n + 1] += ntp[n];
  ebh::fin_blocks(ntp, blk, base);
  return std::max(3.0 * ne, (double)base.back());
}

// Device bytes of one request, estimated before anything is allocated: the
// far-field partial sums of every operator the request builds (above), plus
// the rest -- near-field pairs (about 14 per triangle, with their FP64
// blocks during set-up) and CSR, far-field quadrature data, the Krylov basis
// and work vectors, the mesh -- by a bound that measured requests stay
// under.  Operators with partial sums:
//   P0 Dirichlet: V (FP64, FP32) and the FP64 double layer, 1 value each;
//   P1 Dirichlet: V (FP64, FP32) and the double layer with 3 values, the
//     Calderon preconditioner (FP32 in a mixed solve, else FP64) with a P0
//     V on 3 densities;
//   Neumann: D (FP64, FP32) with a P0 V on 3 densities, the adjoint double
//     layer with 3 values, the dual Calderon preconditioner (FP32 in a
//     mixed solve, else FP64) with a P1 V.
size_t estimate_bytes(int np, int ne, int restart, bool mixed, bool neumann, bool p1,
                      bool precond, double p1_row, double limit)
{
  const double pairs = 14.0 * ne;
  const double near = pairs * (9 * 8 + 6 * 4 + 12 + 8) + 2 * pairs * 12;
  const double far_one = ne * 7.0 * (3 * 12 + 64 + 16) + 4096.0 * 256 * 8 * 2;
  const int nprec = precond && (neumann || p1) ? 1 : 0;
  const int nops = (mixed ? 3 : 2) + nprec;
  const double ops = nops * (near / 2 + far_one);
  const double krylov = (restart + 6.0) * ne * 8;
  const double mesh = np * 24.0 + ne * 100.0;
  const int nsys = mixed ? 2 : 1;
  const double partial = neumann || p1
                             ? (nsys + 1 + nprec) * partial_bytes(ne, p1_row, limit)
                             : (nsys + 1) * partial_bytes(ne, ne, limit);
  return (size_t)(1.25 * (near + ops + krylov + mesh) + partial) + (16u << 20);
}

}  // namespace

struct ebh_server {
  ebh_server_opts o{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<int64_t, std::shared_ptr<Request>> reqs;
  std::multimap<CacheKey, std::unique_ptr<Entry>> idle;  // resident, not in use
  std::vector<std::thread> workers;
  bool stopping = false;
  size_t budget = 0, held = 0, peak = 0;
  int64_t next_id = 1, n_sub = 0, n_done = 0, n_fail = 0, n_active = 0, n_hits = 0;
  uint64_t tick = 0;
  // measured device bytes of all requests and resident meshes
  std::shared_ptr<ebh::MemAccount> acct = std::make_shared<ebh::MemAccount>();
  // evicted meshes, destroyed by a worker after it releases mu (their
  // hipFree and stream destruction must not block submit, wait or stats)
  std::vector<std::unique_ptr<Entry>> trash;

  // Must hold mu.  Evicts least recently used idle meshes until `need` more
  // bytes fit, but only when evicting them all would let it fit: otherwise
  // the cache stays as it is.
  bool evict_for(size_t need)
  {
    if (held + need <= budget) return true;
    size_t idle_bytes = 0;
    for (auto &kv : idle) idle_bytes += kv.second->bytes;
    if (held - idle_bytes + need > budget) return false;
    while (held + need > budget && !idle.empty()) {
      auto lru = idle.begin();
      for (auto it = idle.begin(); it != idle.end(); ++it)
        if (it->second->last_use < lru->second->last_use) lru = it;
      held -= lru->second->bytes;
      trash.push_back(std::move(lru->second));
      idle.erase(lru);
    }
    return held + need <= budget;
  }

  void worker()
  {
    while (true) {
      std::shared_ptr<Request> r;
      std::unique_ptr<Entry> entry;
      std::vector<std::unique_ptr<Entry>> dead;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          if (queue.empty()) return stopping;
          auto &h = queue.front();
          if (h->key) {
            auto it = idle.find(CacheKey(h->key, h->problem));
            if (it != idle.end()) return true;
          }
          return evict_for(h->footprint) || (held == 0);
        });
        if (queue.empty()) return;
        r = queue.front();
        queue.pop_front();
        auto it = r->key ? idle.find(CacheKey(r->key, r->problem)) : idle.end();
        if (it != idle.end()) {
          entry = std::move(it->second);
          idle.erase(it);
          r->res.cache_hit = 1;
          n_hits++;
        } else {
          evict_for(r->footprint);
          held += r->footprint;
        }
        peak = std::max(peak, held);
        n_active++;
        r->t_admit = Clock::now();
        dead.swap(trash);
      }
      dead.clear();
      const bool hit = entry != nullptr;
      bool keep = false;  // a cached mesh that does not match stays cached
      int rc;
      {
        ebh::AccountScope scope(r->acct);
        rc = run(*r, entry, keep);
      }
      if (rc != EBH_OK) r->err = ebh_last_error();
      // bytes stay accounted while the mesh stays resident
      const size_t bytes = hit ? entry->bytes : r->footprint;
      const bool cache = entry && r->key && (rc == EBH_OK || keep);
      // a mesh that is not kept is freed before the request reports done,
      // outside the lock
      if (!cache) entry.reset();
      {
        std::lock_guard<std::mutex> lk(mu);
        r->t_done = Clock::now();
        r->res.status = rc;
        r->res.queue_ms = ms_between(r->t_submit, r->t_admit);
        r->res.total_ms = ms_between(r->t_submit, r->t_done);
        r->res.footprint = r->footprint;
        r->res.device_peak = (size_t)std::max<int64_t>(0, r->acct->peak.load());
        r->done = true;
        n_active--;
        (rc == EBH_OK ? n_done : n_fail)++;
        if (cache) {
          entry->last_use = ++tick;
          entry->bytes = bytes;
          idle.emplace(CacheKey(r->key, r->problem), std::move(entry));
        } else {
          held -= bytes;
        }
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  // b: when given, the right-hand side is formed as soon as its K (or K')
  // exists, while the other operators are still being built
  int build(Request &r, std::unique_ptr<Entry> &e, std::vector<double> *b)
  {
    e.reset(new Entry());
    e->key = r.key;
    e->np = r.np;
    e->ne = r.ne;
    e->hash = r.hash;
    EBH_TRY(ebh_mesh_create(r.np, r.nodes.data(), r.ne, r.elems.data(), &e->mesh));
    const bool neu = r.problem == EBH_NEUMANN;
    const bool p1 = neu || o.dirichlet_p1;
    ebh_op_desc d{};
    d.kernel = neu ? EBH_LAPLACE_HS : EBH_LAPLACE_SL;
    d.test_basis = d.trial_basis = p1 ? EBH_P1 : EBH_P0;
    d.far_precision = EBH_FP64;
    d.line_rule = o.line_rule;
    d.tri_rule = o.tri_rule;
    d.max_rhs = 1;
    // D + m m^T (laplace3dNeumannGPU.py:43-49): constants carry m.u
    d.stab = neu ? 1.0 : 0.0;
    ebh_op_desc d32 = d;
    d32.far_precision = EBH_FP32;
    ebh_op_desc dk = d;
    dk.stab = 0.0;
    dk.kernel = neu ? EBH_LAPLACE_ADL : EBH_LAPLACE_DL;
    dk.test_basis = p1 ? EBH_P1 : EBH_P0;
    dk.trial_basis = neu ? EBH_P0 : EBH_P1;
    // The operators and the Calderon preconditioners depend on the mesh
    // only: they are built on host threads at once (each operator has its
    // own stream; the mesh's near-field pattern cache is locked).  Errors
    // come back with their messages, which are per thread.
    ebh_mesh *mesh = e->mesh;
    using Result = std::pair<int, std::string>;
    std::vector<std::future<Result>> jobs;
    std::shared_ptr<ebh::MemAccount> acct = r.acct;
    const double plim = (double)o.partial_limit;
    ebh::PartialLimitScope limit_here(plim);
    auto spawn = [&](auto fn) {
      auto job = [mesh, fn, acct, plim]() {
        ebh::AccountScope scope(acct);
        ebh::PartialLimitScope limit(plim);
        int rc = hipSetDevice(mesh->device) == hipSuccess ? EBH_OK : EBH_ERR_HIP;
        if (rc == EBH_OK) rc = fn();
        return Result(rc, rc == EBH_OK ? std::string() : std::string(ebh_last_error()));
      };
      try {
        jobs.push_back(std::async(std::launch::async, job));
      } catch (const std::system_error &) {
        // no thread to be had: build it here
        std::promise<Result> done;
        done.set_value(job());
        jobs.push_back(done.get_future());
      }
    };
    Entry *en = e.get();
    if (o.mixed) spawn([=]() { return ebh_op_create(mesh, &d32, &en->A32); });
    const Request *rq = &r;
    spawn([=]() {
      int rc = ebh_op_create(mesh, &dk, &en->K64);
      if (rc == EBH_OK && b)
        rc = neu ? neumann_rhs(*rq, *en, *b) : dirichlet_rhs(*rq, *en, *b);
      return rc;
    });
    const int cal = o.no_precond ? 0 : neu ? 2 : p1 ? 1 : 0;
    // Krylov steps run on the FP32 operator in a mixed solve (the FP64 one
    // only forms residuals), so it needs the FP32 preconditioner alone
    ebh_op **pc = o.mixed ? &en->P32 : &en->P64;
    const int pprec = o.mixed ? EBH_FP32 : EBH_FP64;
    if (cal == 2) {
      // dual Calderon preconditioner Q V Q (Q the Chebyshev inverse of the
      // P1 mass) with a 3-point far and 4-point Gauss near rule: D's small
      // eigenvalues are its smooth modes, which a local inverse cannot reach
      // (its sparse inverse needs more iterations from L5 up)
      spawn([=]() { return ebh_precond_calderon_hs(mesh, pprec, 3, 1, pc); });
    } else if (cal == 1) {
      // P1 Dirichlet: the Calderon preconditioner Q (D + m m^T/(4|Gamma|)) Q
      // (Q the Chebyshev inverse of the P1 mass) with a 1-point far rule
      spawn([=]() { return ebh_precond_calderon(mesh, pprec, 2, 0, 1, pc); });
    }
    int rc = ebh_op_create(mesh, &d, &e->A64);
    std::string msg = rc == EBH_OK ? std::string() : std::string(ebh_last_error());
    for (auto &j : jobs) {
      const Result r2 = j.get();  // every job ends before the entry can go
      if (rc == EBH_OK && r2.first != EBH_OK) {
        rc = r2.first;
        msg = r2.second;
      }
    }
    if (rc != EBH_OK) {
      ebh::set_error(msg);
      return rc;
    }
    EBH_TRY(ebh_solver_create(e->A64, e->A32, o.restart, &e->solver));
    if (cal) {
      EBH_TRY(ebh_solver_set_precond(e->solver, e->P64, e->P32));
    } else if (!o.no_precond) {
      // P0 Dirichlet: near-field sparse approximate inverse of V (2 triangle
      // rings), one FP64 operator for both precisions
      EBH_TRY(ebh_precond_spai(e->A64, 2, 1, &e->P64));
      EBH_TRY(ebh_solver_set_precond(e->solver, e->P64, e->P64));
    }
    // captured Arnoldi steps, replayed by every later solve on this mesh
    // (the capture gate keeps the other workers' allocations out of a
    // capture)
    EBH_TRY(ebh_solver_set(e->solver, o.no_graphs ? 0 : 1, 0));
    return EBH_OK;
  }

  // Dirichlet data u = 1/|x - x0| at the nodes; b = (M/2 + K) u, the first
  // Calderon row V t = (1/2 + K) u (laplace3dDirichletGPU.py:60-70).
  int dirichlet_rhs(const Request &r, Entry &e, std::vector<double> &b)
  {
    const int nrow = o.dirichlet_p1 ? r.np : r.ne;
    std::vector<double> u(r.np), mu_(nrow), ku(nrow);
    for (int n = 0; n < r.np; n++) {
      const double dx = r.nodes[3 * n] - r.x0[0], dy = r.nodes[3 * n + 1] - r.x0[1],
                   dz = r.nodes[3 * n + 2] - r.x0[2];
      u[n] = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
      if (!std::isfinite(u[n])) {
        ebh::set_error("source point lies on the surface");
        return EBH_ERR_NONFINITE;
      }
    }
    EBH_TRY(ebh_mass_apply(e.mesh, o.dirichlet_p1 ? EBH_P1 : EBH_P0, u.data(), mu_.data()));
    EBH_TRY(ebh_op_apply_host(e.K64, u.data(), ku.data(), 1));
    for (int i = 0; i < nrow; i++) b[i] = 0.5 * mu_[i] + ku[i];
    return EBH_OK;
  }

  // Neumann data t = du/dn of u = 1/|x - x0|, the mean of -(x - x0).n/|x -
  // x0|^3 over each triangle by the far-field rule; b = (M^T/2 - K'^T) t,
  // the second Calderon row D u = (1/2 - K') t.
  int neumann_rhs(const Request &r, Entry &e, std::vector<double> &b)
  {
    std::vector<double> rx, ry, rw;
    if (!ebh::tri_rule(o.tri_rule, rx, ry, rw)) {
      ebh::set_error("tri_rule index must be 0..19");
      return EBH_ERR_ARG;
    }
    std::vector<double> t(r.ne), kt(r.np);
    std::fill(b.begin(), b.end(), 0.0);
    for (int i = 0; i < r.ne; i++) {
      const double *P0 = &r.nodes[3 * (r.elems[3 * i] - 1)];
      const double *P1 = &r.nodes[3 * (r.elems[3 * i + 1] - 1)];
      const double *P2 = &r.nodes[3 * (r.elems[3 * i + 2] - 1)];
      double u[3], v[3];
      for (int k = 0; k < 3; k++) { u[k] = P1[k] - P0[k]; v[k] = P2[k] - P0[k]; }
      double nn[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                      u[0] * v[1] - u[1] * v[0]};
      const double J = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
      double acc = 0.0;
      for (size_t q = 0; q < rw.size(); q++) {
        double d[3], r2 = 0.0, dn = 0.0;
        for (int k = 0; k < 3; k++) {
          d[k] = P0[k] + rx[q] * u[k] + ry[q] * v[k] - r.x0[k];
          r2 += d[k] * d[k];
          dn += d[k] * nn[k] / J;
        }
        acc += 2.0 * rw[q] * (-dn / (r2 * std::sqrt(r2)));
      }
      if (!std::isfinite(acc)) {
        ebh::set_error("source point lies on the surface");
        return EBH_ERR_NONFINITE;
      }
      t[i] = acc;
      for (int k = 0; k < 3; k++) b[r.elems[3 * i + k] - 1] += 0.5 * J / 6.0 * acc;
    }
    EBH_TRY(ebh_op_apply_host(e.K64, t.data(), kt.data(), 1));
    for (int n = 0; n < r.np; n++) b[n] -= kt[n];
    return EBH_OK;
  }

  int run(Request &r, std::unique_ptr<Entry> &e, bool &keep)
  {
    const auto t0 = Clock::now();
    if (e && (e->np != r.np || e->ne != r.ne || e->hash != r.hash)) {
      // the key names another mesh: fail this request, keep the cached one
      keep = true;
      r.res.cache_hit = 0;
      ebh::set_error("mesh_key " + std::to_string(r.key) +
                     " is resident for a different mesh");
      return EBH_ERR_ARG;
    }
    r.res.ne = r.problem == EBH_NEUMANN || o.dirichlet_p1 ? r.np : r.ne;
    const bool neu = r.problem == EBH_NEUMANN;
    const int n = neu || o.dirichlet_p1 ? r.np : r.ne;
    std::vector<double> b(n);
    if (!e) {
      // a new mesh: operators, preconditioners and right-hand side
      EBH_TRY(build(r, e, &b));
    } else if (neu) {
      EBH_TRY(neumann_rhs(r, *e, b));
    } else {
      EBH_TRY(dirichlet_rhs(r, *e, b));
    }
    ebh::DevBuf<double> db, dx;
    EBH_TRY(db.upload(b));
    std::vector<double> zero(n, 0.0);
    EBH_TRY(dx.upload(zero));
    const auto t1 = Clock::now();
    ebh_solve_opts so{};
    so.restart = o.restart;
    so.max_iters = o.max_iters;
    so.tol = o.tol;
    so.mixed = o.mixed;
    // Dirichlet: a first FP32 pass to half the target leaves one FP64
    // residual per solve (against two passes to 1e-5: L5 31 -> 25 ms, L6
    // 412 -> 358 ms, one request at a time; to 5e-9 rather than 1e-9: one
    // Arnoldi step fewer on 5 of 8 perturbed L5 meshes, none needing a
    // second pass); Neumann: two passes to 1e-5 take one Arnoldi step fewer
    // than one pass to 1e-9 (L6 480 against 515 ms), so it keeps them
    so.inner_tol = o.inner_tol > 0 ? o.inner_tol : neu ? 1e-5 : 5e-9;
    so.max_outer = 10;
    ebh_solve_stats st{};
    const int rc = ebh_solver_solve(e->solver, db.p, dx.p, &so, &st);
    r.t.resize(n);
    EBH_HIP(ebh::gated([&] { return hipMemcpy(r.t.data(), dx.p, n * sizeof(double), hipMemcpyDeviceToHost); }));
    const auto t2 = Clock::now();
    r.res.iters = st.iters;
    r.res.graph_steps = st.graph_steps;
    r.res.relres = st.relres;
    r.res.setup_ms = ms_between(t0, t1);
    r.res.solve_ms = ms_between(t1, t2);
    return rc;
  }
};

extern "C" {

int ebh_server_create(const ebh_server_opts *opts, ebh_server **out)
{
  if (!out) { ebh::set_error("null argument"); return EBH_ERR_ARG; }
  *out = nullptr;
  ebh_server *s = new ebh_server();
  if (opts) s->o = *opts;
  ebh_server_opts &o = s->o;
  if (o.max_slots <= 0) o.max_slots = 4;
  if (o.restart <= 0) o.restart = 50;
  if (o.tol <= 0) o.tol = 1e-8;
  if (o.max_iters <= 0) o.max_iters = 2000;
  if (!opts) o.mixed = 1;
  if (o.line_rule <= 0) o.line_rule = 4;
  if (o.tri_rule <= 0) o.tri_rule = 4;
  if (o.partial_limit == 0) o.partial_limit = (size_t)512 << 20;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    delete s;
    ebh::set_error("no HIP device");
    return EBH_ERR_HIP;
  }
  s->budget = o.hbm_budget ? o.hbm_budget : (size_t)(0.8 * fr);
  for (int k = 0; k < o.max_slots; k++) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return EBH_OK;
}

int ebh_server_submit(ebh_server *s, int np, const double *nodes, int ne,
                      const int *elems, const double *x0, uint64_t mesh_key,
                      int64_t *request_id)
{
  return ebh_server_submit_bvp(s, EBH_DIRICHLET, np, nodes, ne, elems, x0,
                               mesh_key, request_id);
}

int ebh_server_submit_bvp(ebh_server *s, int problem, int np, const double *nodes,
                          int ne, const int *elems, const double *x0,
                          uint64_t mesh_key, int64_t *request_id)
{
  if (!s || !nodes || !elems || !x0 || !request_id || np < 3 || ne < 4 ||
      (problem != EBH_DIRICHLET && problem != EBH_NEUMANN)) {
    ebh::set_error("bad submit arguments");
    return EBH_ERR_ARG;
  }
  auto r = std::make_shared<Request>();
  r->problem = problem;
  r->np = np;
  r->ne = ne;
  r->nodes.assign(nodes, nodes + (size_t)3 * np);
  r->elems.assign(elems, elems + (size_t)3 * ne);
  if (mesh_key) r->hash = mesh_hash(r->nodes, r->elems);
  for (int k = 0; k < 3; k++) r->x0[k] = x0[k];
  r->key = mesh_key;
  r->acct = std::make_shared<ebh::MemAccount>();
  r->acct->parent = s->acct;
  r->footprint = estimate_bytes(np, ne, s->o.restart, s->o.mixed != 0,
                                problem == EBH_NEUMANN,
                                problem == EBH_NEUMANN || s->o.dirichlet_p1 != 0,
                                s->o.no_precond == 0, padded_p1_row(np, ne, elems),
                                (double)s->o.partial_limit);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->footprint > s->budget) {
      ebh::set_error("request needs about " + std::to_string(r->footprint) +
                     " device bytes, more than the server budget " +
                     std::to_string(s->budget));
      return EBH_ERR_NOMEM;
    }
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    s->n_sub++;
  }
  s->cv_work.notify_all();
  *request_id = r->id;
  return EBH_OK;
}

int ebh_server_wait(ebh_server *s, int64_t id, double timeout_s,
                    ebh_request_result *res, double *t)
{
  if (!s) { ebh::set_error("null server"); return EBH_ERR_ARG; }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) { ebh::set_error("unknown request id"); return EBH_ERR_ARG; }
  auto r = it->second;
  auto pred = [&] { return r->done; };
  if (timeout_s > 0) {
    if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
      ebh::set_error("timed out waiting for the request");
      return EBH_ERR_STATE;
    }
  } else {
    s->cv_done.wait(lk, pred);
  }
  if (res) *res = r->res;
  if (t && r->res.status == EBH_OK) std::copy(r->t.begin(), r->t.end(), t);
  if (r->res.status != EBH_OK) ebh::set_error(r->err);
  s->reqs.erase(it);
  return r->res.status;
}

int ebh_server_stats(ebh_server *s, int64_t *st)
{
  if (!s || !st) { ebh::set_error("null argument"); return EBH_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->mu);
  st[0] = s->n_sub;
  st[1] = s->n_done;
  st[2] = s->n_fail;
  st[3] = s->n_active;
  st[4] = (int64_t)s->queue.size();
  st[5] = (int64_t)s->held;
  st[6] = (int64_t)s->peak;
  st[7] = s->n_hits;
  st[8] = s->acct->live.load();
  st[9] = s->acct->peak.load();
  return EBH_OK;
}

void ebh_server_destroy(ebh_server *s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (auto &w : s->workers) w.join();
  s->idle.clear();
  s->trash.clear();
  delete s;
}

}  // extern "C"
