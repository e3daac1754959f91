// Launch-shape autotuner for the far-field kernels.
//
// The far-field apply has three shape parameters: the workgroup size (target
// triangles per workgroup, 64/128/256), the LDS source tile (source
// triangles staged per barrier, up to 64 so one 64-bit exclusion mask covers
// a tile) and the number of source chunks, expressed here as the total
// workgroup count of the launch.  The tuner times whole applies with HIP
// events on the operator's stream and walks the three parameters by
// coordinate descent from the current shape; operators that admit the
// pair-symmetric schedule then also time it over its step split, and the
// faster schedule wins.  Results go into a process-wide
// table keyed by device, operator kind and mesh size class; operators created
// later under the same key start from the tuned shape.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ebh_internal.h"

namespace ebh {
namespace {

struct Entry {
  int block, tile, grid, mgrid;
  double ms;
  int sched;  // 0: chunked; k >= 1: pair-symmetric with k step ranges
};

std::mutex g_mu;
std::map<std::string, Entry> g_table;

std::string device_tag(int dev)
{
  static std::mutex mu;
  static std::map<int, std::string> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t p;
  std::string tag = "unknown";
  if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
    std::string arch = p.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    tag = arch + "x" + std::to_string(p.multiProcessorCount);
  }
  cache[dev] = tag;
  return tag;
}

// Size class: half-octaves of the triangle count.
std::string key_of(const ebh_op *op)
{
  const int ne = op->mesh->T.ne;
  const int cls = (int)std::lround(2.0 * std::log2((double)std::max(ne, 1)));
  char buf[192];
  std::snprintf(buf, sizeof buf, "%s/k%d/p%d/nt%d/test%d/rb%d/ne%d",
                device_tag(op->mesh->device).c_str(), op->d.kernel,
                op->d.far_precision, op->rule.nt, op->d.test_basis, op->rb, cls);
  return buf;
}

// Best of `reps` whole applies (ms), single or multi-RHS.
int time_apply(ebh_op *op, const double *x, double *y, bool multi, int reps,
               double *ms)
{
  double best = 1e300;
  for (int k = 0; k < reps; k++) {
    if (multi) EBH_TRY(op_apply_multi(op, x, y, op->d.max_rhs, op->stream, true));
    else EBH_TRY(op_apply_one(op, x, y, op->stream, true));
    EBH_HIP(hipEventSynchronize(op->ev[3]));
    float t = 0.f;
    EBH_HIP(hipEventElapsedTime(&t, op->ev[0], op->ev[3]));
    best = std::min(best, (double)t);
  }
  *ms = best;
  return EBH_OK;
}

// workgroups of the far-field launch
int grid_of(const ebh_op *op)
{
  if (op->sym) {
    const int ng = op->chunks / sym_slots(op);
    int nblk = 0;
    for (int M = 0; 4 * M < ng; M++) nblk += ng - 4 * M;
    return nblk * sym_slots(op);
  }
  const int tb = (op->mesh->T.ne + op->block - 1) / op->block;
  return tb * op->chunks;
}

}  // namespace

int tune_max_split(int ne)
{
  const int cls = (int)std::lround(2.0 * std::log2((double)std::max(ne, 1)));
  const std::string tail = "/ne" + std::to_string(cls);
  int best = 0;
  std::lock_guard<std::mutex> lk(g_mu);
  for (const auto &kv : g_table) {
    const std::string &k = kv.first;
    if (k.size() >= tail.size() && k.compare(k.size() - tail.size(), tail.size(), tail) == 0)
      best = std::max(best, kv.second.sched);
  }
  return best;
}

void tune_apply(ebh_op *op)
{
  Entry e;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(key_of(op));
    if (it == g_table.end()) return;
    e = it->second;
  }
  if (e.mgrid > 0) op->mgrid = e.mgrid;
  // a recorded shape was valid when it was tuned; on failure the operator
  // keeps its default shape
  if (e.sched > 0 && sym_eligible(op)) {
    if (op_configure(op, 0, 0, e.sched, true) == EBH_OK) return;
  }
  if (op_configure(op, e.block, e.tile, chunks_for_grid(op, e.block, e.grid)) != EBH_OK)
    (void)op_configure(op, 0, 0, 0);
}

}  // namespace ebh

using namespace ebh;

extern "C" {

int ebh_op_autotune(ebh_op *op, double budget_ms, ebh_tune_result *out)
{
  if (!op || !out) { set_error("null argument"); return EBH_ERR_ARG; }
  if (op->d.kernel == EBH_SPARSE_INVERSE) {
    set_error("a sparse-inverse operator has no far field to tune");
    return EBH_ERR_UNSUPPORTED;
  }
  if (op->inner) {
    EBH_HIP(hipStreamSynchronize(op->stream));
    return ebh_op_autotune(op->inner, budget_ms, out);
  }
  EBH_HIP(hipSetDevice(op->mesh->device));
  EBH_HIP(hipStreamSynchronize(op->stream));
  const int nrhs = op->rb ? op->d.max_rhs : 1;
  DevBuf<double> x, y;
  EBH_TRY(x.alloc((size_t)op->ncols * op->nc * nrhs));
  EBH_TRY(y.alloc((size_t)op->nrows * op->nc * nrhs));
  // zero input: the far-field kernels do the same work for any values
  EBH_HIP(hipMemsetAsync(x.p, 0, x.bytes(), op->stream));
  EBH_HIP(hipStreamSynchronize(op->stream));

  const int reps = 3;
  double spent = 0.0;
  int tried = 0;
  // distinct effective shapes (block, tile, chunks) already timed
  std::map<std::tuple<int, int, int>, double> seen;
  auto eval = [&](int block, int tile, int grid, double *ms) -> int {
    const int chunks = chunks_for_grid(op, block, grid);
    auto key = std::make_tuple(block, tile, chunks);
    auto it = seen.find(key);
    if (it != seen.end()) { *ms = it->second; return EBH_OK; }
    EBH_TRY(op_configure(op, block, tile, chunks));
    EBH_TRY(time_apply(op, x.p, y.p, false, reps, ms));
    spent += *ms * reps;
    tried++;
    seen[key] = *ms;
    return EBH_OK;
  };

  // the starting configuration, then the chunked search from its shape (or
  // from the default chunked shape when it starts symmetric)
  double ms0 = 0.0;
  EBH_TRY(time_apply(op, x.p, y.p, false, 1, &ms0));  // warm-up, not kept
  EBH_TRY(time_apply(op, x.p, y.p, false, reps, &ms0));
  spent += ms0 * reps;
  tried++;
  const bool start_sym = op->sym;
  const int start_ks = op->ksplit;
  int bb = start_sym ? 256 : op->block, bt = start_sym ? 64 : op->tile;
  int bg = start_sym ? 4096 : grid_of(op);
  double best = 1e300;
  const bool chunked = chunked_ok(op);
  if (chunked) EBH_TRY(eval(bb, bt, bg, &best));
  const int blocks[] = {64, 128, 256};
  const int tiles[] = {16, 32, 64};
  const int grids[] = {1024, 2048, 4096, 8192, 16384, 32768};
  for (int round = 0; round < 2 && chunked && spent < budget_ms; round++) {
    for (int g : grids) {
      if (spent >= budget_ms) break;
      double ms;
      EBH_TRY(eval(bb, bt, g, &ms));
      if (ms < best) { best = ms; bg = g; }
    }
    for (int b : blocks) {
      if (spent >= budget_ms) break;
      double ms;
      EBH_TRY(eval(b, bt, bg, &ms));
      if (ms < best) { best = ms; bb = b; }
    }
    for (int t : tiles) {
      if (spent >= budget_ms) break;
      double ms;
      EBH_TRY(eval(bb, t, bg, &ms));
      if (ms < best) { best = ms; bt = t; }
    }
  }
  // pair-symmetric schedule over its step split (stops after two slower
  // splits in a row).  A larger split must win by 2%: each one adds a full
  // set of partial sums to write and reduce.
  int bs = 0;
  if (start_sym && ms0 < best) { best = ms0; bs = start_ks; }
  if (sym_eligible(op)) {
    int worse = 0;
    double sbest = 1e300;
    for (int ks = 1; ks <= 32 && worse < 2 && spent < budget_ms; ks *= 2) {
      // larger splits stop at the partial-sum memory bound
      if (op_configure(op, 0, 0, ks, true) != EBH_OK) break;
      double ms;
      EBH_TRY(time_apply(op, x.p, y.p, false, reps, &ms));
      spent += ms * reps;
      tried++;
      if (ms < 0.98 * sbest) {
        sbest = ms;
        worse = 0;
        if (ms < best) { best = ms; bs = ks; }
      } else {
        worse++;
      }
    }
  }
  if (bs) EBH_TRY(op_configure(op, 0, 0, bs, true));
  else EBH_TRY(op_configure(op, bb, bt, chunks_for_grid(op, bb, bg)));

  // multi-RHS kernel: its workgroup count, with the single-RHS shape fixed
  int bm = 0;
  double mbest = 0.0, mms0 = 0.0;
  if (op->rb) {
    bm = op->mgrid;
    EBH_TRY(time_apply(op, x.p, y.p, true, reps, &mbest));
    mms0 = mbest;
    spent += mbest * reps;
    tried++;
    for (int g : grids) {
      // (the pair-symmetric multi-RHS schedule has no workgroup count to tune)
      if (spent >= budget_ms || op->mts) break;
      if (g == bm) continue;
      op->mgrid = g;
      EBH_TRY(op_configure(op, bb, bt, bs ? bs : op->chunks, bs > 0));
      double ms;
      EBH_TRY(time_apply(op, x.p, y.p, true, reps, &ms));
      spent += ms * reps;
      tried++;
      if (ms < mbest) { mbest = ms; bm = g; }
    }
    op->mgrid = bm;
    EBH_TRY(op_configure(op, bb, bt, bs ? bs : op->chunks, bs > 0));
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[key_of(op)] = Entry{bb, bt, bg, bm, best, bs};
  }
  // the shape in effect (the symmetric schedule fixes block and tile)
  out->block = op->block;
  out->tile = op->tile;
  out->chunks = op->chunks;
  out->grid = bs ? grid_of(op) : bg;
  out->mgrid = bm;
  out->ms = best;
  out->ms_default = ms0;
  out->mms = mbest;
  out->mms_default = mms0;
  out->tried = tried;
  out->schedule = bs ? EBH_SCHED_SYMMETRIC : EBH_SCHED_CHUNKED;
  out->split = bs;
  return EBH_OK;
}

int ebh_tune_save(const char *path)
{
  if (!path) { set_error("null path"); return EBH_ERR_ARG; }
  std::ofstream f(path);
  if (!f) { set_error(std::string("cannot write ") + path); return EBH_ERR_ARG; }
  std::lock_guard<std::mutex> lk(g_mu);
  for (const auto &kv : g_table)
    f << kv.first << ' ' << kv.second.block << ' ' << kv.second.tile << ' '
      << kv.second.grid << ' ' << kv.second.mgrid << ' ' << kv.second.ms << ' '
      << kv.second.sched << '\n';
  return f ? EBH_OK : EBH_ERR_ARG;
}

int ebh_tune_load(const char *path)
{
  if (!path) { set_error("null path"); return EBH_ERR_ARG; }
  std::ifstream f(path);
  if (!f) { set_error(std::string("cannot read ") + path); return EBH_ERR_ARG; }
  std::map<std::string, Entry> add;
  std::string line;
  int n = 0;
  while (std::getline(f, line)) {
    n++;
    if (line.empty()) continue;
    std::istringstream is(line);
    std::string key;
    Entry e;
    e.sched = 0;
    const bool ok = static_cast<bool>(is >> key >> e.block >> e.tile >> e.grid >> e.mgrid >> e.ms);
    if (ok && !(is >> e.sched)) e.sched = 0;  // six-column lines: chunked
    if (!ok || (e.block != 64 && e.block != 128 && e.block != 256) || e.tile < 1 ||
        e.tile > 64 || e.grid < 1 || e.mgrid < 0 || e.sched < 0 || e.sched > 64 ||
        (e.sched & (e.sched - 1))) {
      set_error(std::string(path) + ": malformed line " + std::to_string(n));
      return EBH_ERR_ARG;
    }
    add[key] = e;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (const auto &kv : add) g_table[kv.first] = kv.second;
  return EBH_OK;
}

int ebh_tune_clear(void)
{
  std::lock_guard<std::mutex> lk(g_mu);
  g_table.clear();
  return EBH_OK;
}

}  // extern "C"
