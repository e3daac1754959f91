// Host-side code under the sanitizers: the topology and near-field pattern
// builders (easybem_amd/csrc/host_geometry.cpp, threaded) and the CPU
// oracle (oracle/oracle.c), on valid and invalid meshes.  Built and run by
// tests/test_sanitize.py with -fsanitize=address,undefined and with
// -fsanitize=thread; any sanitizer report fails the run, and the checks
// below (pair counts by closed form, rejections) end it with status 1.
#include <cmath>
#include <cstdio>
#include <string>
#include <algorithm>
#include <utility>
#include <vector>

#include "host_geometry.h"
#ifdef WITH_ORACLE
#include "oracle.h"
#endif

namespace {

int failures = 0;
void expect(bool ok, const char *what)
{
  if (!ok) {
    std::fprintf(stderr, "check failed: %s\n", what);
    failures++;
  }
}

// torus R = 1, r = 0.4 on an nu x nv grid, two triangles per quad, 1-based
void torus(int nu, int nv, std::vector<double> &P, std::vector<int> &T)
{
  P.clear();
  T.clear();
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < nv; j++) {
      const double a = 2 * M_PI * i / nu, b = 2 * M_PI * j / nv;
      P.push_back((1 + 0.4 * std::cos(b)) * std::cos(a));
      P.push_back((1 + 0.4 * std::cos(b)) * std::sin(a));
      P.push_back(0.4 * std::sin(b));
    }
  auto id = [&](int i, int j) { return ((i + nu) % nu) * nv + (j + nv) % nv + 1; };
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < nv; j++) {
      const int a = id(i, j), b = id(i + 1, j), c = id(i + 1, j + 1), d = id(i, j + 1);
      T.insert(T.end(), {a, b, c, a, c, d});
    }
}

void check_mesh(const std::vector<double> &P, const std::vector<int> &T, bool torus_grid)
{
  const int np = (int)P.size() / 3, ne = (int)T.size() / 3;
  ebh::Topology top;
  const std::string err = ebh::build_topology(np, P.data(), ne, T.data(), top);
  expect(err.empty(), "valid mesh accepted");
  if (!err.empty()) return;
  expect(top.n_class[ebh::kCoincident] == ne, "coincident pairs = ne");
  expect(top.n_class[ebh::kEdge] == 3 * (int64_t)ne, "edge pairs = 3 ne");
  if (torus_grid) expect(top.n_class[ebh::kVertex] == 18 * (int64_t)np, "vertex pairs = 18 np");
  const int64_t total = top.n_class[0] + top.n_class[1] + top.n_class[2] + top.n_class[3];
  expect(total == (int64_t)ne * ne, "pair classes cover all pairs");
  // fused-pass row blocks: aligned padded bases, every (t, k) at its own
  // padded position inside its node's block, in incidence order
  const size_t nb = top.fin_blk.size() - 1;
  expect(top.fin_blk.front() == 0 && top.fin_blk.back() == np, "row blocks cover the rows");
  expect(top.fin_pos.size() == nb + 1, "one padded base per block");
  std::vector<int> owner((size_t)top.fin_pos.back(), -1);
  for (size_t b = 0; b < nb; b++) {
    const int r0 = top.fin_blk[b], r1 = top.fin_blk[b + 1];
    const int q0 = top.node_tri_ptr[r0], n = top.node_tri_ptr[r1] - q0;
    expect(r1 > r0 && r1 - r0 <= ebh::kFinRows, "rows per block");
    expect(n <= ebh::kFinPos || r1 - r0 == 1, "incidences per block");
    expect(top.fin_pos[b] % ebh::kFinPos == 0 && top.fin_pos[b + 1] - top.fin_pos[b] >= n,
           "aligned padded block");
    for (int p = q0; p < q0 + n; p++) {
      const int t = top.node_tri[p], k = top.node_tri_k[p];
      const int q = top.tri_pos[(size_t)k * ne + t];
      expect(q == top.fin_pos[b] + (p - q0), "padded position of (t, k)");
      if (q >= 0 && q < (int)owner.size()) owner[q] = p;
    }
  }
  int placed = 0;
  for (int v : owner) placed += v >= 0;
  expect(placed == 3 * ne, "every (t, k) placed once");
  for (int tb = 0; tb < 2; tb++)
    for (int trb = 0; trb < 2; trb++) {
      ebh::NearCSR c;
      ebh::build_near_csr(top, tb, trb, c);
      expect(c.row_ptr.size() == (size_t)(tb ? np : ne) + 1, "CSR rows");
      expect(c.contrib_ptr.size() == c.col.size() + 1, "CSR contributions");
    }
  std::vector<double> curl;
  ebh::surface_curls(top, curl);
  expect(curl.size() == 9 * (size_t)ne, "curl size");
#ifdef WITH_ORACLE
  std::vector<double> V((size_t)ne * ne), K((size_t)ne * np), M((size_t)ne * np);
  expect(orc_laplace_dense(np, P.data(), ne, T.data(), 3, 4, V.data(), K.data(), M.data()) ==
             0, "oracle dense Laplace");
  std::vector<double> Ve(9 * (size_t)ne * ne);
  expect(orc_kelvin_dense(np, P.data(), ne, T.data(), 0.3, 3, 3, Ve.data()) == 0,
         "oracle dense Kelvin");
  // x holds 3 components of the larger DOF space (triangles for P0)
  std::vector<double> x(3 * (size_t)std::max(np, ne), 0.5), y(3 * 8);
  const int rows[8] = {0, 1, 2, 3, 5, 8, 13, 21};
  for (int op = 0; op < 3; op++)
    for (int tb = 0; tb < 2; tb++)
      for (int trb = 0; trb < 2; trb++) {
        if ((op == 1 && trb != 1) || (op == 2 && tb != trb)) continue;
        expect(orc_apply(op, tb, trb, np, P.data(), ne, T.data(), 3, 9, 0.3, x.data(), 8,
                         rows, y.data(), 2) == 0, "oracle apply");
      }
  std::vector<double> oc(9 * (size_t)ne), tt(ne, 1.0), uu(np, 1.0), pts = {0.0, 0.0, 3.0},
      res(1);
  expect(orc_surface_curls(np, P.data(), ne, T.data(), oc.data()) == 0, "oracle curls");
  expect(orc_evaluate(np, P.data(), ne, T.data(), 1, pts.data(), tt.data(), uu.data(), 19,
                      res.data(), 2) == 0, "oracle evaluate");
#endif
}

}  // namespace

// Pass plan of the pair-symmetric multi-RHS schedule (ebh::mt_plan): every
// tile in exactly one pass, passes on the seg-row grid, each pass within the
// limit unless it is one block, and every output's item segments the same,
// in the same order, as in the one-pass plan (so y does not depend on the
// limit).
void check_mt_plan(int ne, double tile_bytes, double item_bytes, double limit, int seg)
{
  ebh::MtPlan p, one;
  ebh::mt_plan(ne, tile_bytes, item_bytes, limit, seg, p);
  ebh::mt_plan(ne, tile_bytes, item_bytes, 1e300, seg, one);
  const int ng = (ne + 63) / 64, np_ = (int)p.rows.size() - 1;
  expect(p.ng == ng && one.rows.size() == 2, "one pass without a limit");
  expect(p.rows.front() == 0 && p.rows.back() == ng, "passes cover the tile rows");
  expect(p.rowoff.back() == ng * (ng + 1) / 2, "tile count");
  // per output group, the (M0, M1) segments in the order they are summed
  std::vector<std::vector<std::pair<int, int>>> segs(ng), segs1(ng);
  auto collect = [&](const ebh::MtPlan &q, std::vector<std::vector<std::pair<int, int>>> &out,
                     bool check) {
    for (size_t pp = 0; pp + 1 < q.rows.size(); pp++) {
      const int I0 = q.rows[pp], I1 = q.rows[pp + 1];
      if (check) {
        expect(I1 > I0, "pass not empty");
        if (pp + 2 < q.rows.size()) expect(I1 % seg == 0, "pass ends on the block grid");
        expect(q.t0[pp] == q.rowoff[I0] && q.t0[pp + 1] == q.rowoff[I1], "pass tiles");
        size_t nt = (size_t)(q.rowoff[I1] - q.rowoff[I0]);
        size_t ni = (size_t)(q.ioff[pp + 1] - q.ioff[pp]);
        const double b = nt * tile_bytes + ni * item_bytes;
        expect(b <= limit || I1 - I0 <= seg, "pass within the limit or one block");
        expect(nt <= q.max_tiles && ni <= q.max_items, "largest pass counts");
        expect(q.koff[pp + 1] - q.koff[pp] == ng - I0 + 1, "item pointers per output");
        // the launch order is a permutation of the pass's items, longest first
        std::vector<char> seen(ni, 0);
        int prev = 1 << 30;
        for (size_t i = 0; i < ni; i++) {
          const int o = q.order[q.ioff[pp] + i];
          const bool ok = o >= 0 && (size_t)o < ni && !seen[o];
          expect(ok, "launch order is a permutation");
          if (!ok) break;
          seen[o] = 1;
          const int *it = &q.item[4 * (size_t)(q.ioff[pp] + o)];
          expect(it[3] - it[2] <= prev, "longest items first");
          prev = it[3] - it[2];
        }
      }
      const int *kp = &q.kptr[q.koff[pp]];
      for (int K = I0; K < ng; K++) {
        int lastM = -1;
        for (int i = kp[K - I0]; i < kp[K - I0 + 1]; i++) {
          const int *it = &q.item[4 * (size_t)(q.ioff[pp] + i)];
          if (check) {
            expect(it[0] == K, "item output group");
            expect(it[3] > it[2] && it[3] - it[2] <= seg, "item length");
            expect(it[2] > lastM, "items in increasing M");
            if (it[1] == 1) {
              expect(it[2] >= I0 && it[3] <= std::min(I1, K), "column item inside the pass");
              expect(it[2] % seg == 0, "column item on the block grid");
            } else {
              expect(K < I1 && it[2] >= K && it[3] <= ng, "row item of a pass row");
            }
          }
          lastM = it[3] - 1;
          out[K].push_back({it[2], it[3]});
        }
      }
    }
  };
  collect(p, segs, true);
  collect(one, segs1, false);
  for (int K = 0; K < ng; K++) {
    int next = 0;
    for (auto &sg : segs[K]) {
      expect(sg.first == next, "output's items cover every source group once, in order");
      next = sg.second;
    }
    expect(next == ng, "output's items reach the last group");
    expect(segs[K] == segs1[K], "the same segments as one pass");
  }
}

int main()
{
  std::vector<double> P;
  std::vector<int> T;
  // multi-RHS pass plans: icosphere L3 and L6 and a ragged mesh, limits
  // from one block per pass to one pass; FP32 and FP64 tiles, 64 RHS
  check_mt_plan(1280, 16384.0, 32768.0, 1.0, 32);
  check_mt_plan(81920, 16384.0, 32768.0, 4.0 * (1 << 30), 32);
  check_mt_plan(81920, 32768.0, 32768.0, 4.0 * (1 << 30), 32);
  check_mt_plan(81920, 16384.0, 32768.0, 256.0 * (1 << 20), 32);
  check_mt_plan(20480, 32768.0, 4096.0, 1 << 20, 32);
  check_mt_plan(20000, 16384.0, 32768.0, 64.0 * (1 << 20), 32);
  check_mt_plan(520, 16384.0, 2048.0, 1e300, 32);
  // valid meshes: torus grids (the threaded builders split their work
  // across host threads from a few hundred triangles up)
  torus(8, 4, P, T);
  check_mesh(P, T, true);
  torus(48, 24, P, T);
  check_mesh(P, T, true);
  // invalid meshes are rejected with a message
  ebh::Topology top;
  {
    std::vector<int> open(T.begin(), T.end() - 3);
    expect(!ebh::build_topology((int)P.size() / 3, P.data(), (int)open.size() / 3,
                                open.data(), top).empty(), "open mesh rejected");
  }
  {
    std::vector<int> bad = T;
    bad[4] = (int)P.size() / 3 + 7;
    expect(!ebh::build_topology((int)P.size() / 3, P.data(), (int)bad.size() / 3,
                                bad.data(), top).empty(), "index out of range rejected");
  }
  {
    std::vector<int> flip = T;
    std::swap(flip[1], flip[2]);
    expect(!ebh::build_topology((int)P.size() / 3, P.data(), (int)flip.size() / 3,
                                flip.data(), top).empty(), "inconsistent orientation rejected");
  }
  {
    std::vector<double> Q = P;
    const int a = T[0] - 1, b = T[1] - 1;
    for (int k = 0; k < 3; k++) Q[3 * b + k] = Q[3 * a + k];
    expect(!ebh::build_topology((int)Q.size() / 3, Q.data(), (int)T.size() / 3, T.data(),
                                top).empty(), "degenerate triangle rejected");
  }
  // quadrature and Sauter-Schwab tables at every index
  for (int i = 0; i < 20; i++) {
    std::vector<double> x, y, w;
    expect(ebh::tri_rule(i, x, y, w), "triangle rule");
    double c[ebh::kMaxTriPoints];
    std::vector<double> x2 = x, y2 = y, w2 = w;
    (void)ebh::ray_order((int)w.size(), x2.data(), y2.data(), w2.data(), c);
#ifdef WITH_ORACLE
    double p[2 * 80], ww[80];
    expect(orc_tri_rule(i, p, ww) == (int)w.size(), "oracle triangle rule");
#endif
  }
  for (int i = 0; i < 19; i++) {
    std::vector<double> p, w;
    expect(ebh::line_rule(i, p, w), "line rule");
  }
  for (int cls = 1; cls <= 3; cls++)
    for (int kd = 1; kd <= 2; kd++)
      for (int nq = 1; nq <= 6; nq++) {
        std::vector<double> W;
        ebh::ss_table(cls, kd, nq, W);
        expect(!W.empty(), "Sauter-Schwab table");
      }
  std::printf("host checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
